// Deterministic mode, wave-parallel: one wavefront (64 lanes) runs one cell.
//
// The cell's trajectory is the reference's, draw for draw (simulate_one_cell,
// src/libmodle/cpu/simulation.cpp:1056-1152, and everything det_sim.h restates serially). What
// makes it parallel:
//
// * Engine words. The cell's Xoshiro256++ stream is produced ahead into an LDS ring: each lane
//   generates a chunk of kWChunk consecutive words and then advances its own engine state by one
//   batch (kWBatch words) with the precomputed GF(2) jump table kDetJump. Word k of the stream
//   sits in ring slot k mod kWRing; `cursor` is the next word the reference would consume.
// * Draws. A step that makes one draw per element (moves, barrier states, releases, bindings,
//   LEF-BAR and LEF-LEF trials) gives every lane one element and evaluates its sampler at the
//   offset the draw would have if every earlier draw took its common word count F. The first
//   lane whose draw took another count (a ziggurat rejection, a bucket rejection) is exact; the
//   lanes after it are evaluated again from its end. Contact attempts work the same way with a
//   cut at the attempt that completes the contact count.
// * Sorts and scans. rank_lefs' order is unique when no two LEFs tie in position and binding
//   epoch; the kept order is checked and repaired by odd-even passes, the LEFs bound this epoch
//   are sorted by counting and merged by binary search. A tie, an overlong repair or a large
//   batch of new LEFs runs the reference's sorters (det_sort.h) on lane 0 instead, from the
//   same input order. Moves are adjusted by segmented prefix-min / prefix-max scans; a unit
//   whose adjusted move reaches a chromosome end (which stops the reference's walk) sends that
//   side through the serial walk.
// * Collisions. LEF-BAR pairs are found by binary search per active barrier (the last
//   successful barrier in the reference's loop order wins), primary LEF-LEF pairs per fwd rank,
//   secondary chains by a walk from each stalled unit with the trials resolved in rank order;
//   bypass fix-ups touch disjoint rank pairs and run in parallel.
//
// The code is written once against a wave policy W (lane(), sync(), ballot(), bcast()); the
// kernel instantiates it with the hardware wavefront (DevWave), and a host build of the same
// header can run it with 64 emulated lanes.
#pragma once
#ifndef MGPU_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "det_jump.h"
#include "det_sim.h"

namespace mgpu {

constexpr uint32_t kWL = 64;                  // lanes per cell
constexpr uint64_t kWChunk = 32;              // words one lane generates per batch
constexpr uint64_t kWBatch = kWL * kWChunk;   // 2048: the jump kDetJump advances
constexpr uint64_t kWRing = 2 * kWBatch;      // ring of engine words (32 KB)
static_assert(kWBatch == kDetJumpSteps, "det_jump.h must advance the engine by one batch");
constexpr uint64_t kWMaxNewCount = 1024;      // larger batches of new LEFs sort on lane 0
constexpr uint32_t kWMaxRepair = 64;          // odd-even passes before the serial sort
constexpr uint64_t kWMaxAttempts = 1ull << 26;  // contact attempts before giving up (error)

#ifndef MGPU_EMU
// The hardware wavefront of a 64-thread block.
struct DevWave {
  __device__ uint32_t lane() const { return threadIdx.x; }
  __device__ void sync() const { __syncthreads(); }
  __device__ uint64_t ballot(bool p) const { return __ballot(p); }
  __device__ uint64_t bcast(uint64_t v, uint32_t src) const {
    const uint32_t lo = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)v, (int)src);
    const uint32_t hi = (uint32_t)__builtin_amdgcn_readlane((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
  }
  // the value of lane src (per lane; src in [0, 64))
  __device__ uint64_t shfl(uint64_t v, uint32_t src) const {
    const uint32_t lo = (uint32_t)__shfl((int)(uint32_t)v, (int)src);
    const uint32_t hi = (uint32_t)__shfl((int)(uint32_t)(v >> 32), (int)src);
    return ((uint64_t)hi << 32) | lo;
  }
};
#endif

// Counters of the exact-but-rare paths, kept when the caller passes a counter array (the tests'
// host build does, to show every path was taken).
enum WDbg : uint32_t {
  kDbgRankSerial,    // rank order resolved by the serial sorters
  kDbgRepairPass,    // odd-even repair passes over the kept order
  kDbgAdjustSerial,  // move adjustment by the serial walk
  kDbgDrawRedo,      // draws evaluated again after a draw took another word count
  kDbgContactCut,    // contact batches cut at an attempt with another word count
  kDbgBypass,        // secondary trials that failed (bypasses)
  kDbgChainDrop,     // chain members dropped after a bypass
  kDbgCount
};
__device__ inline void wdbg(uint64_t* dbg, uint32_t lane, WDbg k, uint64_t v = 1) {
  if (dbg && lane == 0) dbg[k] += v;
}

// ---------------------------------------------------------------------------------------------
// engine words
// ---------------------------------------------------------------------------------------------
__device__ inline void det_jump_batch(uint64_t& s0, uint64_t& s1, uint64_t& s2, uint64_t& s3) {
  uint64_t y0 = 0, y1 = 0, y2 = 0, y3 = 0;
#pragma unroll
  for (int w = 0; w < 4; ++w) {
    const uint64_t x = w == 0 ? s0 : (w == 1 ? s1 : (w == 2 ? s2 : s3));
#pragma unroll 8
    for (int b = 0; b < 64; ++b) {
      const uint64_t m = 0ull - ((x >> b) & 1ull);
      const uint64_t* c = kDetJump[w * 64 + b];
      y0 ^= c[0] & m;
      y1 ^= c[1] & m;
      y2 ^= c[2] & m;
      y3 ^= c[3] & m;
    }
  }
  s0 = y0;
  s1 = y1;
  s2 = y2;
  s3 = y3;
}

// Wave-uniform ring bookkeeping plus this lane's engine state for its next chunk; zn: the normal
// ziggurat tables (x[129] then y[129]) staged in LDS.
struct WRing {
  uint64_t* ring;
  uint64_t cursor, hi;
  uint64_t g0, g1, g2, g3;
  const double* zn;
};

// A sampler's view of the ring from a given word offset.
struct WSrc {
  const uint64_t* ring;
  uint64_t pos, hi;
  const double* zn;
};
__device__ inline uint64_t det_next(WSrc& s) {
  const uint64_t p = s.pos++;
  return p < s.hi ? s.ring[p & (kWRing - 1)] : 0;
}
__device__ inline double zig_nx(const WSrc& g, int i) { return g.zn[i]; }
__device__ inline double zig_ny(const WSrc& g, int i) { return g.zn[129 + i]; }

template <class W>
__device__ inline void wr_init(const W& w, WRing& r, uint64_t seed, uint64_t* states) {
  if (w.lane() == 0) {
    DetRng g;
    det_seed(g, seed);
    for (uint64_t k = 0; k < kWBatch; ++k) {
      if (k % kWChunk == 0) {
        uint64_t* st = states + 4 * (k / kWChunk);
        st[0] = g.s0;
        st[1] = g.s1;
        st[2] = g.s2;
        st[3] = g.s3;
      }
      r.ring[k] = det_next(g);
    }
  }
  w.sync();
  const uint64_t* st = states + 4 * w.lane();
  r.g0 = st[0];
  r.g1 = st[1];
  r.g2 = st[2];
  r.g3 = st[3];
  det_jump_batch(r.g0, r.g1, r.g2, r.g3);
  r.hi = kWBatch;
  r.cursor = 0;
  w.sync();
}

// Keeps at least half a ring of words ahead of the cursor (called at wave-uniform points).
template <class W>
__device__ inline void wr_refill(const W& w, WRing& r) {
  if (r.hi - r.cursor >= kWRing / 2) return;
  w.sync();
  DetRng g;
  g.s0 = r.g0;
  g.s1 = r.g1;
  g.s2 = r.g2;
  g.s3 = r.g3;
  g.draws = 0;
  const uint64_t base = r.hi + (uint64_t)w.lane() * kWChunk;
  for (uint64_t j = 0; j < kWChunk; ++j) r.ring[(base + j) & (kWRing - 1)] = det_next(g);
  det_jump_batch(r.g0, r.g1, r.g2, r.g3);
  r.hi += kWBatch;
  w.sync();
}

__device__ inline uint64_t lanes_below(uint32_t lane) {
  return lane == 0 ? 0ull : (~0ull >> (64 - lane));
}

// One draw for every element k in [0, count) whose flag is set, in element order (element
// index e = k, or count - 1 - k with `reverse`). F: words a draw takes on its common path.
// flag(e) must not depend on what apply() writes in the same call.
template <class W, class Flag, class Eval, class Apply>
__device__ inline void w_draws(const W& w, WRing& r, uint64_t count, bool reverse, uint64_t F,
                               uint32_t& err, uint64_t* dbg, Flag flag, Eval eval, Apply apply) {
  const uint32_t lane = w.lane();
  const uint64_t lt = lanes_below(lane);
  for (uint64_t g = 0; g < count; g += kWL) {
    wr_refill(w, r);
    const uint64_t k = g + lane;
    const uint64_t e = reverse ? count - 1 - k : k;
    const bool need = k < count && flag(e);
    uint64_t pending = w.ballot(need);
    uint64_t val = 0;
    while (pending != 0) {
      const bool mine = (pending >> lane) & 1ull;
      const uint64_t off = r.cursor + (uint64_t)__builtin_popcountll(pending & lt) * F;
      uint64_t used = F;
      if (mine) {
        WSrc src{r.ring, off, r.hi, r.zn};
        val = eval(e, src);
        used = src.pos - off;
      }
      const uint64_t bad = w.ballot(mine && used != F);
      if (bad == 0) {
        if (mine) apply(e, val);
        r.cursor += (uint64_t)__builtin_popcountll(pending) * F;
        break;
      }
      const uint32_t kb = (uint32_t)__builtin_ctzll(bad);
      wdbg(dbg, lane, kDbgDrawRedo);
      if (mine && lane <= kb) apply(e, val);
      const uint64_t end = w.bcast(off + used, kb);
      if (end > r.hi) err |= 1u;
      r.cursor = end;
      pending = kb == 63 ? 0ull : (pending & (~0ull << (kb + 1)));
    }
  }
  w.sync();
}

// One draw on lane 0's behalf (a Poisson or binomial count); the result is wave-uniform.
template <class W, class Eval>
__device__ inline uint64_t w_single(const W& w, WRing& r, uint32_t& err, Eval eval) {
  wr_refill(w, r);
  uint64_t val = 0, used = 0;
  if (w.lane() == 0) {
    WSrc src{r.ring, r.cursor, r.hi, r.zn};
    val = eval(src);
    used = src.pos - r.cursor;
  }
  val = w.bcast(val, 0);
  used = w.bcast(used, 0);
  if (r.cursor + used > r.hi) err |= 1u;
  r.cursor += used;
  return val;
}

// ---------------------------------------------------------------------------------------------
// cell state
// ---------------------------------------------------------------------------------------------
struct WCell {
  uint64_t *rev, *fwd, *ep, *rm, *fm, *rc, *fc, *x0, *x1, *x2;
  uint32_t *rr, *fr, *y0, *y1, *y2, *bsc;
  uint8_t* bst;
  double *cfx, *avg;
  uint64_t* xs;  // 64 x 4 words of wave scratch
  uint64_t* dbg;  // path counters (host builds only; null on the device), see WDbg
  uint64_t n, nact, epoch, burnin_epochs, contacts, hlen;
  bool burnin_done;
  uint32_t err;
  WRing r;
  // phase timers (DetParams::phase): ticks per phase, kept alike on every lane
  bool timed;
  uint64_t tlast;
  uint64_t tacc[12];
};

#ifdef MGPU_EMU
inline uint64_t wclock() { return 0; }
#else
__device__ inline uint64_t wclock() { return wall_clock64(); }
#endif
// Closes phase k at a wave-uniform point: every lane reads the clock (no divergent branch).
__device__ inline void wmark(WCell& s, int k) {
  if (s.timed) {
    const uint64_t t = wclock();
    s.tacc[k] += t - s.tlast;
    s.tlast = t;
  }
}

// LDS the wave always keeps: the engine-word ring, wave scratch and the normal ziggurat tables.
__device__ __host__ inline uint64_t det_wave_fixed_bytes() {
  return ((kWRing * 8 + kWL * 4 * 8 + 2 * 129 * 8) + 255) & ~255ull;
}
// The cell's own arrays: 10 u64 and 5 u32 per LEF, the burn-in histories, per barrier a u32 and
// a state byte.
__device__ __host__ inline uint64_t det_wave_cell_bytes(uint64_t nlefs, uint64_t nbar,
                                                        uint64_t hist) {
  const uint64_t b = nlefs * (10 * 8 + 5 * 4) + ((nbar * 5 + 7) & ~7ull) + 2 * (hist + 1) * 8;
  return (b + 255) & ~255ull;
}
// Everything in LDS.
__device__ __host__ inline uint64_t det_wave_bytes(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  return det_wave_fixed_bytes() + det_wave_cell_bytes(nlefs, nbar, hist);
}

// Where a cell's arrays live: in LDS after the fixed part, in device memory at cell_mem, or
// either, chosen when the cell starts (cell_mem null: LDS). A fixed placement lets the compiler
// address the arrays with LDS or global instructions instead of flat ones.
enum : int { kPlaceLds = 0, kPlaceMem = 1, kPlaceAny = 2 };

// lds: the wave's LDS; cell: the cell's arrays (det_wave_cell_bytes) for kPlaceMem, or for
// kPlaceAny when not null.
template <int kPlace>
__device__ inline void det_wave_carve(WCell& s, uint8_t* lds, uint8_t* cell, uint64_t nlefs,
                                      uint64_t nbar, uint64_t hist) {
  uint64_t* f = reinterpret_cast<uint64_t*>(lds);
  s.r.ring = f;
  f += kWRing;
  s.xs = f;
  f += kWL * 4;
  s.r.zn = reinterpret_cast<const double*>(f);
  uint8_t* const in_lds = lds + det_wave_fixed_bytes();
  uint64_t* p = reinterpret_cast<uint64_t*>(kPlace == kPlaceLds   ? in_lds
                                            : kPlace == kPlaceMem ? cell
                                                                  : (cell ? cell : in_lds));
  s.rev = p;
  s.fwd = p + nlefs;
  s.ep = p + 2 * nlefs;
  s.rm = p + 3 * nlefs;
  s.fm = p + 4 * nlefs;
  s.rc = p + 5 * nlefs;
  s.fc = p + 6 * nlefs;
  s.x0 = p + 7 * nlefs;
  s.x1 = p + 8 * nlefs;
  s.x2 = p + 9 * nlefs;
  double* pd = reinterpret_cast<double*>(p + 10 * nlefs);
  s.cfx = pd;
  s.avg = pd + hist + 1;
  uint32_t* p32 = reinterpret_cast<uint32_t*>(pd + 2 * (hist + 1));
  s.rr = p32;
  s.fr = p32 + nlefs;
  s.y0 = p32 + 2 * nlefs;
  s.y1 = p32 + 3 * nlefs;
  s.y2 = p32 + 4 * nlefs;
  s.bsc = p32 + 5 * nlefs;
  s.bst = reinterpret_cast<uint8_t*>(p32 + 5 * nlefs + nbar);
}

// ---------------------------------------------------------------------------------------------
// wave helpers
// ---------------------------------------------------------------------------------------------
// Copies the elements of src[0, n) satisfying pred to dst in order; returns their count.
template <class W, class P>
__device__ inline uint64_t w_compact(const W& w, const uint32_t* src, uint64_t n, uint32_t* dst,
                                     P pred) {
  const uint32_t lane = w.lane();
  const uint64_t lt = lanes_below(lane);
  uint64_t base = 0;
  for (uint64_t g = 0; g < n; g += kWL) {
    const uint64_t k = g + lane;
    uint32_t v = 0;
    bool p = false;
    if (k < n) {
      v = src[k];
      p = pred(v);
    }
    const uint64_t m = w.ballot(p);
    if (p) dst[base + __builtin_popcountll(m & lt)] = v;
    base += __builtin_popcountll(m);
  }
  w.sync();
  return base;
}

// True on every lane when pred holds for any element in [0, n).
template <class W, class P>
__device__ inline bool w_any(const W& w, uint64_t n, P pred) {
  bool any = false;
  for (uint64_t g = 0; g < n && !any; g += kWL) {
    const uint64_t k = g + w.lane();
    any = w.ballot(k < n && pred(k)) != 0;
  }
  return any;
}

// Sum over lanes (every lane gets it), by a butterfly of lane exchanges.
template <class W>
__device__ inline uint64_t w_sum(const W& w, uint64_t*, uint64_t v) {
  for (uint32_t d = 1; d < kWL; d <<= 1) v += w.shfl(v, w.lane() ^ d);
  return v;
}

// Smallest k in [lo, hi) with pred(k), else hi (every lane gets it; pred monotone not needed).
template <class W, class P>
__device__ inline uint64_t w_first(const W& w, uint64_t lo, uint64_t hi, P pred) {
  for (uint64_t g = lo; g < hi; g += kWL) {
    const uint64_t k = g + w.lane();
    const uint64_t m = w.ballot(k < hi && pred(k));
    if (m) return g + (uint64_t)__builtin_ctzll(m);
  }
  return hi;
}

// ---------------------------------------------------------------------------------------------
// rank_lefs (simulation.cpp:402-487)
// ---------------------------------------------------------------------------------------------
struct WKey {
  const uint64_t* pos;
  const uint64_t* ep;
  bool asc;  // ties by binding epoch ascending (rev) or descending (fwd)
  __device__ bool less(uint32_t a, uint32_t b) const {
    if (pos[a] != pos[b]) return pos[a] < pos[b];
    return asc ? ep[a] < ep[b] : ep[b] < ep[a];
  }
  __device__ bool tie(uint32_t a, uint32_t b) const { return pos[a] == pos[b] && ep[a] == ep[b]; }
};

// Number of elements of arr[0, len) (sorted by k) ordered before x.
__device__ inline uint64_t w_count_less(const uint32_t* arr, uint64_t len, const WKey& k,
                                        uint32_t x) {
  uint64_t lo = 0, hi = len;
  while (lo < hi) {
    const uint64_t mid = (lo + hi) >> 1;
    if (k.less(arr[mid], x)) lo = mid + 1;
    else hi = mid;
  }
  return lo;
}

// Orders rk[0, n) as rank_lefs does. kl, nl, ns: scratch of n entries; tmp: n entries for the
// serial sorters, or the slot marks of the merge.
template <class W>
__device__ inline void w_rank_side(const W& w, WCell& s, uint32_t* rk, const uint64_t* pos,
                                   bool asc, uint64_t n, bool partially_sorted, uint32_t* kl,
                                   uint32_t* nl, uint32_t* ns, uint32_t* tmp) {
  const uint32_t lane = w.lane();
  const WKey key{pos, s.ep, asc};
  const uint64_t epoch = s.epoch;
  // kept and new (bound this epoch) LEFs in rank order, in one pass; rk itself stays as it was
  // until the merge, so the serial sorters below start from it
  uint64_t nk = 0, nn = 0;
  {
    const uint64_t lt = lanes_below(lane);
    for (uint64_t g = 0; g < n; g += kWL) {
      const uint64_t k = g + lane;
      const bool has = k < n;
      const uint32_t v = has ? rk[k] : 0u;
      const bool isnew = has && s.ep[v] == epoch;
      const uint64_t mn = w.ballot(isnew);
      const uint64_t mk = w.ballot(has && !isnew);
      if (isnew) nl[nn + (uint64_t)__builtin_popcountll(mn & lt)] = v;
      else if (has) kl[nk + (uint64_t)__builtin_popcountll(mk & lt)] = v;
      nn += (uint64_t)__builtin_popcountll(mn);
      nk += (uint64_t)__builtin_popcountll(mk);
    }
    w.sync();
  }
  bool serial = nn > kWMaxNewCount;
  if (!serial) {
    // new LEFs: rank by counting. They share the binding epoch, so the key is the position
    // (below 2^32: chromosome ends are) and equal positions tie. Each lane holds the positions
    // of one chunk of 64 new LEFs in a register and the others read them with readlane.
    bool tie = false;
    for (uint64_t q = 0; q < nn; q += kWL) {
      const uint64_t b = q + lane;
      const bool mine = b < nn;
      const uint32_t x = mine ? nl[b] : 0u;
      const uint32_t px = mine ? (uint32_t)pos[x] : 0u;
      uint64_t c = 0;
      bool tq = false;  // a tie of this lane's LEF in this chunk
      for (uint64_t r = 0; r < nn; r += kWL) {
        const uint64_t j = r + lane;
        const uint32_t py = j < nn ? (uint32_t)pos[nl[j]] : 0u;
        const uint32_t m = (uint32_t)(nn - r < kWL ? nn - r : kWL);
        for (uint32_t t = 0; t < m; ++t) {
          const uint32_t y = (uint32_t)w.bcast(py, t);
          c += y < px;
          tq |= y == px && r + t != b;
        }
      }
      // ties found in earlier chunks stay: a lane without a LEF in this chunk adds nothing
      if (mine) {
        ns[c] = x;
        tie |= tq;
      }
    }
    serial = w.ballot(tie) != 0;
  }
  w.sync();
  if (!serial) {
    // kept LEFs: usually still in order; odd-even passes repair a few inversions
    uint32_t passes = 0;
    for (;;) {
      const bool bad = w_any(w, nk > 0 ? nk - 1 : 0, [&](uint64_t a) {
        return !key.less(kl[a], kl[a + 1]);
      });
      if (!bad) break;
      wdbg(s.dbg, lane, kDbgRepairPass);
      if (++passes > kWMaxRepair) {
        serial = true;
        break;
      }
      bool tie = false;
      for (uint32_t par = 0; par < 2; ++par) {
        for (uint64_t a = 2 * lane + par; a + 1 < nk; a += 2 * kWL) {
          const uint32_t u = kl[a], v = kl[a + 1];
          tie |= key.tie(u, v);
          if (key.less(v, u)) {
            kl[a] = v;
            kl[a + 1] = u;
          }
        }
        w.sync();
      }
      if (w.ballot(tie) != 0) {
        serial = true;
        break;
      }
    }
  }
  if (!serial) {
    // merge: each new LEF's slot is its rank plus the kept LEFs before it (new and kept LEFs
    // never tie: their binding epochs differ); the kept LEFs fill the other slots in order
    uint32_t* mark = tmp;
    for (uint64_t k = lane; k < n; k += kWL) mark[k] = 0;
    w.sync();
    for (uint64_t b = lane; b < nn; b += kWL) {
      const uint32_t x = ns[b];
      const uint64_t slot = b + w_count_less(kl, nk, key, x);
      rk[slot] = x;
      mark[slot] = 1;
    }
    w.sync();
    const uint64_t lt = lanes_below(lane);
    uint64_t before = 0;  // new LEFs' slots below this chunk
    for (uint64_t g = 0; g < n; g += kWL) {
      const uint64_t k = g + lane;
      const bool isnew = k < n && mark[k] != 0;
      const uint64_t m = w.ballot(isnew);
      if (k < n && !isnew) rk[k] = kl[k - before - (uint64_t)__builtin_popcountll(m & lt)];
      before += (uint64_t)__builtin_popcountll(m);
    }
    w.sync();
  }
  if (serial) {
    wdbg(s.dbg, lane, kDbgRankSerial);
    if (lane == 0) {
      if (partially_sorted) split_sort(rk, (uint32_t)n, tmp, KeyLess{pos});
      else pdq_sort(rk, (uint32_t)n, KeyLess{pos});
      fix_ties(rk, n, pos, s.ep, asc);
    }
    w.sync();
  }
}

template <class W>
__device__ inline void w_rank(const W& w, WCell& s, uint64_t n, bool partially_sorted) {
  uint32_t* a = reinterpret_cast<uint32_t*>(s.x0);
  uint32_t* b = reinterpret_cast<uint32_t*>(s.x1);
  uint32_t* tmp = reinterpret_cast<uint32_t*>(s.rm);
  w_rank_side(w, s, s.rr, s.rev, true, n, partially_sorted, s.y0, s.y1, a, tmp);
  w_rank_side(w, s, s.fr, s.fwd, false, n, partially_sorted, s.y0, s.y1, b, tmp);
}

// ---------------------------------------------------------------------------------------------
// moves (simulation.cpp:265-399)
// ---------------------------------------------------------------------------------------------
// Serial adjustment of one side, as det_adjust_clamp's loops.
__device__ inline void w_adjust_serial(const DetChrom& ch, WCell& s, uint64_t n, bool rev) {
  if (n == 0) return;
  if (rev) {
    for (uint64_t i = n - 1; i > 0; --i) {
      const uint32_t i1 = s.rr[i - 1], i2 = s.rr[i];
      if (s.ep[i1] == kUnbound || s.ep[i2] == kUnbound) continue;
      if (s.rev[i1] <= ch.start + s.rm[i1] || s.rev[i2] <= ch.start + s.rm[i2]) continue;
      const uint64_t p1 = s.rev[i1] - s.rm[i1];
      const uint64_t p2 = s.rev[i2] - s.rm[i2];
      if (p2 <= p1) s.rm[i1] += (p1 - p2) + 1;
    }
  } else {
    for (uint64_t i = 1; i < n; ++i) {
      const uint32_t i1 = s.fr[i - 1], i2 = s.fr[i];
      if (s.ep[i1] == kUnbound || s.ep[i2] == kUnbound) continue;
      if (s.fwd[i1] + s.fm[i1] > ch.end - 1 || s.fwd[i2] + s.fm[i2] > ch.end - 1) continue;
      const uint64_t p1 = s.fwd[i1] + s.fm[i1];
      const uint64_t p2 = s.fwd[i2] + s.fm[i2];
      if (p1 >= p2) s.fm[i2] += (p1 - p2) + 1;
    }
  }
}

// Adjusts one side's moves so that no unit overtakes its neighbour (the walks of
// simulation.cpp:342-399) with a segmented scan over ranks. rev: target t = pos - move, walked
// from the last rank down, t'[k-1] = min(t[k-1], t'[k] - 1); fwd: t = pos + move, walked up,
// t'[k] = max(t[k], t'[k-1] + 1). With v = t' - k both are a running min / max that restarts
// at the pairs the walk skips: an unbound LEF, or the unit on the not-yet-adjusted side of the
// pair crossing its end. The walk also skips a pair whose already-adjusted unit crosses the
// end; that case (a unit pushed onto a chromosome end) is rare and returns false, and the
// caller then runs the serial walk on the unchanged moves. Each lane scans a block of
// consecutive ranks; blocks pass (value, depends-on-carry) summaries through wave scratch.
template <class W>
__device__ inline bool w_adjust_side(const W& w, const DetChrom& ch, WCell& s, uint64_t n,
                                     bool rev) {
  const uint32_t lane = w.lane();
  int64_t* t0 = reinterpret_cast<int64_t*>(s.x0);  // targets before the walk
  int64_t* t1 = reinterpret_cast<int64_t*>(s.x1);  // after it
  const uint32_t* rk = rev ? s.rr : s.fr;
  const int64_t start = (int64_t)ch.start, endm1 = (int64_t)ch.end - 1;
  // per rank: unbound flags, then whether the walk statically skips pair (k - 1, k) (an
  // unbound LEF, or the not-yet-adjusted side's unit crossing its end), so the serial block
  // loops below read one byte per rank instead of going through the rank order
  uint8_t* ub = reinterpret_cast<uint8_t*>(s.x2);
  uint8_t* br = ub + n;
  for (uint64_t k = lane; k < n; k += kWL) {
    const uint32_t i = rk[k];
    ub[k] = s.ep[i] == kUnbound ? 1 : 0;
    t0[k] = rev ? (int64_t)s.rev[i] - (int64_t)s.rm[i] : (int64_t)s.fwd[i] + (int64_t)s.fm[i];
  }
  w.sync();
  for (uint64_t k = lane; k < n; k += kWL)
    br[k] = k > 0 && (ub[k - 1] != 0 || ub[k] != 0 || (rev ? t0[k - 1] <= start : t0[k] > endm1))
                ? 1
                : 0;
  w.sync();
  auto brk = [&](uint64_t k) -> bool { return br[k] != 0; };  // pair (k - 1, k), k >= 1
  auto comb = [&](int64_t a, int64_t b) -> int64_t {
    return rev ? (a < b ? a : b) : (a > b ? a : b);
  };
  const uint64_t B = (n + kWL - 1) / kWL;
  const uint64_t lo = (uint64_t)lane * B < n ? (uint64_t)lane * B : n;
  const uint64_t hi = lo + B < n ? lo + B : n;
  // summary: the block's outgoing value is dep ? comb(val, carry in) : val
  bool dep = true;
  int64_t val = 0;
  if (lo < hi) {
    if (rev) {
      dep = hi < n && !brk(hi);
      val = t0[hi - 1] - (int64_t)(hi - 1);
      for (uint64_t k = hi - 1; k-- > lo;) {
        const int64_t a = t0[k] - (int64_t)k;
        if (brk(k + 1)) {
          val = a;
          dep = false;
        } else {
          val = comb(a, val);
        }
      }
    } else {
      dep = lo > 0 && !brk(lo);
      val = t0[lo] - (int64_t)lo;
      for (uint64_t k = lo + 1; k < hi; ++k) {
        const int64_t b = t0[k] - (int64_t)k;
        if (brk(k)) {
          val = b;
          dep = false;
        } else {
          val = comb(b, val);
        }
      }
    }
  }
  // carry into each block: a scan, in walk order, of the blocks' summaries composed as
  // (dep_a, v_a) o (dep_b, v_b) = (dep_a && dep_b, dep_a ? comb(v_a, v_b) : v_a); an empty block
  // is the identity (dep, the identity of comb)
  const int64_t ident = rev ? INT64_MAX : INT64_MIN;
  if (lo >= hi) {
    dep = true;
    val = ident;
  }
  for (uint32_t d = 1; d < kWL; d <<= 1) {
    // the partner lane precedes this one in walk order: higher lanes for rev, lower for fwd
    const bool has = rev ? lane + d < kWL : lane >= d;
    const uint32_t src = has ? (rev ? lane + d : lane - d) : lane;
    const int64_t ov = (int64_t)w.shfl((uint64_t)val, src);
    const bool od = w.shfl(dep ? 1 : 0, src) != 0;
    if (has) {
      if (dep) val = comb(val, ov);
      dep = dep && od;
    }
  }
  // exclusive: the composed summary of the blocks before this lane's
  const bool first = rev ? lane + 1 >= kWL : lane == 0;
  const uint32_t prev = first ? lane : (rev ? lane + 1 : lane - 1);
  const int64_t pv = (int64_t)w.shfl((uint64_t)val, prev);
  const int64_t carry = first ? ident : pv;
  if (lo < hi) {
    if (rev) {
      int64_t v = carry;
      for (uint64_t k = hi; k-- > lo;) {
        const int64_t a = t0[k] - (int64_t)k;
        v = (k + 1 >= n || brk(k + 1)) ? a : comb(a, v);
        t1[k] = v + (int64_t)k;
      }
    } else {
      int64_t v = carry;
      for (uint64_t k = lo; k < hi; ++k) {
        const int64_t b = t0[k] - (int64_t)k;
        v = (k == 0 || brk(k)) ? b : comb(b, v);
        t1[k] = v + (int64_t)k;
      }
    }
  }
  w.sync();
  // pairs the walk would skip because the adjusted unit crosses the end
  const bool crossing = w_any(w, n, [&](uint64_t k) {
    if (k == 0 || brk(k)) return false;
    return rev ? t1[k] <= start : t1[k - 1] > endm1;
  });
  if (crossing) return false;
  for (uint64_t k = lane; k < n; k += kWL) {
    const uint32_t i = rk[k];
    if (rev) s.rm[i] = (uint64_t)((int64_t)s.rev[i] - t1[k]);
    else s.fm[i] = (uint64_t)(t1[k] - (int64_t)s.fwd[i]);
  }
  w.sync();
  return true;
}

template <class W>
__device__ inline void w_adjust_clamp(const W& w, const DetChrom& ch, WCell& s, uint64_t n) {
  for (int side = 0; side < 2; ++side) {
    const bool rev = side == 0;
    if (!w_adjust_side(w, ch, s, n, rev)) {
      wdbg(s.dbg, w.lane(), kDbgAdjustSerial);
      if (w.lane() == 0) w_adjust_serial(ch, s, n, rev);
      w.sync();
    }
  }
  for (uint64_t i = w.lane(); i < n; i += kWL) {
    if (s.ep[i] == kUnbound) continue;
    s.rm[i] = umin64(s.rm[i], s.rev[i] - ch.start);
    s.fm[i] = umin64(s.fm[i], ch.end - s.fwd[i] - 1);
  }
  w.sync();
}

template <class W>
__device__ inline void w_generate_moves(const W& w, const DetParams& P, const DetChrom& ch,
                                        WCell& s, uint64_t n) {
  for (int side = 0; side < 2; ++side) {
    const bool rev = side == 0;
    uint64_t* mv = rev ? s.rm : s.fm;
    const double speed = (double)(rev ? (s.burnin_done ? P.rev_speed : P.rev_speed_burnin)
                                      : (s.burnin_done ? P.fwd_speed : P.fwd_speed_burnin));
    const double sd = rev ? P.rev_std : P.fwd_std;
    for (uint64_t i = w.lane(); i < n; i += kWL)
      mv[i] = s.ep[i] == kUnbound ? 0 : (sd == 0.0 ? (uint64_t)round(speed) : 0);
    w.sync();
    if (sd != 0.0) {
      w_draws(
          w, s.r, n, false, 1, s.err, s.dbg, [&](uint64_t i) { return s.ep[i] != kUnbound; },
          [&](uint64_t, WSrc& g) -> uint64_t {
            const double v = det_normal(g, speed, sd);
            return (uint64_t)round(0.0 < v ? v : 0.0);
          },
          [&](uint64_t i, uint64_t v) { mv[i] = v; });
    }
  }
  w_adjust_clamp(w, ch, s, n);
}

// ---------------------------------------------------------------------------------------------
// collisions (simulation_detect_collisions.cpp, simulation_correct_moves.cpp)
// ---------------------------------------------------------------------------------------------
template <class W>
__device__ inline uint64_t w_max(const W& w, uint64_t*, uint64_t v) {
  for (uint32_t d = 1; d < kWL; d <<= 1) {
    const uint64_t o = w.shfl(v, w.lane() ^ d);
    v = o > v ? o : v;
  }
  return v;
}

// Chromosome-end collisions (det_boundaries' two walks): the rev walk stops at the first rank
// past the units sitting on the start that is beyond the first fwd unit or moves onto the
// start; the fwd walk, from the top, at the first bound rank below the units on the end that is
// before the last rev unit or moves onto the end.
template <class W>
__device__ inline void w_boundaries(const W& w, const DetChrom& ch, WCell& s, uint64_t n,
                                    uint64_t& n5, uint64_t& n3) {
  const uint32_t lane = w.lane();
  const uint64_t* posR = s.x0;
  const uint64_t* posF = s.x1;
  const uint64_t first_fwd = posF[0];
  uint64_t lb = 0;
  for (uint64_t j = lane; j < n; j += kWL)
    if (s.ep[s.rr[j]] != kUnbound) lb = j + 1;
  const uint64_t k = w_max(w, s.xs, lb);
  const uint64_t last_rev = posR[k > 0 ? k - 1 : 0];
  const uint64_t sr = w_first(w, 0, n, [&](uint64_t i) {
    const uint64_t pos = posR[i];
    return pos != ch.start && (pos > first_fwd || pos - s.rm[s.rr[i]] == ch.start);
  });
  const bool rend = sr < n && !(posR[sr] > first_fwd);
  // fwd walk from n - 1 down to 1: stop index (largest qualifying rank), or 0 when none
  uint64_t tf = 0;
  for (uint64_t g = 0; g + 1 < n; g += kWL) {
    const uint64_t i = n - 1 - g - lane;  // ranks n-1, n-2, ... in lane order
    bool p = false;
    if (g + lane + 1 < n) {
      const uint32_t idx = s.fr[i];
      if (s.ep[idx] != kUnbound) {
        const uint64_t pos = posF[i];
        p = pos != ch.end - 1 && (pos < last_rev || pos + s.fm[idx] == ch.end - 1);
      }
    }
    const uint64_t m = w.ballot(p);
    if (m) {
      tf = n - 1 - g - (uint64_t)__builtin_ctzll(m);
      break;
    }
  }
  const bool fend = tf >= 1 && !(posF[tf] < last_rev);
  uint64_t c5 = 0, c3 = 0;
  for (uint64_t i = lane; i < sr; i += kWL) {
    if (posR[i] == ch.start) {
      ++c5;
      s.rc[s.rr[i]] = cenc(5, kEvCollision | kEvChrom);
    }
  }
  for (uint64_t i = (tf > 0 ? tf : 0) + 1 + lane; i < n; i += kWL) {
    const uint32_t idx = s.fr[i];
    if (s.ep[idx] == kUnbound) {
      ++c3;
    } else if (posF[i] == ch.end - 1) {
      ++c3;
      s.fc[idx] = cenc(3, kEvCollision | kEvChrom);
    }
  }
  if (lane == 0) {
    if (rend) s.rc[s.rr[sr]] = cenc(5, kEvCollision | kEvChrom);
    if (fend) s.fc[s.fr[tf]] = cenc(3, kEvCollision | kEvChrom);
  }
  n5 = w_sum(w, s.xs, c5) + (rend ? 1 : 0);
  n3 = w_sum(w, s.xs, c3) + (fend ? 1 : 0);
}

// Position buckets over a sorted position array: bucket b covers [start + (b << shift),
// start + ((b + 1) << shift)) and first[b] is the first rank whose position lies in bucket b or
// later (positions at or past the chromosome end count as bucket nbk); first[nbk] closes it.
struct WBuckets {
  uint64_t start, end, nbk;
  uint32_t shift;
  __device__ uint64_t of(uint64_t p) const { return p >= end ? nbk : (p - start) >> shift; }
};

template <class W>
__device__ inline void w_build_buckets(const W& w, const WBuckets& B, const uint64_t* pos,
                                       uint64_t n, uint32_t* first) {
  const uint32_t lane = w.lane();
  for (uint64_t k = lane; k < n; k += kWL) {
    const uint64_t bk = B.of(pos[k]);
    const uint64_t b0 = k == 0 ? 0 : B.of(pos[k - 1]) + 1;
    for (uint64_t b = b0; b <= bk; ++b) first[b] = (uint32_t)k;
  }
  const uint64_t last = n == 0 ? 0 : B.of(pos[n - 1]) + 1;
  for (uint64_t b = (n == 0 ? 0 : last) + lane; b <= B.nbk; b += kWL) first[b] = (uint32_t)n;
  w.sync();
}

// LEF-BAR (det_lef_bar, simulation_detect_collisions.cpp:128-252). The rev walk meets the
// active barriers in ascending order, each with the first rev unit beyond it (from rank
// max(n5, 1) - 1 on) and stops at the first barrier with no such unit; a unit keeps the last
// barrier whose trial succeeds. The fwd walk meets them in descending order with the last fwd
// unit before them (from rank n - max(n3, 1)). Each barrier's search bisects only the ranks of
// its position bucket (about one LEF per bucket), found from a per-epoch bucket index of the
// sorted rev and fwd positions.
template <class W>
__device__ inline void w_lef_bar(const W& w, const DetParams& P, const DetChrom& ch, WCell& s,
                                 uint64_t n, uint64_t n5, uint64_t n3) {
  const uint32_t lane = w.lane();
  const uint64_t nb = ch.nbar;
  const uint64_t* posR = s.x0;
  const uint64_t* posF = s.x1;
  uint32_t* win = s.y0;
  // bucket index: a power of two of buckets below n, both sides' first ranks in x2 (2n u32)
  WBuckets B{ch.start, ch.end, 0, 0};
  uint32_t* firstR = reinterpret_cast<uint32_t*>(s.x2);
  uint32_t* firstF = firstR + n;
  const bool bucketed = n >= 4;
  if (bucketed) {
    uint64_t nbk = 1;
    while (nbk * 2 < n) nbk *= 2;  // nbk + 1 <= n entries per side
    B.nbk = nbk;
    const uint64_t range = ch.end - ch.start;
    while (((range - 1) >> B.shift) >= nbk) ++B.shift;
    w_build_buckets(w, B, posR, n, firstR);
    w_build_buckets(w, B, posF, n, firstF);
  }
  for (int side = 0; side < 2; ++side) {
    const bool rev = side == 0;
    const uint8_t major = rev ? kDirRev : kDirFwd;
    for (uint64_t u = lane; u < n; u += kWL) win[u] = rev ? 0u : ~0u;
    const uint64_t j0 = rev ? (n5 > 0 ? n5 - 1 : 0) : n - (n3 > 0 ? n3 - 1 : 0) - 1;
    for (uint64_t i = lane; i < nb; i += kWL) {
      uint32_t cand = ~0u;
      if (s.bst[i]) {
        const uint64_t bpos = ch.bar_pos[i];
        if (rev) {
          uint64_t lo = j0, hi = n;  // first rank with posR > bpos
          if (bucketed) {
            const uint64_t b = B.of(bpos);
            lo = firstR[b] > j0 ? firstR[b] : j0;
            hi = firstR[b + 1] > lo ? firstR[b + 1] : lo;
          }
          while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (posR[mid] <= bpos) lo = mid + 1;
            else hi = mid;
          }
          if (lo < n) {
            const uint32_t u = s.rr[lo];
            if (s.ep[u] != kUnbound && posR[lo] - bpos <= s.rm[u]) cand = u;
          }
        } else {
          uint64_t lo = 0, hi = j0 + 1;  // ranks in [0, j0] with posF < bpos
          if (bucketed) {
            const uint64_t b = B.of(bpos);
            lo = firstF[b] < j0 + 1 ? firstF[b] : j0 + 1;
            hi = firstF[b + 1] < j0 + 1 ? firstF[b + 1] : j0 + 1;
          }
          while (lo < hi) {
            const uint64_t mid = (lo + hi) >> 1;
            if (posF[mid] < bpos) lo = mid + 1;
            else hi = mid;
          }
          if (lo > 0) {
            const uint32_t u = s.fr[lo - 1];
            if (s.ep[u] != kUnbound && bpos - posF[lo - 1] <= s.fm[u]) cand = u;
          }
        }
      }
      s.bsc[i] = cand;
    }
    w.sync();
    auto pblock = [&](uint64_t i) { return ch.bar_dir[i] == major ? P.pblock_major : P.pblock_minor; };
    auto take = [&](uint64_t i) {
      if (rev) atomicMax(&win[s.bsc[i]], (uint32_t)i + 1);
      else atomicMin(&win[s.bsc[i]], (uint32_t)i);
    };
    w_draws(
        w, s.r, nb, !rev, 1, s.err, s.dbg,
        [&](uint64_t i) {
          const double p = pblock(i);
          return s.bsc[i] != ~0u && p != 1.0 && p != 0.0;
        },
        [&](uint64_t i, WSrc& g) -> uint64_t { return det_bernoulli(g, pblock(i)) ? 1 : 0; },
        [&](uint64_t i, uint64_t v) {
          if (v) take(i);
        });
    for (uint64_t i = lane; i < nb; i += kWL)
      if (s.bsc[i] != ~0u && pblock(i) == 1.0) take(i);
    w.sync();
    for (uint64_t u = lane; u < n; u += kWL) {
      if (rev && win[u] != 0) s.rc[u] = cenc(win[u] - 1, kEvCollision | kEvLefBar);
      if (!rev && win[u] != ~0u) s.fc[u] = cenc(win[u], kEvCollision | kEvLefBar);
    }
    w.sync();
  }
}

__device__ inline void w_primary_apply(const DetChrom& ch, WCell& s, uint64_t p, uint64_t q) {
  const uint32_t ridx = s.rr[q], fidx = s.fr[p];
  const uint64_t rpos = s.x0[q], fpos = s.x1[p];
  uint64_t cp_rev, cp_fwd;
  collision_pos(rpos, fpos, s.rm[ridx], s.fm[fidx], cp_rev, cp_fwd);
  const bool rocc = c_any(s.rc[ridx]);
  const bool focc = c_any(s.fc[fidx]);
  if (!rocc && !focc) {
    s.rc[ridx] = cenc(fidx, kEvCollision | kEvPrimary);
    s.fc[fidx] = cenc(ridx, kEvCollision | kEvPrimary);
  } else if (rocc && !focc) {
    const uint64_t bpos = ch.bar_pos[cidx(s.rc[ridx])];
    if (cp_fwd > bpos) s.rc[ridx] = cenc(fidx, kEvCollision | kEvPrimary);
    s.fc[fidx] = cenc(ridx, kEvCollision | kEvPrimary);
  } else if (!rocc && focc) {
    const uint64_t bpos = ch.bar_pos[cidx(s.fc[fidx])];
    s.rc[ridx] = cenc(fidx, kEvCollision | kEvPrimary);
    if (cp_rev < bpos) s.fc[fidx] = cenc(ridx, kEvCollision | kEvPrimary);
  }
}

// Primary LEF-LEF (det_primary, simulation_detect_collisions.cpp:255-401): the walk pairs the
// last fwd unit of every run of fwd units with the first rev unit after it, for fwd ranks
// p + 1 < n - max(n3, 1) + 1 and rev ranks from n5. Pairs are disjoint, so each is applied on
// its own lane once its trial (in pair order) is known.
template <class W>
__device__ inline void w_primary(const W& w, const DetParams& P, const DetChrom& ch, WCell& s,
                                 uint64_t n, uint64_t n5, uint64_t n3) {
  if (n5 == n || n3 == n) return;
  const uint32_t lane = w.lane();
  const uint64_t* posR = s.x0;
  const uint64_t* posF = s.x1;
  uint32_t* prt = s.y1;
  const uint64_t i2 = n - (n3 > 0 ? n3 - 1 : 0);
  const uint64_t np = i2 > 0 ? i2 - 1 : 0;
  // each lane takes a contiguous run of fwd ranks; posF is sorted, so the first rev rank
  // (from n5) beyond posF[p] only grows along the run and is carried from p to p + 1 (a few
  // ranks walked, a wider gap bisected)
  const uint64_t bp = (np + kWL - 1) / kWL;
  const uint64_t p0 = (uint64_t)lane * bp < np ? (uint64_t)lane * bp : np;
  const uint64_t p1 = p0 + bp < np ? p0 + bp : np;
  uint64_t lo = n5;
  for (uint64_t p = p0; p < p1; ++p) {
    const uint64_t fp = posF[p];
    for (int step = 0; step < 8 && lo < n && posR[lo] <= fp; ++step) ++lo;
    if (lo < n && posR[lo] <= fp) {
      uint64_t hi = n;
      while (lo < hi) {
        const uint64_t mid = (lo + hi) >> 1;
        if (posR[mid] <= fp) lo = mid + 1;
        else hi = mid;
      }
    }
    uint32_t part = ~0u;
    if (lo < n && posF[p + 1] >= posR[lo]) {
      const uint32_t ridx = s.rr[lo], fidx = s.fr[p];
      if (posR[lo] - fp < s.rm[ridx] + s.fm[fidx]) part = (uint32_t)lo;
    }
    prt[p] = part;
  }
  w.sync();
  const double pkeep = 1.0 - P.p_bypass;
  if (P.p_bypass == 0.0) {
    for (uint64_t p = lane; p < np; p += kWL)
      if (prt[p] != ~0u) w_primary_apply(ch, s, p, prt[p]);
  } else if (pkeep != 0.0) {
    w_draws(
        w, s.r, np, false, 1, s.err, s.dbg, [&](uint64_t p) { return prt[p] != ~0u; },
        [&](uint64_t, WSrc& g) -> uint64_t { return det_bernoulli(g, pkeep) ? 1 : 0; },
        [&](uint64_t p, uint64_t v) {
          if (v) w_primary_apply(ch, s, p, prt[p]);
        });
  }
  w.sync();
}

template <class W>
__device__ inline void w_correct(const W& w, const DetChrom& ch, WCell& s, uint64_t n) {
  const uint32_t lane = w.lane();
  for (uint64_t i = lane; i < n; i += kWL) {
    if (c_is(s.rc[i], kEvLefBar)) s.rm[i] = (s.rev[i] - ch.bar_pos[cidx(s.rc[i])]) - 1;
    if (c_is(s.fc[i], kEvLefBar)) s.fm[i] = (ch.bar_pos[cidx(s.fc[i])] - s.fwd[i]) - 1;
  }
  w.sync();
  for (uint64_t k = lane; k < n; k += kWL) {
    const uint32_t ridx = s.rr[k];
    if (!c_is(s.rc[ridx], kEvPrimary)) continue;
    const uint64_t fidx = cidx(s.rc[ridx]);
    if (c_is(s.fc[fidx], kEvPrimary)) {
      uint64_t p1, p2;
      collision_pos(s.rev[ridx], s.fwd[fidx], s.rm[ridx], s.fm[fidx], p1, p2);
      s.rm[ridx] = s.rev[ridx] - p1;
      s.fm[fidx] = p2 - s.fwd[fidx];
    } else if (c_is(s.fc[fidx], kEvLefBar)) {
      s.rm[ridx] = s.rev[ridx] - (s.fwd[fidx] + s.fm[fidx]) - 1;
    }
  }
  w.sync();
  for (uint64_t k = lane; k < n; k += kWL) {
    const uint32_t fidx = s.fr[k];
    if (!c_is(s.fc[fidx], kEvPrimary)) continue;
    const uint64_t ridx = cidx(s.fc[fidx]);
    if (c_is(s.rc[ridx], kEvLefBar)) s.fm[fidx] = (s.rev[ridx] - s.rm[ridx]) - s.fwd[fidx] - 1;
  }
  w.sync();
}

// Secondary LEF-LEF (det_secondary, simulation_detect_collisions.cpp:404-518) on one side. A
// unit whose neighbour (toward the side's walk origin) is stalled, that is not stalled itself
// and whose target reaches the neighbour's, takes a trial: success stalls it next to the
// neighbour, failure marks it as bypassing. Stalls only propagate along consecutive ranks, so
// every candidate belongs to the chain of the stalled unit (seed) it follows. The chains are
// walked assuming every trial succeeds; the trials are then drawn in the walk's rank order, and
// a failure drops the rest of its chain.
template <class W>
__device__ inline void w_secondary_side(const W& w, const DetParams& P, WCell& s, uint64_t n,
                                        uint64_t n5, uint64_t n3, bool rev) {
  const uint32_t lane = w.lane();
  const uint64_t* pos = rev ? s.x0 : s.x1;
  const uint32_t* rk = rev ? s.rr : s.fr;
  uint64_t* col = rev ? s.rc : s.fc;
  uint64_t* mv = rev ? s.rm : s.fm;
  uint32_t* chn = s.y0;   // chain (seed rank) of a candidate rank, ~0 otherwise
  uint32_t* lst = s.y1;   // candidates in draw order
  uint32_t* res = s.y2;   // 1 stalled, 2 bypassing, 0 untouched
  uint64_t* tprev = s.x2; // the neighbour's target a candidate compares with
  for (uint64_t j = lane; j < n; j += kWL) {
    chn[j] = ~0u;
    res[j] = 0;
  }
  w.sync();
  auto seed = [&](uint64_t j) { return c_any(col[rk[j]]); };
  auto target = [&](uint64_t j) { return rev ? pos[j] - mv[rk[j]] : pos[j] + mv[rk[j]]; };
  // rev pairs (i - 1, i) for i >= max(1, n5); fwd pairs (i - 1, i) for i <= top, walked down
  const uint64_t top = n - (n3 > 0 ? n3 - 1 : 0) - 1;
  const uint64_t first = n5 > 1 ? n5 : 1;
  for (uint64_t j = lane; j < n; j += kWL) {
    if (!seed(j)) continue;
    uint64_t tau = target(j);
    if (rev) {
      if (j + 1 < first) continue;
      for (uint64_t i = j + 1; i < n && !seed(i) && target(i) <= tau; ++i) {
        chn[i] = (uint32_t)j;
        tprev[i] = tau;
        tau = pos[i] > tau ? tau + 1 : tau;
      }
    } else {
      if (j < 1 || j > top) continue;
      for (uint64_t i = j; i-- > 0;) {
        if (seed(i) || !(target(i) >= tau)) break;
        chn[i] = (uint32_t)j;
        tprev[i] = tau;
        tau = tau > pos[i] ? tau - 1 : tau;
      }
    }
  }
  w.sync();
  // candidates in draw order: ascending ranks (rev), descending (fwd)
  uint64_t nc = 0;
  {
    const uint64_t lt = lanes_below(lane);
    for (uint64_t g = 0; g < n; g += kWL) {
      const uint64_t k = g + lane;
      const uint64_t j = rev ? k : n - 1 - k;
      const bool p = k < n && chn[j] != ~0u;
      const uint64_t m = w.ballot(p);
      if (p) lst[nc + __builtin_popcountll(m & lt)] = (uint32_t)j;
      nc += __builtin_popcountll(m);
    }
    w.sync();
  }
  const double pkeep = 1.0 - P.p_bypass;
  if (P.p_bypass == 0.0) {
    for (uint64_t c = lane; c < nc; c += kWL) res[lst[c]] = 1;
  } else if (pkeep == 0.0) {
    // every trial fails without a draw: the first candidate of each chain bypasses
    for (uint64_t c = lane; c < nc; c += kWL) {
      const uint64_t j = lst[c];
      if (rev ? chn[j] + 1 == j : chn[j] == j + 1) res[j] = 2;
    }
  } else {
    uint64_t idx = 0;
    while (idx < nc) {
      wr_refill(w, s.r);
      const uint64_t c = idx + lane;
      bool ok = false;
      if (c < nc) {
        WSrc g{s.r.ring, s.r.cursor + lane, s.r.hi, s.r.zn};
        ok = det_bernoulli(g, pkeep);
      }
      const uint64_t fails = w.ballot(c < nc && !ok);
      const uint64_t take = fails ? (uint64_t)__builtin_ctzll(fails) + 1
                                  : (nc - idx < kWL ? nc - idx : kWL);
      if (c < nc && lane < take) res[lst[c]] = ok ? 1 : 2;
      s.r.cursor += take;
      idx += take;
      if (fails) {
        wdbg(s.dbg, lane, kDbgBypass);
        const uint32_t ch_f = chn[lst[idx - 1]];
        const uint64_t from = idx;
        while (idx < nc && chn[lst[idx]] == ch_f) ++idx;
        wdbg(s.dbg, lane, kDbgChainDrop, idx - from);
      }
      w.sync();
    }
  }
  w.sync();
  for (uint64_t i = lane; i < n; i += kWL) {
    if (res[i] == 0) continue;
    const uint64_t nb = rev ? i - 1 : i + 1;  // the stalled neighbour's rank
    const uint32_t u = rk[i], v = rk[nb];
    if (res[i] == 1) {
      col[u] = cenc(v, kEvCollision | kEvSecondary);
      const uint64_t d = rev ? pos[i] - tprev[i] : tprev[i] - pos[i];
      mv[u] = umin64(d, d - 1);
    } else {
      col[u] = cenc(v, kEvSecondary);
    }
  }
  w.sync();
}

// Bypass fix-ups (det_fix_secondary, simulation_detect_collisions.cpp:520-647): a
// bypassing unit never neighbours another (its own trial needed a stalled neighbour, and the
// unit after it sees it unstalled), so the rank pairs each fix-up swaps are disjoint.
template <class W>
__device__ inline void w_fix_secondary(const W& w, const DetChrom& ch, WCell& s, uint64_t n,
                                       uint64_t n5, uint64_t n3) {
  const uint32_t lane = w.lane();
  for (uint64_t i = (n5 > 1 ? n5 : 1) + lane; i < n; i += kWL) {
    const uint32_t r2 = s.rr[i];
    if (!c_avoided(s.rc[r2], kEvSecondary)) continue;
    const uint32_t r1 = s.rr[i - 1];
    const uint64_t pos1 = s.rev[r1] - s.rm[r1];
    s.rm[r2] = s.rev[r2] > pos1 + 1 ? s.rev[r2] - (pos1 + 1) : 0;
    s.rc[r2] = cenc(r1, kEvCollision | kEvSecondary);
    const uint64_t p1 = s.rev[r1], p2 = s.rev[r2];
    s.rev[r1] = umin64(s.fwd[r1], p2);
    s.rev[r2] = umin64(s.fwd[r2], p1);
    uint64_t t = s.rc[r1];
    s.rc[r1] = s.rc[r2];
    s.rc[r2] = t;
    t = s.rm[r1];
    s.rm[r1] = s.rm[r2];
    s.rm[r2] = t;
    s.rr[i - 1] = r2;
    s.rr[i] = r1;
    s.rm[r2] = umin64(s.rev[r2] - ch.start, s.rm[r2]);
    s.rm[r1] = umin64(s.rev[r1] - ch.start, s.rm[r1]);
  }
  w.sync();
  const uint64_t n_active_fwd = n - (n3 > 0 ? n3 - 1 : 0);
  for (uint64_t i = lane; i + 1 < n_active_fwd; i += kWL) {
    const uint32_t f1 = s.fr[i];
    if (!c_avoided(s.fc[f1], kEvSecondary)) continue;
    const uint32_t f2 = s.fr[i + 1];
    const uint64_t pos2 = s.fwd[f2] + s.fm[f2];
    s.fm[f1] = pos2 > s.fwd[f1] + 1 ? pos2 - (s.fwd[f1] + 1) : 0;
    s.fc[f1] = cenc(f2, kEvCollision | kEvSecondary);
    const uint64_t p1 = s.fwd[f1], p2 = s.fwd[f2];
    s.fwd[f1] = umax64(s.rev[f1], p2);
    s.fwd[f2] = umax64(s.rev[f2], p1);
    uint64_t t = s.fc[f1];
    s.fc[f1] = s.fc[f2];
    s.fc[f2] = t;
    t = s.fm[f1];
    s.fm[f1] = s.fm[f2];
    s.fm[f2] = t;
    s.fr[i] = f2;
    s.fr[i + 1] = f1;
    s.fm[f2] = umin64(ch.end - 1 - s.fwd[f2], s.fm[f2]);
    s.fm[f1] = umin64(ch.end - 1 - s.fwd[f1], s.fm[f1]);
  }
  w.sync();
}

template <class W>
__device__ inline void w_process_collisions(const W& w, const DetParams& P, const DetChrom& ch,
                                            WCell& s, uint64_t n) {
  for (uint64_t k = w.lane(); k < n; k += kWL) {
    s.x0[k] = s.rev[s.rr[k]];
    s.x1[k] = s.fwd[s.fr[k]];
  }
  w.sync();
  uint64_t n5, n3;
  w_boundaries(w, ch, s, n, n5, n3);
  wmark(s, 6);
  w_lef_bar(w, P, ch, s, n, n5, n3);
  wmark(s, 7);
  w_primary(w, P, ch, s, n, n5, n3);
  w_correct(w, ch, s, n);
  wmark(s, 8);
  w_secondary_side(w, P, s, n, n5, n3, true);
  w_secondary_side(w, P, s, n, n5, n3, false);
  wmark(s, 9);
  w_fix_secondary(w, ch, s, n, n5, n3);
  wmark(s, 10);
}

// ---------------------------------------------------------------------------------------------
// burn-in, binding, contacts, release
// ---------------------------------------------------------------------------------------------
// dips_ratio_stable over a history of cap values (evaluate_burnin, simulation.cpp:983-1025):
// each lane compares its windows' means, each mean summed in the serial order.
template <class W>
__device__ inline bool w_dips_ratio_stable(const W& w, const double* v, uint64_t cap,
                                           uint64_t win) {
  uint64_t cnt = 0;
  const uint64_t nk = cap - 1 >= win + 1 ? cap - 1 - (win + 1) + 1 : 0;  // k + win + 1 <= cap - 1
  for (uint64_t g = 0; g < nk; g += kWL) {
    const uint64_t k = g + w.lane();
    const bool p = k < nk && mean_of(v + k, win) > mean_of(v + k + 1, win);
    cnt += __builtin_popcountll(w.ballot(p));
  }
  const double r = (double)cnt / (double)(cap - win - cnt);
  return r >= 0.95 && r <= 1.05;
}

// run_burnin (simulation.cpp:1027-1054) for one epoch.
template <class W>
__device__ inline void w_burnin(const W& w, const DetParams& P, const DetChrom& ch, WCell& s) {
  const uint32_t lane = w.lane();
  do {
    ++s.burnin_epochs;
    if (s.nact != s.n) {
      const double rate = ch.bind_rate;
      const uint64_t k = w_single(w, s.r, s.err, [&](WSrc& g) { return det_poisson(g, rate); });
      s.nact = umin64(s.nact + k, s.n);
    } else {
      const uint64_t n = s.nact;
      // loop sizes are integers below 2^53, so the serial double sum equals the integer sum
      uint64_t part = 0;
      for (uint64_t i = lane; i < n; i += kWL) part += s.fwd[i] - s.rev[i];
      const uint64_t total = w_sum(w, s.xs, part);
      const double mean = (double)total / (double)n;
      // the squared deviations summed in LEF order, as the reference's serial double sum: each
      // lane holds one chunk's square in a register, and every lane adds them in order
      double ssd = 0.0;
      for (uint64_t g = 0; g < n; g += kWL) {
        const uint64_t i = g + lane;
        const double d = i < n ? (double)(s.fwd[i] - s.rev[i]) - mean : 0.0;
        const double q = d * d;
        uint64_t qb;
        __builtin_memcpy(&qb, &q, 8);
        const uint32_t m = (uint32_t)(n - g < kWL ? n - g : kWL);
        for (uint32_t t = 0; t < m; ++t) {
          const uint64_t vb = w.bcast(qb, t);
          double v;
          __builtin_memcpy(&v, &vb, 8);
          ssd = ssd + v;
        }
      }
      const uint64_t cap = P.hist_len;
      if (s.hlen == cap) {
        // both histories move down one slot, all lanes: a block of values is read, then
        // written one slot lower (a block's writes never reach the next block's reads)
        for (uint64_t b = 1; b < cap; b += 4 * kWL) {
          double va[4], vc[4];
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint64_t i = b + (uint64_t)u * kWL + lane;
            va[u] = i < cap ? s.avg[i] : 0.0;
            vc[u] = i < cap ? s.cfx[i] : 0.0;
          }
          w.sync();
#pragma unroll
          for (int u = 0; u < 4; ++u) {
            const uint64_t i = b + (uint64_t)u * kWL + lane;
            if (i < cap) {
              s.avg[i - 1] = va[u];
              s.cfx[i - 1] = vc[u];
            }
          }
          w.sync();
        }
      }
      if (lane == 0) {
        const double sd = sqrt(ssd / (double)n);
        const uint64_t h = s.hlen == cap ? cap - 1 : s.hlen;
        s.avg[h] = mean;
        s.cfx[h] = sd / mean;
      }
      s.hlen = s.hlen == P.hist_len ? P.hist_len : s.hlen + 1;
      w.sync();
      bool done = s.hlen == P.hist_len;
      if (done) done = w_dips_ratio_stable(w, s.cfx, P.hist_len, P.window);
      if (done) done = w_dips_ratio_stable(w, s.avg, P.hist_len, P.window);
      done = done && s.epoch > P.min_burnin_epochs;
      if (!done && s.epoch >= P.max_burnin_epochs) done = true;
      s.burnin_done = done;
    }
  } while (s.nact == 0);
}

// register_contacts_loop / _tad (register_contacts.cpp:129-198): attempts run 64 at a time, at
// the offsets they would have if every earlier attempt took its full word count (LEF index,
// two GEV draws with noise, two positions for TAD contacts); the batch is cut at the first
// attempt that took another count or at the attempt that completes the count. With `occ`, the
// same attempts as loop contacts register 1D LEF occupancy instead
// (register_1d_lef_occupancy, register_contacts.cpp:200-232): each accepted attempt adds one to
// the occupancy bins of both of its (noisified) unit positions.
template <class W>
__device__ inline void w_register(const W& w, const DetParams& P, const DetChrom& ch, WCell& s,
                                  uint64_t nreg, bool tad, bool occ = false) {
  const uint32_t lane = w.lane();
  const uint64_t sp = ch.start + 1, epp = ch.end - 1;
  const double spd = (double)sp, epd = (double)epp;
  const uint64_t n = s.nact;
  const uint64_t range = n - 1;
  const uint64_t bucket = bucket_of(range);
  const uint64_t F = (range == 0 ? 0 : 1) + (P.noisify ? 2 : 0) + (tad ? 2 : 0);
  uint64_t attempts = 0;
  while (nreg != 0) {
    wr_refill(w, s.r);
    const uint64_t off = s.r.cursor + (uint64_t)lane * F;
    WSrc g{s.r.ring, off, s.r.hi, s.r.zn};
    bool ok = false;
    uint64_t q1 = 0, q2 = 0;
    do {
      const uint64_t i = det_uniform_b(g, 0, range, bucket);
      if (s.ep[i] == kUnbound) break;
      const uint64_t r = s.rev[i], f = s.fwd[i];
      if (!(r > sp && r < epp && f > sp && f < epp)) break;
      const double a = (double)r - (P.noisify ? det_gev(g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
      const double b = (double)f + (P.noisify ? det_gev(g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
      const double p1 = b < a ? b : a;
      const double p2 = a < b ? b : a;
      if (!(p1 >= spd && p2 >= spd && p1 < epd && p2 < epd)) break;
      q1 = (uint64_t)p1;
      q2 = (uint64_t)p2;
      if (tad) {
        const uint64_t lo = q1, hi = q2;
        q1 = det_uniform(g, lo, hi);
        q2 = det_uniform(g, lo, hi);
      }
      ok = true;
    } while (false);
    const uint64_t used = g.pos - off;
    const uint64_t bad = w.ballot(used != F);
    const uint64_t succ = w.ballot(ok);
    const uint32_t kb = bad ? (uint32_t)__builtin_ctzll(bad) : 64;
    if (bad) wdbg(s.dbg, lane, kDbgContactCut);
    const uint64_t upto = kb >= 63 ? ~0ull : ((2ull << kb) - 1);  // lanes <= kb
    uint64_t valid = bad ? upto : ~0ull;
    const uint64_t sv = succ & valid;
    uint32_t last = 0xFFFFFFFFu;  // lane whose success completes the count
    if ((uint64_t)__builtin_popcountll(sv) >= nreg) {
      uint64_t m = sv;
      for (uint64_t t = 1; t < nreg; ++t) m &= m - 1;
      last = (uint32_t)__builtin_ctzll(m);
      valid = last == 63 ? ~0ull : ((2ull << last) - 1);
    }
    if (ok && ((valid >> lane) & 1ull)) {
      if (occ) {
        atomicAdd(&ch.occupancy[(q1 - sp) / P.bin_size], 1ULL);
        atomicAdd(&ch.occupancy[(q2 - sp) / P.bin_size], 1ULL);
      } else {
        band_add(ch, (q1 - sp) / P.bin_size, (q2 - sp) / P.bin_size);
      }
    }
    const uint64_t nsucc = (uint64_t)__builtin_popcountll(succ & valid);
    nreg -= nsucc;
    if (last != 0xFFFFFFFFu) {
      s.r.cursor = w.bcast(g.pos, last);
    } else if (bad) {
      s.r.cursor = w.bcast(g.pos, kb);
    } else {
      s.r.cursor += kWL * F;
    }
    if (s.r.cursor > s.r.hi) s.err |= 1u;
    attempts += kWL;
    if (attempts > kWMaxAttempts) {
      s.err |= 2u;
      break;
    }
  }
  w.sync();
}

template <class W>
__device__ inline void w_sample_contacts(const W& w, const DetParams& P, const DetChrom& ch,
                                         WCell& s, uint64_t target) {
  uint64_t n = ch.cpe;
  if (P.target_contact_density > 0.0) n = umin64(n, target - s.contacts);
  if (n == 0) return;
  uint64_t nloop;
  if (P.tad_to_loop_ratio == 0) {
    nloop = n;
  } else if (isinf(P.tad_to_loop_ratio)) {
    nloop = 0;
  } else {
    const double q = ch.qn ? ch.qn[n] : -1.0;
    const double pl = 1.0 / (P.tad_to_loop_ratio + 1.0);
    nloop = w_single(w, s.r, s.err, [&](WSrc& g) {
      return (uint64_t)det_binomial(g, (int64_t)n, pl, q);
    });
  }
  w_register(w, P, ch, s, nloop, false);
  w_register(w, P, ch, s, n - nloop, true);
  if (ch.occupancy) w_register(w, P, ch, s, n, false, true);
  s.contacts += n;
}

// dump_stats (simulation.cpp:1166-1224) as det_log_state records it: stalls are collisions that
// occurred, the typed counts sum both sides, and the average loop size is the sequential double
// sum over the active LEFs divided by their count. Loop sizes are integers below 2^32, so while
// their total stays below 2^53 every partial sum is exact in any order and the wave's integer
// sum gives the same double; a larger total is summed on lane 0 in LEF order.
template <class W>
__device__ inline void w_log_state(const W& w, const DetParams& P, const DetChrom& ch, WCell& s,
                                   uint32_t item) {
  const uint32_t lane = w.lane();
  uint64_t ab = 0, sr = 0, sf = 0, sb = 0, lb = 0, pr = 0, se = 0, len = 0;
  for (uint64_t i = lane; i < ch.nbar; i += kWL) ab += s.bst[i] == 1;
  for (uint64_t i = lane; i < s.nact; i += kWL) {
    const uint64_t rc = s.rc[i], fc = s.fc[i];
    const bool a = c_any(rc), b = c_any(fc);
    sr += a;
    sf += b;
    sb += a && b;
    lb += (uint64_t)c_is(rc, kEvLefBar) + (uint64_t)c_is(fc, kEvLefBar);
    pr += (uint64_t)c_is(rc, kEvPrimary) + (uint64_t)c_is(fc, kEvPrimary);
    se += (uint64_t)c_is(rc, kEvSecondary) + (uint64_t)c_is(fc, kEvSecondary);
    len += s.fwd[i] - s.rev[i];
  }
  ab = w_sum(w, s.xs, ab);
  sr = w_sum(w, s.xs, sr);
  sf = w_sum(w, s.xs, sf);
  sb = w_sum(w, s.xs, sb);
  lb = w_sum(w, s.xs, lb);
  pr = w_sum(w, s.xs, pr);
  se = w_sum(w, s.xs, se);
  len = w_sum(w, s.xs, len);
  if (lane == 0) {
    double acc = (double)len;
    if (len >= (1ull << 53)) {
      acc = 0.0;
      for (uint64_t i = 0; i < s.nact; ++i) acc = acc + (double)(s.fwd[i] - s.rev[i]);
    }
    const unsigned long long slot = atomicAdd(P.log.count, 1ULL);
    if (slot < P.log.cap) {
      StateRecord r{};
      r.item = item;
      r.epoch = (uint32_t)s.epoch;
      r.burnin = s.burnin_done ? 0u : 1u;
      r.nact = (uint32_t)s.nact;
      r.active_barriers = (uint32_t)ab;
      r.stalls_rev = (uint32_t)sr;
      r.stalls_fwd = (uint32_t)sf;
      r.stalls_both = (uint32_t)sb;
      r.lef_bar = (uint32_t)lb;
      r.primary = (uint32_t)pr;
      r.secondary = (uint32_t)se;
      r.avg_loop_size = s.nact == 0 ? 0.0 : acc / (double)s.nact;
      P.log.rec[slot] = r;
    }
  }
  w.sync();
}

template <class W>
__device__ inline void w_release(const W& w, const DetParams& P, const DetChrom& ch, WCell& s) {
  const double base = s.burnin_done ? P.prob_release : P.prob_release_burnin;
  auto prob = [&](uint64_t i) {
    const int h = (c_is(s.rc[i], kEvLefBar) && ch.bar_dir[cidx(s.rc[i])] == kDirRev) +
                  (c_is(s.fc[i], kEvLefBar) && ch.bar_dir[cidx(s.fc[i])] == kDirFwd);
    const double aff = h == 0 ? 1.0 : (h == 1 ? P.inv_soft : P.inv_hard);
    return aff * base;
  };
  w_draws(
      w, s.r, s.nact, false, 1, s.err, s.dbg,
      [&](uint64_t i) { return s.ep[i] != kUnbound && prob(i) != 0.0; },
      [&](uint64_t i, WSrc& g) -> uint64_t { return det_bernoulli(g, prob(i)) ? 1 : 0; },
      [&](uint64_t i, uint64_t v) {
        if (v) s.rev[i] = s.fwd[i] = s.ep[i] = kUnbound;
      });
}

// simulate_one_cell (simulation.cpp:1056-1152) on one wavefront. base: LDS of det_wave_bytes(),
// or of det_wave_fixed_bytes() when cell_mem (det_wave_cell_bytes() of device memory) holds the
// cell's arrays (kPlaceMem, or kPlaceAny with cell_mem not null).
template <class W, int kPlace = kPlaceAny>
__device__ inline void det_wave_cell(const W& w, const DetParams& P, const DetChrom& ch,
                                     uint8_t* base, uint8_t* cell_mem, uint64_t hist,
                                     uint64_t seed, uint64_t target,
                                     uint64_t* trace, uint64_t max_trace, uint64_t* stats,
                                     uint32_t item, uint64_t* dbg = nullptr) {
  const uint32_t lane = w.lane();
  WCell s;
  det_wave_carve<kPlace>(s, base, cell_mem, ch.nlefs, ch.nbar, hist);
  s.dbg = dbg;
  const uint64_t N = ch.nlefs;
  s.n = N;
  s.nact = 0;
  s.hlen = 0;
  s.epoch = 0;
  s.burnin_epochs = 0;
  s.contacts = 0;
  s.burnin_done = false;
  s.err = 0;
  s.timed = P.phase != nullptr;
  s.tlast = s.timed ? wclock() : 0;
  for (int k = 0; k < 12; ++k) s.tacc[k] = 0;
  for (uint64_t i = lane; i < N; i += kWL) {
    s.rev[i] = s.fwd[i] = s.ep[i] = kUnbound;
    s.rr[i] = s.fr[i] = (uint32_t)i;
  }
  {
    double* zn = const_cast<double*>(s.r.zn);
    for (uint32_t i = lane; i < 129; i += kWL) {
      zn[i] = kZigNormalX[i];
      zn[129 + i] = kZigNormalY[i];
    }
  }
  wr_init(w, s.r, seed, s.xs);
  // initial barrier states: Bernoulli(occupancy), no draw where the occupancy is 0
  auto occ = [&](uint64_t i) {
    const double sa = ch.bar_stp_active[i], si = ch.bar_stp_inactive[i];
    double o = 0.0;
    if (sa + si != 0) {
      const double tia = 1.0 - si, tai = 1.0 - sa;
      o = tia / (tia + tai);
      o = o < 0.0 ? 0.0 : (o > 1.0 ? 1.0 : o);
    }
    return o;
  };
  for (uint64_t i = lane; i < ch.nbar; i += kWL) s.bst[i] = 0;
  w.sync();
  w_draws(
      w, s.r, ch.nbar, false, 1, s.err, s.dbg, [&](uint64_t i) { return occ(i) != 0.0; },
      [&](uint64_t i, WSrc& g) -> uint64_t { return det_bernoulli(g, occ(i)) ? 1 : 0; },
      [&](uint64_t i, uint64_t v) { s.bst[i] = v ? 1 : 0; });
  if (P.skip_burnin) {
    s.nact = N;
    s.burnin_done = true;
  }
  const uint64_t bind_range = ch.end - 1 - ch.start;
  const uint64_t bind_bucket = bucket_of(bind_range);
  for (;; ++s.epoch) {
    const bool go = P.target_contact_density >= 0 ? s.contacts < target
                                                  : s.epoch - s.burnin_epochs < ch.target_epochs;
    if (!go) break;
    if (!s.burnin_done) w_burnin(w, P, ch, s);
    wmark(s, 0);
    const uint64_t n = s.nact;
    // select + bind: every unbound LEF of the active span draws a position in [start, end-1]
    if (bind_range == 0) {
      for (uint64_t i = lane; i < n; i += kWL) {
        if (s.ep[i] != kUnbound) continue;
        s.rev[i] = s.fwd[i] = ch.start;
        s.ep[i] = s.epoch;
      }
      w.sync();
    } else {
      w_draws(
          w, s.r, n, false, 1, s.err, s.dbg, [&](uint64_t i) { return s.ep[i] == kUnbound; },
          [&](uint64_t, WSrc& g) -> uint64_t {
            return det_uniform_b(g, ch.start, bind_range, bind_bucket);
          },
          [&](uint64_t i, uint64_t v) {
            s.rev[i] = s.fwd[i] = v;
            s.ep[i] = s.epoch;
          });
    }
    wmark(s, 1);
    w_rank(w, s, n, s.epoch != 0);
    wmark(s, 2);
    if (s.burnin_done) {
      w_sample_contacts(w, P, ch, s, target);
      wmark(s, 3);
      if (target != 0 && s.contacts >= target) break;
    }
    w_generate_moves(w, P, ch, s, n);
    wmark(s, 4);
    w_draws(
        w, s.r, ch.nbar, false, 1, s.err, s.dbg, [&](uint64_t) { return true; },
        [&](uint64_t, WSrc& g) -> uint64_t {
          const double u = det_canonical(g);
          uint64_t b;
          __builtin_memcpy(&b, &u, 8);
          return b;
        },
        [&](uint64_t i, uint64_t b) {
          double u;
          __builtin_memcpy(&u, &b, 8);
          if (s.bst[i] == 0 && u > ch.bar_stp_inactive[i]) s.bst[i] = 1;
          else if (s.bst[i] == 1 && u > ch.bar_stp_active[i]) s.bst[i] = 0;
        });
    wmark(s, 5);
    for (uint64_t i = lane; i < n; i += kWL) s.rc[i] = s.fc[i] = 0;
    w.sync();
    w_process_collisions(w, P, ch, s, n);
    for (uint64_t i = lane; i < n; i += kWL) {
      if (s.ep[i] == kUnbound) continue;
      s.rev[i] -= s.rm[i];
      s.fwd[i] += s.fm[i];
    }
    w.sync();
    if (P.log.rec) w_log_state(w, P, ch, s, item);
    w_release(w, P, ch, s);
    if (trace && s.epoch < max_trace) {
      if (lane == 0) {
        uint64_t h = 0x84222325cbf29ce4ULL ^ s.r.cursor;
        for (uint64_t i = 0; i < s.nact; ++i) {
          h = (h ^ s.rev[i]) * 0x100000001b3ULL;
          h = (h ^ s.fwd[i]) * 0x100000001b3ULL;
        }
        trace[s.epoch] = h ^ (s.contacts << 1);
      }
      w.sync();
    }
    wmark(s, 11);
    if (s.err) break;
  }
  if (lane == 0 && s.timed)
    for (int k = 0; k < 12; ++k) atomicAdd(&P.phase[k], (unsigned long long)s.tacc[k]);
  if (lane == 0 && stats) {
    stats[0] = s.epoch;
    stats[1] = s.burnin_epochs;
    stats[2] = s.contacts;
    stats[3] = s.err ? ~0ull : s.r.cursor;
  }
  w.sync();
}

}  // namespace mgpu

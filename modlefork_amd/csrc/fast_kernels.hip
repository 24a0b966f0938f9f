// Fast-mode kernels: one workgroup per cell behind a device work queue (see fast_sim.h).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <mutex>
#include <string>

#include "fast_kernels.h"
#include "fast_math.h"
#include "fast_sim.h"

namespace mgpu {

namespace {
// The lane's index in the workgroup, read through a volatile move: every index computation
// derived from it stays inside the epoch phase that uses it. With threadIdx.x itself the compiler
// hoists dozens of per-lane index and address values out of the epoch loop, where they occupy
// VGPRs for the whole cell and spill.
__device__ inline uint32_t tidx() {
  uint32_t t;
  asm volatile("v_mov_b32 %0, %1" : "=v"(t) : "v"((uint32_t)threadIdx.x));
  return t;
}

// ---------------------------------------------------------------------------------------------
// block primitives (NT lanes, wave64)
// ---------------------------------------------------------------------------------------------
struct OpMaxI {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a > b ? a : b; }
};
struct OpMinI {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a < b ? a : b; }
};
struct OpAddI {
  __device__ int32_t operator()(int32_t a, int32_t b) const { return a + b; }
};
struct OpAddU {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a + b; }
};
struct OpMaxU {
  __device__ uint32_t operator()(uint32_t a, uint32_t b) const { return a > b ? a : b; }
};

// A 32-bit value of lane (lane - shift) through DPP (lanes with no source keep `ident`): the
// data-parallel-primitive moves run on the VALU, not on the LDS pipe as ds_bpermute does.
template <int kCtrl, int kRowMask, class T>
__device__ __forceinline__ T dpp_move(T v, T ident) {
  static_assert(sizeof(T) == 4, "32-bit values");
  return __builtin_bit_cast(T, __builtin_amdgcn_update_dpp(__builtin_bit_cast(int, ident),
                                                           __builtin_bit_cast(int, v), kCtrl,
                                                           kRowMask, 0xf, false));
}

// Inclusive scan across the 64 lanes of a wave: shifts by 1, 2, 4 and 8 within each row of 16
// lanes (row_shr), then lane 15 of a row into the next row (row_bcast:15, rows 1 and 3) and lane
// 31 into the upper half (row_bcast:31, rows 2 and 3).
template <class T, class Op>
__device__ __forceinline__ T wave_inclusive(T x, Op op, T ident) {
  x = op(x, dpp_move<0x111, 0xf>(x, ident));
  x = op(x, dpp_move<0x112, 0xf>(x, ident));
  x = op(x, dpp_move<0x114, 0xf>(x, ident));
  x = op(x, dpp_move<0x118, 0xf>(x, ident));
  x = op(x, dpp_move<0x142, 0xa>(x, ident));
  x = op(x, dpp_move<0x143, 0xc>(x, ident));
  return x;
}

// The lane below's value (wave_shr:1), ident in lane 0.
template <class T>
__device__ __forceinline__ T wave_shift1(T x, T ident) {
  return dpp_move<0x138, 0xf>(x, ident);
}

// Exclusive scan of one value per lane across the block; sh holds NT / 64 entries.
template <int NT, class T, class Op>
__device__ T block_exclusive(T v, Op op, T ident, T* sh) {
  const int lane = tidx() & 63, w = tidx() >> 6;
  const T incl = wave_inclusive(v, op, ident);
  if (lane == 63) sh[w] = incl;
  __syncthreads();
  T pre = ident;
  for (int k = 0; k < w; ++k) pre = op(pre, sh[k]);
  const T ex = wave_shift1(incl, ident);
  __syncthreads();
  return op(pre, ex);
}

// Exclusive scans of two values per lane across the block (a with opa, b with opb) sharing one
// pair of barriers; sh holds 2 * NT / 64 entries.
template <int NT, class T, class OpA, class OpB>
__device__ void block_exclusive2(T& a, T& b, OpA opa, OpB opb, T ia, T ib, T* sh) {
  const int lane = tidx() & 63, w = tidx() >> 6;
  const T xa = wave_inclusive(a, opa, ia), xb = wave_inclusive(b, opb, ib);
  if (lane == 63) {
    sh[w] = xa;
    sh[NT / 64 + w] = xb;
  }
  __syncthreads();
  T pa = ia, pb = ib;
  for (int k = 0; k < w; ++k) {
    pa = opa(pa, sh[k]);
    pb = opb(pb, sh[NT / 64 + k]);
  }
  const T ea = wave_shift1(xa, ia), eb = wave_shift1(xb, ib);
  __syncthreads();
  a = opa(pa, ea);
  b = opb(pb, eb);
}

// Sum over the block: a butterfly within each wave, then every wave reduces the NT / 64 wave
// sums the same way (a sequential loop over them, unrolled, held 16 u64 loads in flight and
// made this the register-pressure peak of the epoch loop).
template <int NT>
__device__ uint64_t block_sum_u64(uint64_t v, uint64_t* sh) {
  constexpr int NW = NT / 64;
  for (int d = 32; d >= 1; d >>= 1) v += __shfl_xor(v, d, 64);
  if ((tidx() & 63) == 0) sh[tidx() >> 6] = v;
  __syncthreads();
  uint64_t t = (tidx() & 63) < NW ? sh[tidx() & 63] : 0;
  for (int d = NW / 2; d >= 1; d >>= 1) t += __shfl_xor(t, d, 64);
  __syncthreads();
  return t;
}

// Three sums over the block with one pair of barriers; sh holds 3 * NT / 64 entries.
template <int NT>
__device__ void block_sum3_u64(uint64_t& a, uint64_t& b, uint64_t& c, uint64_t* sh) {
  constexpr int NW = NT / 64;
  for (int d = 32; d >= 1; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  const int lane = tidx() & 63, w = tidx() >> 6;
  if (lane == 0) {
    sh[w] = a;
    sh[NW + w] = b;
    sh[2 * NW + w] = c;
  }
  __syncthreads();
  a = lane < NW ? sh[lane] : 0;
  b = lane < NW ? sh[NW + lane] : 0;
  c = lane < NW ? sh[2 * NW + lane] : 0;
  for (int d = NW / 2; d >= 1; d >>= 1) {
    a += __shfl_xor(a, d, 64);
    b += __shfl_xor(b, d, 64);
    c += __shfl_xor(c, d, 64);
  }
  __syncthreads();
}

__device__ inline uint64_t pos_key(const uint32_t* pos, uint32_t lef) {
  return ((uint64_t)pos[lef] << 32) | lef;
}

// ---------------------------------------------------------------------------------------------
// per-cell working set (LDS, or a global slab for cells that do not fit)
// ---------------------------------------------------------------------------------------------
constexpr uint32_t kNewCap = 1024;  // LEFs merged into the rank order per batch

struct Cell {
  uint32_t *rev, *fwd, *rm, *fm, *rc, *fc;
  uint16_t *rr, *fr;
  uint32_t* scratch;  // scratch_bytes(n): new-LEF keys, sorted positions, chain words
  uint8_t* bst;
  double *cfx, *avg;  // burn-in history
};

__device__ __host__ inline uint64_t align16(uint64_t x) { return (x + 15) & ~15ULL; }

// n u32 words, or the u64 keys of one merge batch (at most min(n, kNewCap) keys, padded to a
// power of two for the bitonic sort), whichever is larger
__device__ __host__ inline uint64_t scratch_bytes(uint32_t n) {
  uint64_t m = 1;
  while (m < (n < kNewCap ? n : kNewCap)) m <<= 1;
  return 4ull * n > 8ull * m ? 4ull * n : 8ull * m;
}

__device__ __host__ inline uint64_t fast_cell_bytes(uint32_t n, uint32_t nbar, uint32_t hist) {
  uint64_t b = align16(4ull * n) * 6 + align16(2ull * n) * 2;
  b += align16(scratch_bytes(n));
  b += align16(nbar) + align16(8ull * (hist + 1)) * 2;
  return b;
}

__device__ inline void carve(Cell& c, uint8_t* base, uint32_t n, uint32_t nbar, uint32_t hist) {
  uint8_t* p = base;
  auto take = [&](uint64_t bytes) {
    uint8_t* r = p;
    p += align16(bytes);
    return r;
  };
  c.rev = (uint32_t*)take(4ull * n);
  c.fwd = (uint32_t*)take(4ull * n);
  c.rm = (uint32_t*)take(4ull * n);
  c.fm = (uint32_t*)take(4ull * n);
  c.rc = (uint32_t*)take(4ull * n);
  c.fc = (uint32_t*)take(4ull * n);
  c.rr = (uint16_t*)take(2ull * n);
  c.fr = (uint16_t*)take(2ull * n);
  c.scratch = (uint32_t*)take(scratch_bytes(n));
  c.bst = take(nbar);
  c.cfx = (double*)take(8ull * (hist + 1));
  c.avg = (double*)take(8ull * (hist + 1));
}

// Working sets too large for LDS keep the arrays every phase gathers from (positions, ranks,
// barrier states, burn-in history) in LDS and the per-epoch temporaries (moves, collisions,
// scratch) in the workgroup's slab of device memory.
__device__ __host__ inline uint64_t hybrid_lds_bytes(uint32_t n, uint32_t nbar, uint32_t hist) {
  return align16(4ull * n) * 2 + align16(2ull * n) * 2 + align16(nbar) +
         align16(8ull * (hist + 1)) * 2;
}


__device__ inline void carve_hybrid(Cell& c, uint8_t* lds, uint8_t* slab, uint32_t n, uint32_t nbar,
                                    uint32_t hist) {
  uint8_t* p = lds;
  uint8_t* q = slab;
  auto take = [](uint8_t*& at, uint64_t bytes) {
    uint8_t* r = at;
    at += align16(bytes);
    return r;
  };
  c.rev = (uint32_t*)take(p, 4ull * n);
  c.fwd = (uint32_t*)take(p, 4ull * n);
  c.rr = (uint16_t*)take(p, 2ull * n);
  c.fr = (uint16_t*)take(p, 2ull * n);
  c.bst = take(p, nbar);
  c.cfx = (double*)take(p, 8ull * (hist + 1));
  c.avg = (double*)take(p, 8ull * (hist + 1));
  c.rm = (uint32_t*)take(q, 4ull * n);
  c.fm = (uint32_t*)take(q, 4ull * n);
  c.rc = (uint32_t*)take(q, 4ull * n);
  c.fc = (uint32_t*)take(q, 4ull * n);
  c.scratch = (uint32_t*)take(q, scratch_bytes(n));
}

// where a launch keeps its cells' working sets
constexpr int kStoreSlab = 0, kStoreLds = 1, kStoreHybrid = 2;

struct Shared {
  int32_t shi[kFastMaxWaves];
  uint32_t shu[kFastMaxWaves];
  uint64_t phase[kFastPhases];
  uint64_t tick;
  uint64_t item;
  uint64_t burnin_epochs;
  uint64_t shx[3 * kFastMaxWaves];
  uint32_t nact, hlen, hhead, burnin_done, nnew, nkept, n5, n3, stable, navoid, navoid_fwd;
};

// bitonic sort of keys[0, m) ascending, m a power of two
// A value every lane of the block holds equal, moved to scalar registers (lane 0's copy).
__device__ inline uint64_t uniform64(uint64_t v) {
  const uint32_t lo = __builtin_amdgcn_readfirstlane((uint32_t)v);
  const uint32_t hi = __builtin_amdgcn_readfirstlane((uint32_t)(v >> 32));
  return ((uint64_t)hi << 32) | lo;
}

// LEF-BAR records a fwd barrier collision as kCkBarFwdTmp with the index mirrored, so that
// atomicMax keeps the closest (lowest-index) barrier; this gives the plain kCkBar word.
__device__ inline uint32_t fc_final(uint32_t v) {
  return ck_kind(v) == kCkBarFwdTmp ? ck(kCkBar, kCkIdxMask - ck_idx(v)) : v;
}

template <int NT>
__device__ void bitonic(uint64_t* keys, uint32_t m) {
  for (uint32_t k = 2; k <= m; k <<= 1) {
    for (uint32_t j = k >> 1; j > 0; j >>= 1) {
      for (uint32_t i = tidx(); i < m; i += NT) {
        const uint32_t l = i ^ j;
        if (l > i) {
          const uint64_t a = keys[i], b = keys[l];
          if ((a > b) == ((i & k) == 0)) {
            keys[i] = b;
            keys[l] = a;
          }
        }
      }
      __syncthreads();
    }
  }
}

// Sorts LEF ids a[0, n) by position with odd-even transposition passes (stable: equal
// positions keep their order); returns after two consecutive passes without a swap. Cheap on
// the nearly sorted orders extrusion leaves (a few adjacent inversions per epoch): one check
// pass first, then a handful of passes. rerank_both runs the same passes on both sides at once.
template <int NT>
__device__ void odd_even_sort(uint32_t* a, uint32_t n, const uint32_t* pos) {
  int bad = 0;
  for (uint32_t i = tidx(); i + 1 < n; i += NT) bad |= pos[a[i + 1]] < pos[a[i]];
  if (!__syncthreads_or(bad)) return;
  int quiet = 0;
  for (uint32_t pass = 0; quiet < 2; ++pass) {
    int swapped = 0;
    for (uint32_t i = 2 * tidx() + (pass & 1); i + 1 < n; i += 2 * NT) {
      const uint32_t x = a[i], y = a[i + 1];
      if (pos[y] < pos[x]) {
        a[i] = y;
        a[i + 1] = x;
        swapped = 1;
      }
    }
    quiet = __syncthreads_or(swapped) ? 0 : quiet + 1;
  }
}

// Sorts the k distinct keys of keys[0, k) ascending: by counting ranks (one lane per key, every
// lane reading the same key at a time) when the block has a lane per key, else bitonic over the
// next power of two (keys[k, m) padded with the largest key).
template <int NT>
__device__ void sort_new_keys(uint64_t* keys, uint32_t k) {
  if (k <= NT && k <= 512) {
    uint64_t mine = 0;
    uint32_t r = 0;
    if (tidx() < k) {
      mine = keys[tidx()];
      for (uint32_t j = 0; j < k; ++j) r += keys[j] < mine;
    }
    __syncthreads();
    if (tidx() < k) keys[r] = mine;
    __syncthreads();
    return;
  }
  uint32_t m = 1;
  while (m < k) m <<= 1;
  for (uint32_t i = k + tidx(); i < m; i += NT) keys[i] = ~0ULL;
  __syncthreads();
  bitonic<NT>(keys, m);
}

// Merges the kept LEF ids kept[0, nk) (sorted by position) with the new keys keys[0, k)
// (sorted by position, then id) into out(i, lef) (a merge path: each lane owns a contiguous run
// of the kept order, finds where its run starts among the new keys once, then walks). At equal
// positions the new LEFs go after the kept ones, or before them when kNewFirst. The new keys
// that fall before kept entry a (and after a - 1) are written by a's lane, those after the
// last kept entry by the last lane.
template <int NT, bool kNewFirst, class Out>
__device__ void merge_kept_new(const uint32_t* kept, uint32_t nk, const uint64_t* keys, uint32_t k,
                               const uint32_t* pos, Out out) {
  if (nk == 0) {
    for (uint32_t b = tidx(); b < k; b += NT) out(b, (uint32_t)keys[b]);
    return;
  }
  // whether new key q goes before a kept LEF at position p (true on a prefix of the keys)
  auto before = [](uint64_t q, uint32_t p) {
    return kNewFirst ? (uint32_t)(q >> 32) <= p : (uint32_t)(q >> 32) < p;
  };
  const uint32_t chunk = (nk + NT - 1) / NT;
  const uint32_t ab = tidx() * chunk, ae = min(nk, ab + chunk);
  if (ab >= ae) return;
  // lo: new keys that go before kept[ab - 1] (0 for the first run)
  uint32_t lo = 0;
  if (ab > 0) {
    const uint32_t p = pos[kept[ab - 1]];
    uint32_t hi = k;
    while (lo < hi) {
      const uint32_t mid = (lo + hi) >> 1;
      if (before(keys[mid], p)) lo = mid + 1;
      else hi = mid;
    }
  }
  for (uint32_t a = ab; a < ae; ++a) {
    const uint32_t lef = kept[a];
    const uint32_t p = pos[lef];
    while (lo < k && before(keys[lo], p)) {
      out(lo + a, (uint32_t)keys[lo]);
      ++lo;
    }
    out(a + lo, lef);
  }
  if (ae == nk) {
    for (; lo < k; ++lo) out(lo + nk, (uint32_t)keys[lo]);
  }
}

static_assert(kFastPhases == kFastPhaseCount, "phase count of the timers and of the host API");
__device__ unsigned long long g_fast_phase[kFastPhases];

// Phase timers: called by every lane right after the barrier that ends a phase; lane 0 adds the
// wall-clock ticks since the previous mark to the phase's total.
__device__ inline void mark(const FastParams& P, Shared& S, int ph) {
  if (P.timers && tidx() == 0) {
    const uint64_t t = wall_clock64();
    S.phase[ph] += t - S.tick;
    S.tick = t;
  }
}

// merge_kept_new for both unit sides at once, their searches and walks interleaved in each lane
// (independent dependency chains): kept orders kr (rev positions, new keys after kept ones at
// ties) and kf (fwd positions, new keys first at ties), results in c.rr and c.fr.
template <int NT>
__device__ void merge_both(const uint32_t* kr, const uint32_t* kf, uint32_t nk,
                           const uint64_t* keys, uint32_t k, Cell& c) {
  if (nk == 0) {
    for (uint32_t b = tidx(); b < k; b += NT) c.rr[b] = c.fr[b] = (uint16_t)keys[b];
    return;
  }
  auto before_r = [](uint64_t q, uint32_t p) { return (uint32_t)(q >> 32) < p; };
  auto before_f = [](uint64_t q, uint32_t p) { return (uint32_t)(q >> 32) <= p; };
  const uint32_t chunk = (nk + NT - 1) / NT;
  const uint32_t ab = tidx() * chunk, ae = min(nk, ab + chunk);
  if (ab >= ae) return;
  uint32_t lor = 0, lof = 0;
  if (ab > 0) {
    const uint32_t pr = c.rev[kr[ab - 1]], pf = c.fwd[kf[ab - 1]];
    uint32_t hir = k, hif = k;
    while (lor < hir || lof < hif) {
      if (lor < hir) {
        const uint32_t mid = (lor + hir) >> 1;
        if (before_r(keys[mid], pr)) lor = mid + 1;
        else hir = mid;
      }
      if (lof < hif) {
        const uint32_t mid = (lof + hif) >> 1;
        if (before_f(keys[mid], pf)) lof = mid + 1;
        else hif = mid;
      }
    }
  }
  for (uint32_t a = ab; a < ae; ++a) {
    const uint32_t lr = kr[a], lf = kf[a];
    const uint32_t pr = c.rev[lr], pf = c.fwd[lf];
    while (lor < k && before_r(keys[lor], pr)) {
      c.rr[lor + a] = (uint16_t)keys[lor];
      ++lor;
    }
    c.rr[a + lor] = (uint16_t)lr;
    while (lof < k && before_f(keys[lof], pf)) {
      c.fr[lof + a] = (uint16_t)keys[lof];
      ++lof;
    }
    c.fr[a + lof] = (uint16_t)lf;
  }
  if (ae == nk) {
    for (; lor < k; ++lor) c.rr[lor + nk] = (uint16_t)keys[lor];
    for (; lof < k; ++lof) c.fr[lof + nk] = (uint16_t)keys[lof];
  }
}

// Merges the LEFs flagged in is_new into one unit side's order: tmp_a[0, nkept) holds the kept
// entries of last epoch's order (compacted stably); the result goes to ranks[0, nact). More
// than kNewCap new LEFs are merged in batches through tmp_b (n-sized scratch). kNewFirst: new
// LEFs go before kept ones at equal positions (the fwd side).
template <int NT, bool kNewFirst>
__device__ void merge_new(uint16_t* ranks, const uint32_t* pos, uint32_t nact, uint32_t nkept,
                          const uint32_t* is_new, uint32_t* tmp_a, uint32_t* tmp_b,
                          uint64_t* keys, Shared& S) {
  // Extrusion keeps neighbours ordered almost always, but the merge below needs the kept order
  // sorted exactly.
  odd_even_sort<NT>(tmp_a, nkept, pos);
  // new LEFs, merged batch by batch (one batch unless more than kNewCap were bound)
  const uint32_t batches = (nkept + kNewCap > nact || nact - nkept <= kNewCap)
                               ? 1u
                               : (nact + kNewCap - 1) / kNewCap;
  bool written = false;
  for (uint32_t bt = 0; bt < batches; ++bt) {
    const uint32_t lo_id = batches == 1 ? 0 : bt * kNewCap;
    const uint32_t hi_id = batches == 1 ? nact : min(nact, lo_id + kNewCap);
    if (tidx() == 0) S.nnew = 0;
    __syncthreads();
    for (uint32_t l = lo_id + tidx(); l < hi_id; l += NT) {
      if (is_new[l]) keys[atomicAdd(&S.nnew, 1u)] = pos_key(pos, l);
    }
    __syncthreads();
    const uint32_t k = S.nnew;
    if (k == 0) continue;
    sort_new_keys<NT>(keys, k);
    // the last batch merges straight into the rank array (nothing reads it any more)
    if (bt + 1 == batches) {
      merge_kept_new<NT, kNewFirst>(tmp_a, nkept, keys, k, pos,
                                    [&](uint32_t i, uint32_t lef) { ranks[i] = (uint16_t)lef; });
      written = true;
      __syncthreads();
    } else {
      merge_kept_new<NT, kNewFirst>(tmp_a, nkept, keys, k, pos,
                                    [&](uint32_t i, uint32_t lef) { tmp_b[i] = lef; });
      __syncthreads();
      nkept += k;
      for (uint32_t i = tidx(); i < nkept; i += NT) tmp_a[i] = tmp_b[i];
      __syncthreads();
    }
  }
  if (!written) {
    for (uint32_t i = tidx(); i < nact; i += NT) ranks[i] = (uint16_t)tmp_a[i];
    __syncthreads();
  }
}

// Re-ranks both unit sides (rank_lefs for rev and fwd units). ranks[0, nact_old) hold last
// epoch's orders; LEFs flagged in c.rm were (re)bound this epoch. Both kept orders are
// compacted in one pass (their counts packed into one scan). A LEF bound this epoch sits at one
// position on both sides (rev == fwd), so the new LEFs have the same keys on both sides: they
// are collected and sorted once and merged into both orders in one pass. Epochs that bind more
// than kNewCap LEFs (the first ones) merge each side in batches instead.
// Ties (rank_lefs orders equal positions by binding epoch: earliest first on the rev side,
// latest first on the fwd side): kept LEFs keep last epoch's order among themselves, and the
// LEFs bound this epoch, the latest bound, go after them on the rev side and before them on the
// fwd side. So a kept order with equal positions stays sorted, and the units stacked at a
// chromosome end need no sorting pass.
template <int NT>
__device__ void rerank_both(Cell& c, uint32_t nact, uint32_t nact_old, uint64_t* keys,
                            Shared& S, const FastParams& P) {
  const uint32_t* is_new = c.rm;
  uint32_t* kr = c.rc;
  uint32_t* kf = c.fc;
  {
    const uint32_t ch = (nact_old + NT - 1) / NT;
    const uint32_t b = tidx() * ch, e = min(nact_old, b + ch);
    // kept counts of this lane's chunk: rev side in the low 16 bits, fwd side in the high 16
    // (at most 65535 LEFs per cell, so the low half cannot carry into the high one)
    uint32_t cnt = 0;
    for (uint32_t i = b; i < e; ++i)
      cnt += (is_new[c.rr[i]] == 0 ? 1u : 0u) + (is_new[c.fr[i]] == 0 ? 0x10000u : 0u);
    const uint32_t off = block_exclusive<NT>(cnt, OpAddU{}, 0u, S.shu);
    uint32_t orv = off & 0xFFFFu, ofw = off >> 16;
    for (uint32_t i = b; i < e; ++i) {
      const uint32_t lr = c.rr[i], lf = c.fr[i];
      if (is_new[lr] == 0) kr[orv++] = lr;
      if (is_new[lf] == 0) kf[ofw++] = lf;
    }
    if (tidx() == NT - 1) S.nkept = orv;  // the same LEFs are kept on both sides
    if (tidx() == 0) S.nnew = 0;
    __syncthreads();
  }
  const uint32_t nkept = S.nkept;
  if (nact - nkept > kNewCap) {
    merge_new<NT, false>(c.rr, c.rev, nact, nkept, is_new, kr, c.fm, keys, S);
    mark(P, S, 2);
    merge_new<NT, true>(c.fr, c.fwd, nact, nkept, is_new, kf, c.fm, keys, S);
    mark(P, S, 3);
    return;
  }
  // both kept orders must be sorted exactly for the merge; one check pass covers both
  int bad = 0;
  for (uint32_t i = tidx(); i + 1 < nkept; i += NT)
    bad |= (c.rev[kr[i + 1]] < c.rev[kr[i]]) | (c.fwd[kf[i + 1]] < c.fwd[kf[i]]);
  for (uint32_t l = tidx(); l < nact; l += NT) {
    if (is_new[l]) keys[atomicAdd(&S.nnew, 1u)] = pos_key(c.rev, l);
  }
  if (__syncthreads_or(bad)) {
    // odd-even transposition passes over both orders at once (stable), until two consecutive
    // passes swap nothing on either side
    int quiet = 0;
    for (uint32_t pass = 0; quiet < 2; ++pass) {
      int swapped = 0;
      for (uint32_t i = 2 * tidx() + (pass & 1); i + 1 < nkept; i += 2 * NT) {
        const uint32_t x = kr[i], y = kr[i + 1];
        if (c.rev[y] < c.rev[x]) {
          kr[i] = y;
          kr[i + 1] = x;
          swapped = 1;
        }
        const uint32_t u = kf[i], v = kf[i + 1];
        if (c.fwd[v] < c.fwd[u]) {
          kf[i] = v;
          kf[i + 1] = u;
          swapped = 1;
        }
      }
      quiet = __syncthreads_or(swapped) ? 0 : quiet + 1;
    }
  }
  const uint32_t k = S.nnew;
  if (k == 0) {
    for (uint32_t i = tidx(); i < nact; i += NT) {
      c.rr[i] = (uint16_t)kr[i];
      c.fr[i] = (uint16_t)kf[i];
    }
  } else {
    sort_new_keys<NT>(keys, k);
    merge_both<NT>(kr, kf, nkept, keys, k, c);
  }
  __syncthreads();
  mark(P, S, 2);
}

// Generalised extreme value noise (genextreme_value_distribution.hpp:87-105) in double
// precision from a 53-bit uniform: mu + sigma * (1 - (-ln u)^xi) / xi, written with expm1 so the
// small-xi cancellation stays exact; xi == 0 takes the reference's (mu - sigma) * ln(-ln u).
__device__ inline double gev_noise(const FastParams& P, uint32_t a, uint32_t b) {
  const double y = ln_d(-ln_d(u01d_open(a, b)));
  if (P.gev_xi == 0.0) return (P.gev_mu - P.gev_sigma) * y;
  return P.gev_mu - P.gev_sigma * expm1_d(P.gev_xi * y) / P.gev_xi;
}

__device__ inline void band_add(const FastChrom& ch, uint32_t row, uint32_t col) {
  const uint32_t i = row > col ? row - col : col - row;
  const uint32_t j = row > col ? row : col;
  if (i >= ch.nrows) {
    atomicAdd(ch.missed, 1ULL);
    return;
  }
  atomicAdd(&ch.band[(uint64_t)j * ch.nrows + i], 1u);
}

// Poisson(mean) LEF activations of burn-in epoch `epoch`; `sub` numbers the repeated draws of
// an epoch that activated none (the first epoch repeats until one LEF is active).
// exp_neg and sd are e^-mean and sqrt(mean), computed once per chromosome on the host.
__device__ inline uint32_t poisson(const Rng& R, double mean, double exp_neg, double sd,
                                   uint32_t epoch, uint32_t sub) {
  const U4 d = R.draw(0, epoch, kSPoisson, sub);
  if (mean > 500.0) {  // normal approximation far from the inversion range
    const double r = sqrt(-2.0 * ln_d(u01d_open(d.x, d.y)));
    double sn, cs;
    sincos_2pi_d(u01d(d.z, d.w), &sn, &cs);
    const double v = mean + sd * r * cs;
    return v < 0 ? 0u : (uint32_t)(v + 0.5);
  }
  // inversion by sequential search
  const double u = ((double)d.x * 4294967296.0 + (double)d.y) * (1.0 / 18446744073709551616.0);
  double p = exp_neg, s = p;
  uint32_t k = 0;
  while (u > s && k < 4096) {
    ++k;
    p *= mean / (double)k;
    s += p;
  }
  return k;
}

// The secondary LEF-LEF pass of both sides at once (a side's pass reads and writes only that
// side's ranks, positions, moves and collisions). Chain index j runs along the direction in
// which stalls propagate (rev: rank order; fwd: reverse rank order). Seeds are units already
// stalled; unit j behind seed s stalls when every unit in (s, j] reaches its predecessor's
// target and keeps its collision (does not bypass). w(j) = (s + 1) << 16 | (last failure + 1)
// packs both, so a single prefix max yields the last failure inside j's own chain.
// Each lane owns the same contiguous run of j on both sides, and the two sides share every
// scan's barriers. Seeds are re-read where needed (a unit is rewritten only at its own j, after
// its seed test, and never when it is a seed) and only the 16-bit failure flags are stored, so
// both sides fit in scratch (F: 2 x n u16). Returns the bypassing units' ranks in F[0, na[0])
// (rev, chain order) and F[n, n + na[1]) (fwd, chain order).
template <int NT>
__device__ void secondary_both(const FastParams& P, const Rng& R, uint32_t epoch, Cell& c,
                               Shared& S, uint32_t n, uint32_t lo_rev, uint32_t lo_fwd,
                               uint16_t* F, uint32_t* na) {
  const uint32_t chunk = (n + NT - 1) / NT;
  const uint32_t jb = tidx() * chunk, je = min(n, jb + chunk);
  auto unit = [&](int sd, uint32_t j) -> uint32_t { return sd == 0 ? c.rr[j] : c.fr[n - 1 - j]; };
  auto coll = [&](int sd) -> uint32_t* { return sd == 0 ? c.rc : c.fc; };
  auto target = [&](int sd, uint32_t l) -> int32_t {
    return sd == 0 ? (int32_t)c.rev[l] - (int32_t)c.rm[l] : (int32_t)c.fwd[l] + (int32_t)c.fm[l];
  };
  auto seed = [&](int sd, uint32_t j) -> uint32_t {
    return ck_stalled(coll(sd)[unit(sd, j)]) ? j + 1 : 0u;
  };
  // the last seed of this lane's run, per side. A run of at most 32 units keeps its seed flags
  // in a bit mask, which the later passes read instead of the collisions.
  const bool use_mask = chunk <= 32;
  uint32_t mask0 = 0, mask1 = 0;
  uint32_t e0 = 0, e1 = 0;
  if (use_mask) {
    for (uint32_t j = jb; j < je; ++j) {
      mask0 |= (seed(0, j) ? 1u : 0u) << (j - jb);
      mask1 |= (seed(1, j) ? 1u : 0u) << (j - jb);
    }
    e0 = mask0 ? jb + (31u - (uint32_t)__clz((int)mask0)) + 1 : 0u;
    e1 = mask1 ? jb + (31u - (uint32_t)__clz((int)mask1)) + 1 : 0u;
  } else {
    for (uint32_t j = je; j-- > jb;) {
      if (e0 == 0) e0 = seed(0, j);
      if (e1 == 0) e1 = seed(1, j);
      if (e0 && e1) break;
    }
  }
  auto seed_at = [&](int sd, uint32_t j) -> uint32_t {
    if (use_mask) return (((sd == 0 ? mask0 : mask1) >> (j - jb)) & 1u) ? j + 1 : 0u;
    return seed(sd, j);
  };
  uint32_t* shu2 = reinterpret_cast<uint32_t*>(S.shx);
  block_exclusive2<NT>(e0, e1, OpMaxU{}, OpMaxU{}, 0u, 0u, shu2);
  const uint32_t s0 = e0, s1p = e1;  // seed prefixes before this run, kept for the last pass
  // failure flags; the packed w's prefix max over this run
  uint32_t m0 = 0, m1 = 0;
  for (uint32_t j = jb; j < je; ++j) {
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      uint32_t& run = sd == 0 ? e0 : e1;
      const uint32_t sj = seed_at(sd, j);
      run = sj > run ? sj : run;
      const uint32_t s1 = run;  // seed + 1, 0 if none
      uint32_t fail = 0;
      if (s1 != 0 && s1 - 1 != j) {
        bool ok = j >= (sd == 0 ? lo_rev : lo_fwd);
        if (ok) {
          const uint32_t l = unit(sd, j);
          const int32_t ts = target(sd, unit(sd, s1 - 1));
          const int32_t dist = (int32_t)j - (int32_t)(s1 - 1);
          const bool reach = sd == 0 ? target(0, l) <= ts + dist - 1 : target(1, l) >= ts - dist + 1;
          // the bypass trial only for units that reach the one ahead of them
          ok = reach && bern(R.draw(l, epoch, sd == 0 ? kSSecondary : kSSecondary + 16).x,
                             P.t_bypass_keep);
        }
        fail = ok ? 0 : j + 1;
      }
      F[sd * n + j] = (uint16_t)fail;
      const uint32_t w = (s1 << 16) | fail;
      uint32_t& m = sd == 0 ? m0 : m1;
      m = w > m ? w : m;
    }
  }
  block_exclusive2<NT>(m0, m1, OpMaxU{}, OpMaxU{}, 0u, 0u, shu2);
  // corrections; m is the prefix max of w through j, prev through j - 1
  e0 = s0;
  e1 = s1p;
  uint32_t navoid = 0;  // rev count in the low 16 bits, fwd in the high 16
  for (uint32_t j = jb; j < je; ++j) {
#pragma unroll
    for (int sd = 0; sd < 2; ++sd) {
      uint32_t& run = sd == 0 ? e0 : e1;
      uint32_t& m = sd == 0 ? m0 : m1;
      const uint32_t sj = seed_at(sd, j);
      run = sj > run ? sj : run;
      const uint32_t prev = m;
      const uint32_t w = (run << 16) | F[sd * n + j];
      m = w > m ? w : m;
      const uint32_t s1 = m >> 16;
      if (s1 == 0 || s1 - 1 == j || j < (sd == 0 ? lo_rev : lo_fwd)) continue;
      if (sj) continue;  // already stalled (its seed flag)
      const uint32_t l = unit(sd, j);
      uint32_t* cl = coll(sd);
      const uint32_t s = s1 - 1;
      const uint32_t lead = unit(sd, j - 1);
      const int32_t dist = (int32_t)j - (int32_t)s;
      const int32_t ts = target(sd, unit(sd, s));
      const uint32_t last_fail = m & 0xFFFFu;
      if (last_fail == 0) {
        // stalled one bp behind the unit ahead of it in the chain
        if (sd == 0) {
          const int32_t mm = (int32_t)c.rev[l] - (ts + dist);
          c.rm[l] = mm > 0 ? (uint32_t)mm : 0u;
        } else {
          const int32_t mm = (ts - dist) - (int32_t)c.fwd[l];
          c.fm[l] = mm > 0 ? (uint32_t)mm : 0u;
        }
        cl[l] = ck(kCkSecondary, lead);
      } else if (last_fail == j + 1 && (prev & 0xFFFFu) == 0) {
        const bool reach = sd == 0 ? target(0, l) <= ts + dist - 1 : target(1, l) >= ts - dist + 1;
        if (reach) {
          cl[l] = ck(kCkAvoided, lead);
          navoid += sd == 0 ? 1u : 0x10000u;
        }
      }
    }
  }
  // bypassing units, compacted in chain order (the scan's barriers also end every lane's reads
  // of F, which the lists overwrite)
  const uint32_t off = block_exclusive<NT>(navoid, OpAddU{}, 0u, S.shu);
  uint32_t o0 = off & 0xFFFFu, o1 = off >> 16;
  if (navoid & 0xFFFFu) {
    for (uint32_t j = jb; j < je; ++j)
      if (ck_kind(c.rc[unit(0, j)]) == kCkAvoided) F[o0++] = (uint16_t)j;
  }
  if (navoid >> 16) {
    for (uint32_t j = jb; j < je; ++j)
      if (ck_kind(c.fc[unit(1, j)]) == kCkAvoided) F[n + o1++] = (uint16_t)(n - 1 - j);
  }
  if (tidx() == NT - 1) {
    S.navoid = o0;
    S.navoid_fwd = o1;
  }
  __syncthreads();
  na[0] = S.navoid;
  na[1] = S.navoid_fwd;
}

// fix_secondary_lef_lef_collisions (simulation_detect_collisions.cpp:520-647) for the rev unit
// at rank k: the bypassing unit swaps places with the one it passed.
__device__ inline void fix_one_rev(Cell& c, const FastChrom& ch, uint32_t k) {
  const uint32_t r2 = c.rr[k];
  if (k == 0 || ck_kind(c.rc[r2]) != kCkAvoided) return;
  const uint32_t r1 = c.rr[k - 1];
  const uint32_t p1t = c.rev[r1] - c.rm[r1];
  c.rm[r2] = c.rev[r2] > p1t + 1 ? c.rev[r2] - (p1t + 1) : 0;
  c.rc[r2] = ck(kCkSecondary, r1);
  const uint32_t a1 = c.rev[r1], b1 = c.rev[r2];
  c.rev[r1] = min(c.fwd[r1], b1);
  c.rev[r2] = min(c.fwd[r2], a1);
  uint32_t t = c.rc[r1];
  c.rc[r1] = c.rc[r2];
  c.rc[r2] = t;
  t = c.rm[r1];
  c.rm[r1] = c.rm[r2];
  c.rm[r2] = t;
  c.rr[k - 1] = (uint16_t)r2;
  c.rr[k] = (uint16_t)r1;
  c.rm[r2] = min(c.rev[r2] - ch.start, c.rm[r2]);
  c.rm[r1] = min(c.rev[r1] - ch.start, c.rm[r1]);
}

// The same for the fwd unit at rank k, which swaps with the unit at rank k + 1.
__device__ inline void fix_one_fwd(Cell& c, const FastChrom& ch, uint32_t n, uint32_t k) {
  const uint32_t f1 = c.fr[k];
  if (k + 1 >= n || ck_kind(c.fc[f1]) != kCkAvoided) return;
  const uint32_t f2 = c.fr[k + 1];
  const uint32_t p2t = c.fwd[f2] + c.fm[f2];
  c.fm[f1] = p2t > c.fwd[f1] + 1 ? p2t - (c.fwd[f1] + 1) : 0;
  c.fc[f1] = ck(kCkSecondary, f2);
  const uint32_t a1 = c.fwd[f1], b1 = c.fwd[f2];
  c.fwd[f1] = max(c.rev[f1], b1);
  c.fwd[f2] = max(c.rev[f2], a1);
  uint32_t t = c.fc[f1];
  c.fc[f1] = c.fc[f2];
  c.fc[f2] = t;
  t = c.fm[f1];
  c.fm[f1] = c.fm[f2];
  c.fm[f2] = t;
  c.fr[k] = (uint16_t)f2;
  c.fr[k + 1] = (uint16_t)f1;
  c.fm[f2] = min(ch.end - 1 - c.fwd[f2], c.fm[f2]);
  c.fm[f1] = min(ch.end - 1 - c.fwd[f1], c.fm[f1]);
}

// The fix-ups of one side, in the reference's order: ascending rank. list holds the bypassing
// units' ranks in chain order (ascending for rev, descending for fwd). A fix-up at rank k touches
// ranks k - 1 and k (rev) or k and k + 1 (fwd) only, so fix-ups depend on each other only along
// runs of consecutive ranks: each run goes to one lane, which walks it in ascending rank, and
// distinct runs (separated by at least one rank) proceed in parallel. serial != 0 walks the
// whole list on lane 0 instead (the order of the reference loop, kept to check the runs against).
template <int NT>
__device__ void fix_avoided(Cell& c, const FastChrom& ch, uint32_t n, const uint16_t* list,
                            uint32_t cnt, bool rev_side, bool serial) {
  if (serial) {
    if (tidx() == 0) {
      if (rev_side) {
        for (uint32_t a = 0; a < cnt; ++a) fix_one_rev(c, ch, list[a]);
      } else {
        for (uint32_t a = cnt; a-- > 0;) fix_one_fwd(c, ch, n, list[a]);
      }
    }
    return;
  }
  for (uint32_t a = tidx(); a < cnt; a += NT) {
    if (rev_side) {
      if (a != 0 && list[a - 1] + 1 == list[a]) continue;  // not the lowest rank of its run
      for (uint32_t b = a;; ++b) {
        fix_one_rev(c, ch, list[b]);
        if (b + 1 >= cnt || list[b + 1] != list[b] + 1) break;
      }
    } else {
      if (a + 1 < cnt && list[a + 1] + 1 == list[a]) continue;
      for (uint32_t b = a;; --b) {
        fix_one_fwd(c, ch, n, list[b]);
        if (b == 0 || list[b - 1] != list[b] + 1) break;
      }
    }
  }
}

// dump_stats (simulation.cpp:1166-1224) for one epoch, on the state after the move corrections
// and before extrusion applies them: stalls are collisions that occurred, the typed counts sum
// both sides, loop sizes are taken at the extruded positions. Lane 0 writes the record.
template <int NT>
__device__ __attribute__((always_inline)) void log_state(const FastParams& P, const FastChrom& ch,
                                                    Cell& c, Shared& S, uint32_t n,
                                                    uint32_t epoch, bool bdone, uint32_t item) {
  uint64_t st_rev = 0, st_fwd = 0, st_both = 0, n_bar = 0, n_pri = 0, n_sec = 0, loops = 0;
  for (uint32_t i = tidx(); i < n; i += NT) {
    const uint32_t kr = ck_kind(c.rc[i]), kf = ck_kind(c.fc[i]);
    const bool sr = ck_stalled(c.rc[i]), sf = ck_stalled(c.fc[i]);
    st_rev += sr;
    st_fwd += sf;
    st_both += sr && sf;
    n_bar += (kr == kCkBar) + (kf == kCkBar);
    n_pri += (kr == kCkPrimary) + (kf == kCkPrimary);
    n_sec += (kr == kCkSecondary) + (kf == kCkSecondary);
    loops += (uint64_t)((c.fwd[i] + c.fm[i]) - (c.rev[i] - c.rm[i]));
  }
  uint64_t n_act = 0;
  for (uint32_t b = tidx(); b < ch.nbar; b += NT) n_act += c.bst[b];
  // counts < 2^17 and loop sums < 2^48: two counts share one 64-bit reduction
  const uint64_t a0 = block_sum_u64<NT>(st_rev | (st_fwd << 32), S.shx);
  const uint64_t a1 = block_sum_u64<NT>(st_both | (n_act << 32), S.shx);
  const uint64_t a2 = block_sum_u64<NT>(n_bar | (n_pri << 32), S.shx);
  const uint64_t a3 = block_sum_u64<NT>(n_sec, S.shx);
  const uint64_t a4 = block_sum_u64<NT>(loops, S.shx);
  if (tidx() == 0) {
    const unsigned long long slot = atomicAdd(P.log.count, 1ULL);
    if (slot < P.log.cap) {
      StateRecord r{};
      r.item = item;
      r.epoch = epoch;
      r.burnin = bdone ? 0u : 1u;
      r.nact = n;
      r.stalls_rev = (uint32_t)a0;
      r.stalls_fwd = (uint32_t)(a0 >> 32);
      r.stalls_both = (uint32_t)a1;
      r.active_barriers = (uint32_t)(a1 >> 32);
      r.lef_bar = (uint32_t)a2;
      r.primary = (uint32_t)(a2 >> 32);
      r.secondary = (uint32_t)a3;
      r.avg_loop_size = n ? (double)a4 / (double)n : 0.0;
      P.log.rec[slot] = r;
    }
  }
}

// ---------------------------------------------------------------------------------------------
// the cell loop
// ---------------------------------------------------------------------------------------------
// kOut: the optional outputs (1D occupancy, state log) are compiled in; the plain variant keeps
// their code, and the registers it would hold, out of the epoch loop.
template <int NT, bool kOut>
__device__ __attribute__((always_inline)) void fast_cell(const FastParams& P, const FastChrom& ch,
                                                         Cell& c, Shared& S,
                          uint64_t seed, uint64_t target, uint64_t* stats, uint32_t item) {
  const uint32_t N = ch.nlefs, NB = ch.nbar;
  const Rng R{(uint32_t)seed, (uint32_t)(seed >> 32)};
  target = uniform64(target);
  for (uint32_t i = tidx(); i < N; i += NT) {
    c.rev[i] = c.fwd[i] = kFreePos;
    c.rr[i] = c.fr[i] = (uint16_t)i;
  }
  for (uint32_t b = tidx(); b < NB; b += NT) {
    const U4 d = R.draw(b, 0xFFFFFFFFu, kSBarInit);
    const uint32_t t = ch.bar_t_occ[b];
    c.bst[b] = (d.x < t || t == kThresholdOne) ? 1 : 0;
  }
  if (tidx() == 0) {
    S.nact = P.skip_burnin ? N : 0;
    S.burnin_done = P.skip_burnin ? 1 : 0;
    S.hlen = 0;
    S.hhead = 0;
    S.burnin_epochs = 0;
    for (int k = 0; k < kFastPhases; ++k) S.phase[k] = 0;
    S.tick = P.timers ? wall_clock64() : 0;
  }
  __syncthreads();
  const uint32_t span = ch.end - ch.start;  // bind positions in [start, end - 1]
  // Position buckets for the LEF-BAR and primary searches, kept in the burn-in history's memory
  // once burn-in is over: T[q] = the first rank whose position lies in bucket q or later
  // (bucket = (position - start) >> bshift, T[nbk] = n), so a search narrows to one bucket.
  uint16_t* const T = reinterpret_cast<uint16_t*>(c.cfx);
  const uint32_t nbk_cap = (uint32_t)align16(8ull * (P.hist_len + 1)) - 1;  // cfx and avg, u16
  uint32_t bshift = 0;
  while (bshift < 31 && ((span - 1) >> bshift) + 1 > nbk_cap) ++bshift;
  const uint32_t nbk = min(((span - 1) >> bshift) + 1, nbk_cap);
  auto bucket = [&](uint32_t p) { return min((p - ch.start) >> bshift, nbk - 1); };
  uint32_t nact_prev = 0;
  uint64_t contacts = 0;  // registered so far, the same in every lane (kept in scalar registers)
  uint32_t epoch = 0;
  for (;; ++epoch) {
    // ---- stop test and burn-in (simulation.cpp:1087-1098, 1027-1054) ----
    // every lane holds the contact count and tests on its own (the same values everywhere);
    // lane 0 changes burnin_epochs only after later barriers
    if (P.target_contact_density >= 0 ? contacts >= target
                                      : (uint64_t)epoch - S.burnin_epochs >= ch.target_epochs)
      break;
    if (!S.burnin_done) {
      const bool activating = S.nact != N;
      __syncthreads();  // every lane has read nact before lane 0 changes it
      if (activating) {
        if (tidx() == 0) {
          uint32_t sub = 0;
          do {
            ++S.burnin_epochs;
            S.nact = min(S.nact + poisson(R, ch.bind_rate, ch.bind_exp_neg, ch.bind_sd, epoch,
                                          sub++), N);
          } while (S.nact == 0);
        }
        __syncthreads();
      } else {
        // mean and population standard deviation of the loop sizes, unbound LEFs counting as
        // 0 (compute_loop_size_stats, simulation.cpp:957-981), from exact integer sums so they
        // do not depend on the reduction order: x < 2^32, per-lane sums of x^2 are split into
        // 32-bit halves that cannot overflow across the block
        uint64_t sx = 0, sq_lo = 0, sq_hi = 0;
        for (uint32_t i = tidx(); i < N; i += NT) {
          const uint64_t x = c.rev[i] == kFreePos ? 0 : (uint64_t)(c.fwd[i] - c.rev[i]);
          const uint64_t x2 = x * x;
          sx += x;
          sq_lo += x2 & 0xFFFFFFFFull;
          sq_hi += x2 >> 32;
        }
        block_sum3_u64<NT>(sx, sq_lo, sq_hi, S.shx);
        const unsigned __int128 sxx = ((unsigned __int128)sq_hi << 32) + sq_lo;
        const unsigned __int128 num = (unsigned __int128)N * sxx - (unsigned __int128)sx * sx;
        const double num_d = (double)(uint64_t)(num >> 64) * 18446744073709551616.0 +
                             (double)(uint64_t)num;  // N^2 * variance
        const double mean = (double)sx / (double)N;
        const double sd = sqrt(num_d / ((double)N * (double)N));
        const uint32_t cap = P.hist_len, w = P.window;
        // the history is a ring: logical entry i (0 = oldest) sits at (hhead + i) % cap, so a
        // full history drops its oldest entry without moving the others
        if (tidx() == 0) {
          ++S.burnin_epochs;
          uint32_t slot;
          if (S.hlen == cap) {
            slot = S.hhead;
            S.hhead = S.hhead + 1 == cap ? 0 : S.hhead + 1;
          } else {
            slot = S.hlen++;
          }
          c.avg[slot] = mean;
          c.cfx[slot] = sd / mean;
        }
        __syncthreads();
        uint32_t stable = 0;
        if (S.hlen == cap) {
          // windows whose mean drops, for the CV and for the mean (evaluate_burnin :983-1025);
          // both counts (at most cap <= 4096) share one scan
          const uint32_t head = S.hhead;
          int32_t d2 = 0;
          for (uint32_t k = tidx(); k + w + 1 <= cap - 1; k += NT) {
            double a1 = 0, a2 = 0, b1 = 0, b2 = 0;
            uint32_t q = head + k >= cap ? head + k - cap : head + k;
            for (uint32_t j = 0; j < w; ++j) {
              const uint32_t q1 = q + 1 == cap ? 0 : q + 1;
              a1 += c.cfx[q];
              a2 += c.cfx[q1];
              b1 += c.avg[q];
              b2 += c.avg[q1];
              q = q1;
            }
            d2 += (a1 > a2 ? 1 : 0) + (b1 > b2 ? 1 << 16 : 0);
          }
          const int32_t tot = block_exclusive<NT>(d2, OpAddI{}, 0, S.shi) + d2;
          const int32_t nc = tot & 0xFFFF, na = tot >> 16;
          if (tidx() == NT - 1) {
            const double r1 = (double)nc / (double)((int32_t)(cap - w) - nc);
            const double r2 = (double)na / (double)((int32_t)(cap - w) - na);
            S.stable = (r1 >= 0.95 && r1 <= 1.05 && r2 >= 0.95 && r2 <= 1.05) ? 1 : 0;
          }
          __syncthreads();
          stable = S.stable;
        }
        if (tidx() == 0) {
          bool done = stable && epoch > P.min_burnin_epochs;
          // --max-burnin-epochs, and a hard stop so a cell that never stabilises cannot keep a
          // workgroup busy indefinitely
          if (!done && (epoch >= P.max_burnin_epochs || epoch >= kFastBurninCap)) done = true;
          S.burnin_done = done ? 1 : 0;
        }
        __syncthreads();
      }
    }
    const uint32_t n = S.nact;
    const bool bdone = S.burnin_done != 0;
    mark(P, S, 0);

    // ---- bind (simulation_impl.hpp:28-89) ----
    // rm holds the "bound this epoch" flags until moves are drawn. A LEF released last epoch was
    // rebound, and flagged, right at its release (its bind draw is keyed by this epoch), so
    // this pass only binds the LEFs activated this epoch (all of them in the first epoch); it is
    // skipped once burn-in has activated every LEF.
    if (n > nact_prev) {
      for (uint32_t i = nact_prev + tidx(); i < n; i += NT) {
        uint32_t flag = 0;
        if (c.rev[i] == kFreePos) {
          const U4 d = R.draw(i, epoch, kSBind);
          c.rev[i] = c.fwd[i] = ch.start + ubounded(d.x, span);
          flag = 1;
        }
        c.rm[i] = flag;
      }
      __syncthreads();
    }
    mark(P, S, 1);
    // ---- ranks (rank_lefs, simulation.cpp:402-487) ----
    uint64_t* keys = reinterpret_cast<uint64_t*>(c.scratch);
    rerank_both<NT>(c, n, nact_prev, keys, S, P);
    nact_prev = n;

    // ---- contacts (sample_and_register_contacts, register_contacts.cpp:94-198) ----
    if (bdone) {
      uint64_t want = ch.cpe;
      if (P.target_contact_density > 0.0) want = min(want, target - contacts);
      const uint32_t sp = ch.start + 1, ep = ch.end - 1;
      uint32_t dropped = 0, dropped_occ = 0;
      for (uint32_t s = tidx(); s < want; s += NT) {
        uint32_t attempt = 0;
        for (; attempt < kFastMaxAttempts; ++attempt) {
          const U4 d = R.draw(s, epoch, kSContact, attempt);
          const uint32_t i = ubounded(d.x, n);
          const uint32_t r = c.rev[i], f = c.fwd[i];
          if (!(r > sp && r < ep && f > sp && f < ep)) continue;
          double p1 = (double)r, p2 = (double)f;
          if (P.noisify) {
            const U4 g = R.draw(s, epoch, kSContact, attempt | 0x40000000u);
            p1 -= gev_noise(P, g.x, g.y);
            p2 += gev_noise(P, g.z, g.w);
          }
          if (p2 < p1) {
            const double t = p1;
            p1 = p2;
            p2 = t;
          }
          if (!(p1 >= (double)sp && p2 < (double)ep)) continue;
          uint32_t q1 = (uint32_t)p1, q2 = (uint32_t)p2;
          if ((uint64_t)d.w >= P.t_loop) {  // TAD contact: two uniform points inside the loop
            const U4 e = R.draw(s, epoch, kSContact, attempt | 0x80000000u);
            const uint32_t range = q2 - q1 + 1;
            const uint32_t a = q1 + ubounded(e.x, range);
            q2 = q1 + ubounded(e.y, range);
            q1 = a;
          }
          band_add(ch, (q1 - sp) / P.bin_size, (q2 - sp) / P.bin_size);
          break;
        }
        dropped += attempt == kFastMaxAttempts;
        // 1D LEF occupancy (register_1d_lef_occupancy, register_contacts.cpp:200-232): one more
        // LEF per sampling event, both of its (noisified) units
        if (kOut && ch.occupancy) {
          uint32_t a2 = 0;
          for (; a2 < kFastMaxAttempts; ++a2) {
            const U4 d = R.draw(s, epoch, kSOccupancy, a2);
            const uint32_t i = ubounded(d.x, n);
            const uint32_t r = c.rev[i], f = c.fwd[i];
            if (!(r > sp && r < ep && f > sp && f < ep)) continue;
            double p1 = (double)r, p2 = (double)f;
            if (P.noisify) {
              const U4 g = R.draw(s, epoch, kSOccupancy, a2 | 0x40000000u);
              p1 -= gev_noise(P, g.x, g.y);
              p2 += gev_noise(P, g.z, g.w);
            }
            if (p2 < p1) {
              const double t = p1;
              p1 = p2;
              p2 = t;
            }
            if (!(p1 >= (double)sp && p2 < (double)ep)) continue;
            atomicAdd(&ch.occupancy[((uint32_t)p1 - sp) / P.bin_size], 1ULL);
            atomicAdd(&ch.occupancy[((uint32_t)p2 - sp) / P.bin_size], 1ULL);
            break;
          }
          dropped_occ += a2 == kFastMaxAttempts;
        }
      }
      // a sample that found no LEF inside the chromosome in kFastMaxAttempts draws is counted,
      // so that band + missed + dropped equals the contacts the cell was due to register
      if (dropped) atomicAdd(&ch.dropped[0], (unsigned long long)dropped);
      if (dropped_occ) atomicAdd(&ch.dropped[1], (unsigned long long)dropped_occ);
      // no barrier: the phases up to extrusion do not write the positions sampled here, and a
      // cell that ends here meets the barrier after it before the next cell is set up
      contacts = uniform64(contacts + want);
      mark(P, S, 4);
      if (target != 0 && contacts >= target) break;
    }

    // ---- moves (generate_moves, simulation.cpp:265-323) ----
    // round(max(0, N(v, sigma))) in double precision, both sides from one Box-Muller pair of
    // 53-bit uniforms; moves are capped at kFastMaxMove (beyond any chromosome fast mode takes)
    const double vr = bdone ? P.rev_speed : P.rev_speed_burnin;
    const double vf = bdone ? P.fwd_speed : P.fwd_speed_burnin;
    const bool noisy_moves = P.rev_std != 0.0 || P.fwd_std != 0.0;
    for (uint32_t i = tidx(); i < n; i += NT) {
      double a = vr, b = vf;
      if (noisy_moves) {
        const U4 d = R.draw(i, epoch, kSMove);
        const double rad = sqrt(-2.0 * ln_d(u01d_open(d.x, d.y)));
        double sn, cs;
        sincos_2pi_d(u01d(d.z, d.w), &sn, &cs);
        if (P.rev_std != 0.0) a += P.rev_std * rad * cs;
        if (P.fwd_std != 0.0) b += P.fwd_std * rad * sn;
      }
      c.rm[i] = (uint32_t)fmin(round(fmax(a, 0.0)), (double)kFastMaxMove);
      c.fm[i] = (uint32_t)fmin(round(fmax(b, 0.0)), (double)kFastMaxMove);
    }
    __syncthreads();
    mark(P, S, 5);
    // adjust (simulation.cpp:342-399): rev targets must fall strictly along the rev order and
    // fwd targets rise strictly along the fwd order: T[k] = target(rr[k]) + k takes a reverse
    // prefix min, Tf[k] = target(fr[k]) - k a prefix max, and the targets are then clamped to
    // the chromosome. Both sides share one scan; each lane scans a contiguous run of ranks (from
    // the top for rev) and applies the prefix to the moves as it goes.
    {
      int32_t* T = reinterpret_cast<int32_t*>(c.scratch);
      int32_t* Tf = reinterpret_cast<int32_t*>(c.fc);  // collisions are cleared below
      for (uint32_t k = tidx(); k < n; k += NT) {
        const uint32_t l = c.rr[k], l2 = c.fr[k];
        T[k] = (int32_t)c.rev[l] - (int32_t)c.rm[l] + (int32_t)k;
        Tf[k] = (int32_t)c.fwd[l2] + (int32_t)c.fm[l2] - (int32_t)k;
      }
      __syncthreads();
      const uint32_t chunk = (n + NT - 1) / NT;
      const uint32_t ib = tidx() * chunk, ie = min(n, ib + chunk);
      int32_t ar = INT32_MAX, af = INT32_MIN;
      for (uint32_t i = ib; i < ie; ++i) {
        ar = min(ar, T[n - 1 - i]);
        af = max(af, Tf[i]);
      }
      block_exclusive2<NT>(ar, af, OpMinI{}, OpMaxI{}, INT32_MAX, INT32_MIN,
                           reinterpret_cast<int32_t*>(S.shx));
      for (uint32_t i = ib; i < ie; ++i) {
        const uint32_t k = n - 1 - i;
        ar = min(ar, T[k]);
        const uint32_t l = c.rr[k];
        int32_t t = ar - (int32_t)k;
        if (t < (int32_t)ch.start) t = (int32_t)ch.start;
        c.rm[l] = c.rev[l] - (uint32_t)t;
        c.rc[k] = 0;
        af = max(af, Tf[i]);
        const uint32_t l2 = c.fr[i];
        int32_t t2 = af + (int32_t)i;
        if (t2 > (int32_t)(ch.end - 1)) t2 = (int32_t)(ch.end - 1);
        c.fm[l2] = (uint32_t)t2 - c.fwd[l2];
        c.fc[i] = 0;  // Tf[i] was last read just above, by this lane
      }
      __syncthreads();
    }
    mark(P, S, 6);

    // ---- barrier states (extrusion_barriers.cpp:145-161) ----
    // both thresholds are loaded ahead of the draw, so their latency hides behind it
    for (uint32_t b = tidx(); b < NB; b += NT) {
      const uint32_t t_in = ch.bar_t_inactive[b], t_act = ch.bar_t_active[b];
      const U4 d = R.draw(b, epoch, kSBarrier);
      const uint8_t st = c.bst[b];
      if (st == 0 && d.x >= t_in && t_in != kThresholdOne) c.bst[b] = 1;
      else if (st == 1 && d.x >= t_act && t_act != kThresholdOne) c.bst[b] = 0;
    }
    // ---- units at the chromosome ends (simulation_detect_collisions.cpp:30-125) ----
    if (tidx() == 0) {
      const uint32_t first_fwd = c.fwd[c.fr[0]];
      const uint32_t last_rev = c.rev[c.rr[n - 1]];
      uint32_t n5 = 0;
      for (uint32_t k = 0; k < n; ++k) {
        const uint32_t l = c.rr[k];
        if (c.rev[l] == ch.start) {
          ++n5;
          c.rc[l] = ck(kCkChrom, 5);
        } else if (c.rev[l] > first_fwd) {
          break;
        } else if (c.rev[l] - c.rm[l] == ch.start) {
          c.rc[l] = ck(kCkChrom, 5);
          ++n5;
          break;
        }
      }
      uint32_t n3 = 0;
      for (uint32_t k = n - 1; k > 0; --k) {
        const uint32_t l = c.fr[k];
        if (c.fwd[l] == ch.end - 1) {
          ++n3;
          c.fc[l] = ck(kCkChrom, 3);
        } else if (c.fwd[l] < last_rev) {
          break;
        } else if (c.fwd[l] + c.fm[l] == ch.end - 1) {
          c.fc[l] = ck(kCkChrom, 3);
          ++n3;
          break;
        }
      }
      S.n5 = n5;
      S.n3 = n3;
    }
    // fwd positions in rank order for the first LEF-BAR pass, filled beside lane 0's walk, with
    // their bucket index once burn-in is over
    uint32_t* sorted = c.scratch;
    const bool use_bk = bdone && n > 0;
    auto fill_sorted = [&](const uint32_t* pos, const uint16_t* ranks) {
      for (uint32_t k = tidx(); k < n; k += NT) {
        const uint32_t pk = pos[ranks[k]];
        sorted[k] = pk;
        if (use_bk) {
          const uint32_t q = bucket(pk);
          const uint32_t q0 = k == 0 ? 0u : bucket(pos[ranks[k - 1]]) + 1;
          for (uint32_t b = q0; b <= q; ++b) T[b] = (uint16_t)k;
          if (k == n - 1)
            for (uint32_t b = q + 1; b <= nbk; ++b) T[b] = (uint16_t)n;
        }
      }
    };
    fill_sorted(c.fwd, c.fr);
    __syncthreads();
    mark(P, S, 7);
    const uint32_t n5 = S.n5, n3 = S.n3;

    // ---- LEF-BAR (simulation_detect_collisions.cpp:128-252): per active barrier, the nearest
    // unit moving towards it; the closest blocking barrier wins ----
    // fwd side first, so the rev positions in rank order stay in `sorted` for the primary step
    for (int it = 0; it < 2; ++it) {
      const int side = 1 - it;
      if (side == 0) {  // the fwd pass ended with a barrier, so `sorted` and T are free
        fill_sorted(c.rev, c.rr);
        __syncthreads();
      }
      const uint32_t rev_lo = n5 ? n5 - 1 : 0;
      const uint32_t fwd_top = n - (n3 ? n3 - 1 : 0);
      // the next barrier's position is loaded while this one is searched
      uint32_t bp_next = tidx() < NB ? ch.bar_pos[tidx()] : 0u;
      for (uint32_t b = tidx(); b < NB; b += NT) {
        const uint32_t bp = bp_next;
        if (b + NT < NB) bp_next = ch.bar_pos[b + NT];
        if (!c.bst[b]) continue;
        if (side == 0) {
          // first rank at or after rev_lo with a position beyond bp; it lies in bp's bucket
          // range [T[q], T[q + 1]]
          uint32_t lo = rev_lo, hi = n;
          if (use_bk) {
            const uint32_t q = bucket(bp);
            lo = max(lo, (uint32_t)T[q]);
            hi = max(lo, (uint32_t)T[q + 1]);
          }
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sorted[mid] <= bp) lo = mid + 1;
            else hi = mid;
          }
          if (lo >= n) continue;
          const uint32_t l = c.rr[lo];
          if (sorted[lo] - bp > c.rm[l]) continue;
          const U4 d = R.draw(b, epoch, kSLefBar);
          if (bern(d.x, ch.bar_dir[b] == 1 ? P.t_pblock_major : P.t_pblock_minor))
            atomicMax(&c.rc[l], ck(kCkBar, b));
        } else {
          // first rank below fwd_top at or beyond bp, searched within bp's bucket range
          uint32_t lo = 0, hi = fwd_top;
          if (use_bk) {
            const uint32_t q = bucket(bp);
            lo = min((uint32_t)T[q], fwd_top);
            hi = min((uint32_t)T[q + 1], fwd_top);
          }
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sorted[mid] < bp) lo = mid + 1;
            else hi = mid;
          }
          if (lo == 0) continue;
          const uint32_t l = c.fr[lo - 1];
          if (bp - sorted[lo - 1] > c.fm[l]) continue;
          const U4 d = R.draw(b, epoch, kSLefBar);
          if (bern(d.y, ch.bar_dir[b] == 2 ? P.t_pblock_major : P.t_pblock_minor))
            atomicMax(&c.fc[l], ck(kCkBarFwdTmp, kCkIdxMask - b));
        }
      }
      __syncthreads();
    }

    // ---- primary LEF-LEF (simulation_detect_collisions.cpp:255-401): a fwd unit that is the
    // last one before a rev unit pairs with that rev unit ----
    // `sorted` holds the rev positions in rank order (the LEF-BAR pass above ended with them);
    // fwd barrier collisions still carry the LEF-BAR encoding until the move corrections rewrite
    // them, so they are read through fc_final
    mark(P, S, 8);
    if (n5 < n && n3 < n) {
      const uint32_t i2 = n - (n3 ? n3 - 1 : 0);
      // contiguous runs of fwd ranks per lane: the first rev unit beyond each fwd unit only moves
      // forward along the run (both orders are sorted by position)
      const uint32_t kmax = i2 > 0 ? i2 - 1 : 0;
      const uint32_t ckr = (kmax + NT - 1) / NT;
      const uint32_t k_begin = tidx() * ckr, k_end = min(kmax, k_begin + ckr);
      uint32_t lo = n5;
      for (uint32_t k = k_begin; k < k_end; ++k) {
        const uint32_t fl = c.fr[k];
        const uint32_t fp = c.fwd[fl];
        if (k == k_begin) {
          uint32_t hi = n;
          if (use_bk) {  // T indexes the rev positions (the last LEF-BAR fill)
            const uint32_t q = bucket(fp);
            lo = max(lo, (uint32_t)T[q]);
            hi = max(lo, (uint32_t)T[q + 1]);
          }
          while (lo < hi) {
            const uint32_t mid = (lo + hi) >> 1;
            if (sorted[mid] <= fp) lo = mid + 1;
            else hi = mid;
          }
        } else {
          while (lo < n && sorted[lo] <= fp) ++lo;
        }
        if (lo >= n) continue;
        const uint32_t rp = sorted[lo];
        if (c.fwd[c.fr[k + 1]] < rp) continue;  // not the last fwd unit before rp
        const uint32_t rl = c.rr[lo];
        const uint32_t delta = rp - fp;
        if (!(delta < c.rm[rl] + c.fm[fl])) continue;
        const U4 d = R.draw(k, epoch, kSPrimary);
        if (!bern(d.x, P.t_bypass_keep)) continue;
        const uint32_t rv = c.rc[rl], fv = fc_final(c.fc[fl]);
        const bool rocc = ck_stalled(rv), focc = ck_stalled(fv);
        // collision site (compute_lef_lef_collision_pos, simulation.cpp:514-542)
        const uint32_t rel = c.rm[rl] + c.fm[fl];
        const uint32_t cpos = fp + (uint32_t)rint((double)c.fm[fl] * ((double)delta / (double)rel));
        const uint32_t cp_rev = cpos == fp ? cpos + 1 : cpos;
        const uint32_t cp_fwd = cpos == fp ? cpos : cpos - 1;
        if (!rocc && !focc) {
          c.rc[rl] = ck(kCkPrimary, fl);
          c.fc[fl] = ck(kCkPrimary, rl);
        } else if (rocc && !focc) {
          if (ck_kind(rv) == kCkBar && cp_fwd > ch.bar_pos[ck_idx(rv)]) c.rc[rl] = ck(kCkPrimary, fl);
          c.fc[fl] = ck(kCkPrimary, rl);
        } else if (!rocc && focc) {
          c.rc[rl] = ck(kCkPrimary, fl);
          if (ck_kind(fv) == kCkBar && cp_rev < ch.bar_pos[ck_idx(fv)]) c.fc[fl] = ck(kCkPrimary, rl);
        }
      }
    }
    __syncthreads();
    mark(P, S, 9);

    // ---- move corrections (simulation_correct_moves.cpp:20-122) ----
    // One pass: LEF i's barrier corrections, then i as the rev unit and as the fwd unit of a
    // primary pair. The reference's three passes read a partner's barrier-corrected move only
    // as a position (fwd + corrected fwd move = barrier - 1, rev - corrected rev move =
    // barrier + 1), taken here from the barrier itself, so no pass waits for another. Primary
    // pairs are one-to-one (one fwd unit per rev unit), so every move is written by one lane.
    for (uint32_t i = tidx(); i < n; i += NT) {
      const uint32_t rci = c.rc[i], fci = fc_final(c.fc[i]);
      c.fc[i] = fci;
      if (ck_kind(rci) == kCkBar) c.rm[i] = (c.rev[i] - ch.bar_pos[ck_idx(rci)]) - 1;
      if (ck_kind(fci) == kCkBar) c.fm[i] = (ch.bar_pos[ck_idx(fci)] - c.fwd[i]) - 1;
      if (ck_kind(rci) == kCkPrimary) {
        const uint32_t f = ck_idx(rci);
        const uint32_t fcf = fc_final(c.fc[f]);
        if (ck_kind(fcf) == kCkPrimary) {
          const uint32_t rp = c.rev[i], fp = c.fwd[f];
          const uint32_t rel = c.rm[i] + c.fm[f];
          const uint32_t cpos = fp + (uint32_t)rint((double)c.fm[f] * ((double)(rp - fp) / (double)rel));
          c.rm[i] = rp - (cpos == fp ? cpos + 1 : cpos);
          c.fm[f] = (cpos == fp ? cpos : cpos - 1) - fp;
        } else if (ck_kind(fcf) == kCkBar) {
          c.rm[i] = c.rev[i] - (ch.bar_pos[ck_idx(fcf)] - 1) - 1;
        }
      }
      if (ck_kind(fci) == kCkPrimary) {
        const uint32_t rcr = c.rc[ck_idx(fci)];
        if (ck_kind(rcr) == kCkBar) c.fm[i] = (ch.bar_pos[ck_idx(rcr)] + 1) - c.fwd[i] - 1;
      }
    }
    __syncthreads();
    mark(P, S, 10);

    // ---- secondary LEF-LEF (simulation_detect_collisions.cpp:404-518) ----
    // Both sides' passes first (neither reads what the other writes), then the fix-ups, rev
    // first as in process_collisions (simulation.cpp:925-955): the rev fix-ups read fwd
    // positions and the fwd fix-ups rev positions, as of the fix-ups before them.
    {
      uint16_t* F = reinterpret_cast<uint16_t*>(c.scratch);
      uint32_t na[2];
      secondary_both<NT>(P, R, epoch, c, S, n, max(1u, n5), (n3 ? n3 - 1 : 0) + 1, F, na);
      mark(P, S, 11);
      if (na[0]) {
        fix_avoided<NT>(c, ch, n, F, na[0], true, P.serial_fixups);
        __syncthreads();
      }
      if (na[1]) {
        fix_avoided<NT>(c, ch, n, F + n, na[1], false, P.serial_fixups);
        __syncthreads();
      }
      mark(P, S, 13);
    }

    if constexpr (kOut) {
      if (P.log.rec) log_state<NT>(P, ch, c, S, n, epoch, bdone, item);
    }
    // ---- extrude and release (simulation.cpp:489-512, 544-592) ----
    {
      const double base = bdone ? ch.prob_release : ch.prob_release_burnin;
      // the three release thresholds (no, one soft, two or one hard barrier stall), once
      const uint64_t t_rel[3] = {thresh(1.0 * base), thresh(P.inv_soft * base),
                                 thresh(P.inv_hard * base)};
      for (uint32_t i = tidx(); i < n; i += NT) {
        c.rev[i] -= c.rm[i];
        c.fwd[i] += c.fm[i];
        const int h = (ck_kind(c.rc[i]) == kCkBar && ch.bar_dir[ck_idx(c.rc[i])] == 1) +
                      (ck_kind(c.fc[i]) == kCkBar && ch.bar_dir[ck_idx(c.fc[i])] == 2);
        const U4 d = R.draw(i, epoch, kSRelease);
        // a released LEF binds again at once, with the draw the next epoch's bind would make
        // (simulate_one_cell rebinds every released LEF at the start of the next epoch)
        uint32_t rebound = 0;
        if (bern(d.x, h == 0 ? t_rel[0] : (h == 1 ? t_rel[1] : t_rel[2]))) {
          const U4 e = R.draw(i, epoch + 1, kSBind);
          c.rev[i] = c.fwd[i] = ch.start + ubounded(e.x, span);
          rebound = 1;
        }
        c.rm[i] = rebound;  // the next epoch's "bound this epoch" flag
      }
    }
    __syncthreads();
    mark(P, S, 14);
  }
  if (tidx() == 0 && stats) {
    stats[0] = epoch;
    stats[1] = S.burnin_epochs;
    stats[2] = contacts;
    stats[3] = 0;
  }
  if (P.timers && tidx() == 0)
    for (int k = 0; k < kFastPhases; ++k) atomicAdd(&g_fast_phase[k], (unsigned long long)S.phase[k]);
}

// Persistent workers: each workgroup takes cells off the queue until it is empty. kStore selects
// where the working set lives (LDS, a per-workgroup slab in global memory, or the hybrid of the
// two); it is a template argument so arrays in LDS are addressed with LDS instructions. Every size is compiled
// for four waves per SIMD (at most 128 VGPRs), so 16 waves share a CU at every workgroup size
// whose cells fit 1024 / NT to a CU: the epoch loop waits on LDS and atomics, and the waves
// of other workgroups hide that latency.
template <int NT, int kStore, bool kOut>
__global__ void __launch_bounds__(NT) __attribute__((amdgpu_waves_per_eu(4)))
    fast_cells_kernel(FastParams P, const FastChrom* chroms, const WorkItem* items,
                      const uint32_t* order, uint64_t nitems, unsigned long long* next,
                      uint8_t* slab, uint64_t slab_stride, uint64_t* cell_stats) {
  extern __shared__ __align__(16) uint8_t lds[];
  __shared__ Shared S;
  for (;;) {
    if (tidx() == 0) S.item = atomicAdd(next, 1ULL);
    __syncthreads();
    // the same in every lane: in scalar registers, so the work item and the chromosome
    // descriptor are scalar loads and stay in SGPRs through the cell
    const uint64_t q = uniform64(S.item);
    __syncthreads();
    if (q >= nitems) return;
    const uint64_t it = order[q];
    const WorkItem w = items[it];
    const FastChrom ch = chroms[w.chrom];
    Cell c;
    if constexpr (kStore == kStoreLds) {
      carve(c, lds, ch.nlefs, ch.nbar, P.hist_len);
    } else if constexpr (kStore == kStoreSlab) {
      carve(c, slab + (uint64_t)blockIdx.x * slab_stride, ch.nlefs, ch.nbar, P.hist_len);
    } else {
      carve_hybrid(c, lds, slab + (uint64_t)blockIdx.x * slab_stride, ch.nlefs, ch.nbar,
                   P.hist_len);
    }
    fast_cell<NT, kOut>(P, ch, c, S, w.seed, w.target,
                        cell_stats ? cell_stats + 4 * it : nullptr, (uint32_t)it);
    __syncthreads();
  }
}

// Clears n 32-bit words. Runs and captured graphs clear bands and counters with this kernel
// rather than with memset nodes.
__global__ void zero_words_kernel(uint32_t* p, uint64_t n) {
  for (uint64_t e = (uint64_t)blockIdx.x * blockDim.x + tidx(); e < n;
       e += (uint64_t)gridDim.x * blockDim.x)
    p[e] = 0u;
}

// fast_math.h's functions on n inputs: fn 0 ln, 1 expm1, 2 sin and cos of 2 pi x (out holds 2n:
// sines then cosines).
__global__ void fast_math_kernel(int fn, const double* x, double* out, int n) {
  const int i = blockIdx.x * blockDim.x + threadIdx.x;
  if (i >= n) return;
  if (fn == 0) {
    out[i] = ln_d(x[i]);
  } else if (fn == 1) {
    out[i] = expm1_d(x[i]);
  } else {
    double sn, cs;
    sincos_2pi_d(x[i], &sn, &cs);
    out[i] = sn;
    out[n + i] = cs;
  }
}

// Fast mode's band increments for (rows[i], cols[i]), in order, on one lane.
__global__ void fast_band_kernel(FastChrom ch, const uint32_t* rows, const uint32_t* cols,
                                 uint64_t n) {
  if (threadIdx.x != 0 || blockIdx.x != 0) return;
  for (uint64_t i = 0; i < n; ++i) band_add(ch, rows[i], cols[i]);
}

__global__ void philox_kat_kernel(const uint32_t* in, uint32_t* out, int n) {
  const int i = blockIdx.x * blockDim.x + tidx();
  if (i >= n) return;
  const U4 r = philox(U4{in[6 * i], in[6 * i + 1], in[6 * i + 2], in[6 * i + 3]}, in[6 * i + 4],
                      in[6 * i + 5]);
  out[4 * i] = r.x;
  out[4 * i + 1] = r.y;
  out[4 * i + 2] = r.z;
  out[4 * i + 3] = r.w;
}

thread_local std::string g_fast_err;

// 32-bit Bernoulli threshold of probability p: an event of probability p happens when a 32-bit
// draw is below it. kThresholdOne stands for p = 1 (every p >= 1 - 2^-32), which no draw may
// fail, so a barrier whose self-transition probability is 1 (a barrier a perturbation deletes:
// stp_active 0, stp_inactive 1) never changes state, as in the reference.
uint32_t t32(double p) {
  if (p <= 0.0) return 0;
  const double v = p * 4294967296.0;
  return v >= 4294967295.0 ? 0xFFFFFFFFu : (uint32_t)v;
}

}  // namespace

// One launch of the persistent kernel over a contiguous run of the work order.
struct Tier {
  uint32_t threads = 0;    // workgroup size
  uint32_t per_cu = 1;     // workgroups one CU holds by LDS (2 for device-memory working sets)
  uint64_t shmem = 0;      // dynamic LDS per workgroup; 0: working sets in the slab
  int store = kStoreLds;   // kStoreLds, kStoreSlab or kStoreHybrid
  uint64_t begin = 0, count = 0;  // cells order[begin, begin + count)
};

struct FastImpl {
  FastParams P;
  std::vector<FastChrom> chroms;
  FastChrom* d_chroms = nullptr;
  uint32_t* d_order = nullptr;
  unsigned long long* d_next = nullptr;  // work-queue counters, one per size tier
  // the size tiers' launches run on streams of their own, forked from and joined back into the
  // caller's stream, so a tier's cells start on CUs that the tier before has finished with
  hipStream_t side[5] = {};
  hipEvent_t fork = nullptr, join[5] = {};
  uint8_t* d_slab = nullptr;
  uint64_t slab_stride = 0;
  uint32_t num_cus = 1;
  uint32_t threads = 0;          // 0: one launch per size tier; else one launch at this size
  bool global = false;           // every working set in the slab, one launch
  bool fits_lds = true;          // every cell's working set fits in LDS
  uint64_t max_lefs = 1;
  uint64_t max_bytes = 0;        // largest working set of one cell
  uint64_t max_grid = 0;         // workgroups the slab has room for
  std::vector<Tier> tiers;       // size tiers, in launch order (their runs tile the order)
  std::vector<uint32_t> order_host;  // the work order (a copy of d_order)
  bool occupancy = false;        // chromosomes keep a 1D LEF occupancy histogram
  void* d_slab_own = nullptr;    // slab allocated when a job that fits LDS is moved to memory
  uint64_t nitems = 0;
  void* d_mem = nullptr;
};

constexpr uint64_t kLdsTotal = 160 * 1024;                          // LDS of one CU
constexpr uint64_t kLdsReserve = sizeof(Shared) + 512;               // static LDS and slack
constexpr uint64_t kLdsLimit = kLdsTotal - kLdsReserve;              // largest working set in LDS

// Workgroups of one persistent launch: as many as the CU holds (LDS, 2048 lanes, at most 16 per
// CU), never more than there are cells.
uint32_t launch_grid(const FastImpl* impl, uint32_t threads, uint32_t per_cu, uint64_t cells) {
  const uint64_t k = std::min<uint64_t>({per_cu, 2048u / threads, 16u});
  return (uint32_t)std::min<uint64_t>(cells, std::max<uint64_t>(1, k) * impl->num_cus);
}

// The launch a job makes when it is not split into tiers: the whole order at one workgroup
// size, from LDS when every cell fits there (sized for the largest), else from the slab.
Tier single_launch(const FastImpl* impl) {
  Tier t;
  const bool lds = impl->fits_lds && !impl->global;
  t.threads = impl->threads ? impl->threads : kFastDefaultThreads;
  t.shmem = lds ? impl->max_bytes : 0;
  t.store = lds ? kStoreLds : kStoreSlab;
  t.per_cu = lds ? (uint32_t)std::max<uint64_t>(1, kLdsTotal / (impl->max_bytes + kLdsReserve)) : 2u;
  t.begin = 0;
  t.count = impl->nitems;
  return t;
}

bool tiered(const FastImpl* impl) { return impl->threads == 0 && !impl->global; }

template <int NT, bool kOut>
hipError_t launch_fast_out(hipStream_t stream, FastImpl* impl, const WorkItem* d_items,
                           uint64_t* d_cell_stats, uint32_t grid, uint64_t shmem, int store,
                           const uint32_t* order, uint64_t nitems, unsigned long long* next,
                           uint8_t* slab) {
  if (store == kStoreLds)
    hipLaunchKernelGGL((fast_cells_kernel<NT, kStoreLds, kOut>), dim3(grid), dim3(NT), shmem,
                       stream, impl->P, impl->d_chroms, d_items, order, nitems, next,
                       nullptr, (uint64_t)0, d_cell_stats);
  else if (store == kStoreHybrid)
    hipLaunchKernelGGL((fast_cells_kernel<NT, kStoreHybrid, kOut>), dim3(grid), dim3(NT), shmem,
                       stream, impl->P, impl->d_chroms, d_items, order, nitems, next,
                       slab, impl->slab_stride, d_cell_stats);
  else
    hipLaunchKernelGGL((fast_cells_kernel<NT, kStoreSlab, kOut>), dim3(grid), dim3(NT), 0, stream,
                       impl->P, impl->d_chroms, d_items, order, nitems, next,
                       slab, impl->slab_stride, d_cell_stats);
  return hipGetLastError();
}

template <int NT>
hipError_t launch_fast(hipStream_t stream, FastImpl* impl, const WorkItem* d_items,
                       uint64_t* d_cell_stats, uint32_t grid, uint64_t shmem, int store,
                       const uint32_t* order, uint64_t nitems, unsigned long long* next,
                       uint8_t* slab) {
  if (impl->occupancy || impl->P.log.rec)
    return launch_fast_out<NT, true>(stream, impl, d_items, d_cell_stats, grid, shmem, store,
                                     order, nitems, next, slab);
  return launch_fast_out<NT, false>(stream, impl, d_items, d_cell_stats, grid, shmem, store,
                                    order, nitems, next, slab);
}

// Launches the persistent kernel with tier t's shape over order[0, nitems) on work counter
// `slot` (cleared first, on the same stream, so launches in a row reuse it, unless `cleared` says
// the caller has cleared it already; launches that run at once use different slots) and, for
// working sets in device memory, the slab's workgroup slots from slab_slot on (launches that run
// at once use disjoint slots).
hipError_t launch_tier(hipStream_t stream, FastImpl* impl, const Tier& t, const WorkItem* d_items,
                       uint64_t* d_cell_stats, const uint32_t* order, uint64_t nitems,
                       uint32_t slot = 0, uint64_t slab_slot = 0, bool cleared = false) {
  if (nitems == 0) return hipSuccess;
  unsigned long long* next = impl->d_next + slot;
  uint8_t* slab = impl->d_slab ? impl->d_slab + slab_slot * impl->slab_stride : nullptr;
  hipError_t e = cleared ? hipSuccess : launch_zero_words(stream, next, 8);
  if (e != hipSuccess) return e;
  const uint32_t grid = launch_grid(impl, t.threads, t.per_cu, nitems);
  switch (t.threads) {
    case 64: return launch_fast<64>(stream, impl, d_items, d_cell_stats, grid, t.shmem, t.store, order, nitems, next, slab);
    case 128: return launch_fast<128>(stream, impl, d_items, d_cell_stats, grid, t.shmem, t.store, order, nitems, next, slab);
    case 256: return launch_fast<256>(stream, impl, d_items, d_cell_stats, grid, t.shmem, t.store, order, nitems, next, slab);
    case 512: return launch_fast<512>(stream, impl, d_items, d_cell_stats, grid, t.shmem, t.store, order, nitems, next, slab);
    case 1024: return launch_fast<1024>(stream, impl, d_items, d_cell_stats, grid, t.shmem, t.store, order, nitems, next, slab);
    default: return hipErrorInvalidValue;
  }
}

// The dynamic-LDS limit of every LDS instantiation, set to the most a workgroup may take, so no
// launch changes kernel attributes (launches may be recorded into a graph).
template <int NT>
hipError_t set_lds_limit() {
  const void* fns[] = {reinterpret_cast<const void*>(fast_cells_kernel<NT, kStoreLds, false>),
                       reinterpret_cast<const void*>(fast_cells_kernel<NT, kStoreLds, true>),
                       reinterpret_cast<const void*>(fast_cells_kernel<NT, kStoreHybrid, false>),
                       reinterpret_cast<const void*>(fast_cells_kernel<NT, kStoreHybrid, true>)};
  for (const void* f : fns) {
    const hipError_t e =
        hipFuncSetAttribute(f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)kLdsLimit);
    if (e != hipSuccess) return e;
  }
  return hipSuccess;
}

hipError_t set_lds_limits() {
  hipError_t e = set_lds_limit<64>();
  if (e == hipSuccess) e = set_lds_limit<128>();
  if (e == hipSuccess) e = set_lds_limit<256>();
  if (e == hipSuccess) e = set_lds_limit<512>();
  if (e == hipSuccess) e = set_lds_limit<1024>();
  return e;
}

// Size tiers: a cell whose working set leaves room for k = 1, 2 or 4 of them in one CU's LDS
// runs in the tier of that k, at 1024 / k lanes, so every tier keeps 16 waves on each CU. A cell
// too large for LDS whose positions, ranks and barrier states fit there runs hybrid (those in
// LDS, the per-epoch temporaries in the slab; one workgroup of 1024 lanes per CU); larger cells
// work wholly from the slab (two workgroups of kFastDefaultThreads per CU). Tier of chromosome
// c: 0 slab, 1 k = 1, 2 k = 2, 3 k = 4, 4 hybrid.
constexpr uint32_t kTiers = 5;

uint32_t tier_of(uint64_t bytes, uint64_t hybrid_bytes) {
  if (bytes > kLdsLimit) return hybrid_bytes <= kLdsLimit ? 4 : 0;
  const uint64_t k = kLdsTotal / (bytes + kLdsReserve);
  return k >= 4 ? 3 : (k >= 2 ? 2 : 1);
}

Tier tier_shape(uint32_t tier) {
  Tier t;
  if (tier == 0) {
    t.threads = kFastDefaultThreads;
    t.per_cu = 2;
    t.store = kStoreSlab;
  } else if (tier == 4) {
    t.threads = 1024;
    t.per_cu = 1;
    t.store = kStoreHybrid;
  } else {
    t.per_cu = 1u << (tier - 1);
    t.threads = 1024 / t.per_cu;
    t.store = kStoreLds;
  }
  return t;
}

uint64_t fast_cell_bytes_host(uint32_t n, uint32_t nbar, uint32_t hist) {
  return fast_cell_bytes(n, nbar, hist);
}

namespace {

// The job-wide kernel parameters of a parameter block; false (with g_fast_err set) when fast mode
// cannot take them.
bool fill_fast_params(const mgpu_params& p, FastParams& P) {
  P = FastParams{};
  P.bin_size = (uint32_t)p.bin_size;
  P.rev_speed = (double)p.rev_speed;
  P.fwd_speed = (double)p.fwd_speed;
  P.rev_speed_burnin = (double)p.rev_speed_burnin;
  P.fwd_speed_burnin = (double)p.fwd_speed_burnin;
  P.rev_std = p.rev_speed_std;
  P.fwd_std = p.fwd_speed_std;
  P.t_bypass_keep = thresh(1.0 - p.p_bypass);
  P.t_pblock_major = thresh(p.pblock_major);
  P.t_pblock_minor = thresh(p.pblock_minor);
  P.inv_soft = 1.0 / p.soft_stall_mult;
  P.inv_hard = 1.0 / p.hard_stall_mult;
  P.t_loop = std::isinf(p.tad_to_loop_ratio) ? 0 : thresh(1.0 / (p.tad_to_loop_ratio + 1.0));
  P.gev_mu = p.gev_mu;
  P.gev_sigma = p.gev_sigma;
  P.gev_xi = p.gev_xi;
  P.noisify = p.noisify;
  P.skip_burnin = p.skip_burnin;
  P.hist_len = (uint32_t)p.burnin_history_length;
  P.window = (uint32_t)p.burnin_smoothing_window;
  P.min_burnin_epochs = p.min_burnin_epochs;
  P.max_burnin_epochs = p.max_burnin_epochs;
  P.target_contact_density = p.target_contact_density;
  if (P.hist_len > 4096) {
    g_fast_err = "fast mode supports burn-in histories of up to 4096 epochs";
    return false;
  }
  return true;
}

// The device view of one chromosome whose barrier tables (positions and the three thresholds,
// 4 x nbar u32) start at bars; band, counters, direction table and occupancy from d.
FastChrom make_fast_chrom(const FastParams& P, const DetChrom& d, uint32_t* bars,
                          double release, double release_burnin) {
  FastChrom f{};
  f.start = (uint32_t)d.start;
  f.end = (uint32_t)d.end;
  f.nlefs = (uint32_t)d.nlefs;
  f.nbar = (uint32_t)d.nbar;
  f.bar_pos = bars;
  f.bar_t_active = bars + d.nbar;
  f.bar_t_inactive = bars + 2 * d.nbar;
  f.bar_t_occ = bars + 3 * d.nbar;
  f.bar_dir = d.bar_dir;
  f.band = d.band;
  f.missed = d.missed;
  f.dropped = d.dropped;
  f.nrows = (uint32_t)d.nrows;
  f.cpe = (uint32_t)d.cpe;
  f.target_epochs = d.target_epochs;
  f.bind_rate = d.bind_rate;
  f.bind_exp_neg = std::exp(-d.bind_rate);
  f.bind_sd = std::sqrt(d.bind_rate);
  f.prob_release = release;
  f.prob_release_burnin = release_burnin;
  f.lds_bytes = fast_cell_bytes(f.nlefs, f.nbar, P.hist_len);
  f.occupancy = d.occupancy;
  return f;
}

// The 32-bit barrier tables of one chromosome (positions, then the thresholds of the active and
// inactive self-transitions and of the stationary occupancy) into hp[0, 4 * n).
void fill_barrier_tables(uint32_t* hp, uint64_t n, const uint64_t* pos, const double* sa,
                         const double* si) {
  for (uint64_t i = 0; i < n; ++i) {
    hp[i] = (uint32_t)pos[i];
    hp[n + i] = t32(sa[i]);
    hp[2 * n + i] = t32(si[i]);
    const double occ =
        (sa[i] + si[i] == 0) ? 0.0 : (1.0 - si[i]) / ((1.0 - si[i]) + (1.0 - sa[i]));
    hp[3 * n + i] = t32(occ < 0.0 ? 0.0 : (occ > 1.0 ? 1.0 : occ));
  }
}

}  // namespace

int fast_job_create(const mgpu_params& p, const std::vector<DetChrom>& dchroms,
                    const mgpu_barriers* host_bars, const std::vector<double>& release,
                    const std::vector<WorkItem>& items, FastJob* job) {
  auto* impl = new FastImpl();
  FastParams& P = impl->P;
  if (!fill_fast_params(p, P)) {
    delete impl;
    return -1;
  }

  const uint64_t nch = dchroms.size();
  uint64_t max_bytes = 0, nbar_total = 0, max_lefs = 1;
  for (const auto& d : dchroms) {
    max_lefs = std::max<uint64_t>(max_lefs, d.nlefs);
    if (d.nlefs > 65535) {
      g_fast_err = "fast mode supports up to 65535 LEFs per chromosome";
      delete impl;
      return -1;
    }
    nbar_total += d.nbar;
    max_bytes = std::max<uint64_t>(max_bytes, fast_cell_bytes((uint32_t)d.nlefs, (uint32_t)d.nbar,
                                                              P.hist_len));
  }
  int dev = 0;
  hipDeviceProp_t prop;
  hipError_t e = hipGetDevice(&dev);
  if (e == hipSuccess) e = hipGetDeviceProperties(&prop, dev);
  if (e == hipSuccess) e = set_lds_limits();
  if (e != hipSuccess) {
    g_fast_err = std::string("fast-mode set-up failed: ") + hipGetErrorString(e);
    delete impl;
    return -1;
  }
  impl->fits_lds = max_bytes <= kLdsLimit;
  impl->max_bytes = max_bytes;
  impl->num_cus = (uint32_t)prop.multiProcessorCount;
  impl->nitems = items.size();
  impl->max_lefs = max_lefs;
  impl->threads = 0;
  // the slab holds one working set (sized for the largest cell) per workgroup of the largest
  // grid any launch makes; it exists from the start when some cell does not fit in LDS
  impl->slab_stride = (max_bytes + 255) & ~255ULL;
  const uint64_t max_grid = std::min<uint64_t>(items.size(), 16ull * impl->num_cus);
  impl->max_grid = max_grid;

  const uint64_t bar_bytes = (4 * nbar_total + 64) * 4;
  const uint64_t chrom_bytes = (nch * sizeof(FastChrom) + 255) & ~255ULL;
  const uint64_t order_bytes = (items.size() * 4 + 255) & ~255ULL;
  const uint64_t slab_bytes = impl->fits_lds ? 0 : impl->slab_stride * max_grid;
  e = hipMalloc(&impl->d_mem, bar_bytes + chrom_bytes + order_bytes + 256 + slab_bytes);
  if (e != hipSuccess) {
    g_fast_err = std::string("fast-mode allocation failed: ") + hipGetErrorString(e);
    delete impl;
    return -1;
  }
  uint8_t* base = static_cast<uint8_t*>(impl->d_mem);
  uint32_t* bars = reinterpret_cast<uint32_t*>(base);
  impl->d_chroms = reinterpret_cast<FastChrom*>(base + bar_bytes);
  impl->d_order = reinterpret_cast<uint32_t*>(base + bar_bytes + chrom_bytes);
  impl->d_next = reinterpret_cast<unsigned long long*>(base + bar_bytes + chrom_bytes + order_bytes);
  impl->d_slab = impl->fits_lds ? nullptr : base + bar_bytes + chrom_bytes + order_bytes + 256;

  // 32-bit barrier tables: positions and the transition / occupancy thresholds
  std::vector<uint32_t> hb(4 * nbar_total + 64, 0);
  uint64_t off = 0;
  for (uint64_t c = 0; c < nch; ++c) {
    const DetChrom& d = dchroms[c];
    fill_barrier_tables(hb.data() + off, d.nbar, host_bars[c].pos, host_bars[c].stp_active,
                        host_bars[c].stp_inactive);
    const FastChrom f = make_fast_chrom(P, d, bars + off, release[2 * c], release[2 * c + 1]);
    impl->occupancy |= d.occupancy != nullptr;
    impl->chroms.push_back(f);
    off += 4 * d.nbar;
  }
  e = hipMemcpy(bars, hb.data(), hb.size() * 4, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = hipMemcpy(impl->d_chroms, impl->chroms.data(), nch * sizeof(FastChrom),
                  hipMemcpyHostToDevice);
  // work order: by size tier, longest cells first within a tier (chromosomes with more LEFs
  // take longer per epoch)
  std::vector<uint32_t> tier(nch);
  for (uint64_t c = 0; c < nch; ++c)
    tier[c] = tier_of(impl->chroms[c].lds_bytes,
                      hybrid_lds_bytes(impl->chroms[c].nlefs, impl->chroms[c].nbar, P.hist_len));
  std::vector<uint32_t> order(items.size());
  for (uint32_t i = 0; i < order.size(); ++i) order[i] = i;
  std::stable_sort(order.begin(), order.end(), [&](uint32_t a, uint32_t b) {
    const uint32_t ca = items[a].chrom, cb = items[b].chrom;
    if (tier[ca] != tier[cb]) return tier[ca] < tier[cb];
    return dchroms[ca].nlefs > dchroms[cb].nlefs;
  });
  for (uint32_t k = 0; k < kTiers; ++k) {
    Tier t = tier_shape(k);
    uint64_t bytes = 0;
    for (uint64_t i = 0; i < order.size(); ++i) {
      const uint32_t c = items[order[i]].chrom;
      if (tier[c] != k) continue;
      if (t.count++ == 0) t.begin = i;
      const FastChrom& f = impl->chroms[c];
      bytes = std::max<uint64_t>(bytes, k == 4 ? hybrid_lds_bytes(f.nlefs, f.nbar, P.hist_len)
                                               : f.lds_bytes);
    }
    t.shmem = k == 0 ? 0 : bytes;
    if (t.count) impl->tiers.push_back(t);
  }
  // The tiers' launches run on streams of their own, but the runtime spreads streams over only
  // a few hardware queues (four by default), so a third concurrent launch shares a queue and
  // waits for the launch before it there. With all three LDS tiers present, the four-per-CU
  // cells therefore join the two-per-CU launch (they fit its LDS allocation and run at its
  // width, 512 lanes, which times the same as 256 for them): two launches, each on its own queue.
  const auto at = [&](int store, uint32_t per_cu) {
    for (size_t i = 0; i < impl->tiers.size(); ++i)
      if (impl->tiers[i].store == store && impl->tiers[i].per_cu == per_cu) return (int)i;
    return -1;
  };
  const int t1 = at(kStoreLds, 1), t2 = at(kStoreLds, 2), t4 = at(kStoreLds, 4);
  if (t1 >= 0 && t2 >= 0 && t4 == t2 + 1 &&
      impl->tiers[t2].begin + impl->tiers[t2].count == impl->tiers[t4].begin) {
    impl->tiers[t2].count += impl->tiers[t4].count;
    impl->tiers.erase(impl->tiers.begin() + t4);
  }
  if (e == hipSuccess)
    e = hipMemcpy(impl->d_order, order.data(), order.size() * 4, hipMemcpyHostToDevice);
  impl->order_host = order;
  if (e != hipSuccess) {
    g_fast_err = std::string("fast-mode upload failed: ") + hipGetErrorString(e);
    (void)hipFree(impl->d_mem);
    delete impl;
    return -1;
  }
  impl->nitems = items.size();
  job->impl = impl;
  job->nitems = items.size();
  job->nchroms = nch;
  return 0;
}

hipError_t fast_job_run(hipStream_t stream, FastJob* job, const WorkItem* d_items,
                        uint64_t* d_cell_stats, int timers, hipEvent_t* ev, uint32_t nev) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl) return hipErrorInvalidValue;
  impl->P.timers = timers ? 1u : 0u;
  const std::vector<Tier> runs =
      tiered(impl) ? impl->tiers : std::vector<Tier>{single_launch(impl)};
  const bool mark = ev != nullptr && nev >= 2 * runs.size();
  if (runs.size() == 1) {
    const Tier& t = runs[0];
    hipError_t e = mark ? hipEventRecord(ev[0], stream) : hipSuccess;
    if (e == hipSuccess)
      e = launch_tier(stream, impl, t, d_items, d_cell_stats, impl->d_order + t.begin, t.count);
    if (e == hipSuccess && mark) e = hipEventRecord(ev[1], stream);
    return e;
  }
  // several tiers: each on its own stream, in the tier order (the hardware dispatches their
  // workgroups in launch order, so the largest cells still start first), each stream waiting for
  // the work before the run and the caller's stream for all of them. Tiers whose working sets
  // use the slab take disjoint workgroup slots of it.
  std::vector<uint64_t> slab_slot(runs.size(), 0);
  uint64_t slots = 0;
  for (size_t k = 0; k < runs.size(); ++k) {
    if (runs[k].store == kStoreLds) continue;
    slab_slot[k] = slots;
    slots += launch_grid(impl, runs[k].threads, runs[k].per_cu, runs[k].count);
  }
  if (slots > impl->max_grid) return hipErrorInvalidValue;  // sized for every tier at once
  hipError_t e = hipSuccess;
  if (!impl->fork) {
    e = hipEventCreateWithFlags(&impl->fork, hipEventDisableTiming);
    for (size_t k = 0; k < 5 && e == hipSuccess; ++k) {
      e = hipStreamCreateWithFlags(&impl->side[k], hipStreamNonBlocking);
      if (e == hipSuccess) e = hipEventCreateWithFlags(&impl->join[k], hipEventDisableTiming);
    }
    if (e != hipSuccess) return e;
  }
  // every tier's work counter is cleared before the fork, so each side stream holds its launch
  // alone and the events around it time the launch itself
  e = launch_zero_words(stream, impl->d_next, 8 * runs.size());
  if (e == hipSuccess) e = hipEventRecord(impl->fork, stream);
  for (size_t k = 0; k < runs.size() && e == hipSuccess; ++k) {
    const Tier& t = runs[k];
    e = hipStreamWaitEvent(impl->side[k], impl->fork, 0);
    if (e == hipSuccess && mark) e = hipEventRecord(ev[2 * k], impl->side[k]);
    if (e == hipSuccess)
      e = launch_tier(impl->side[k], impl, t, d_items, d_cell_stats, impl->d_order + t.begin,
                      t.count, (uint32_t)k, slab_slot[k], true);
    if (e == hipSuccess && mark) e = hipEventRecord(ev[2 * k + 1], impl->side[k]);
    if (e == hipSuccess) e = hipEventRecord(impl->join[k], impl->side[k]);
    if (e == hipSuccess) e = hipStreamWaitEvent(stream, impl->join[k], 0);
  }
  return e;
}

namespace {
// Milliseconds of one launch of shape t over order[0, n).
hipError_t time_launch(hipStream_t stream, FastImpl* impl, const Tier& t, const WorkItem* d_items,
                       uint64_t* d_cell_stats, const uint32_t* order, uint64_t n, hipEvent_t a,
                       hipEvent_t b, float* ms) {
  hipError_t e = hipEventRecord(a, stream);
  if (e == hipSuccess) e = launch_tier(stream, impl, t, d_items, d_cell_stats, order, n);
  if (e == hipSuccess) e = hipEventRecord(b, stream);
  if (e == hipSuccess) e = hipEventSynchronize(b);
  if (e == hipSuccess) e = hipEventElapsedTime(ms, a, b);
  return e;
}
}  // namespace

int fast_job_autotune(hipStream_t stream, FastJob* job, const WorkItem* d_items,
                      uint64_t* d_cell_stats, uint64_t sample, float* ms_out) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl) return -1;
  if (sample == 0 || sample > impl->nitems) sample = impl->nitems;
  // the runs timed: the tiers (or the single launch of a job kept in device memory), each on
  // every k-th cell of its part of the order, so a sample spans the chromosome sizes
  std::vector<Tier> runs = tiered(impl) ? impl->tiers : std::vector<Tier>{single_launch(impl)};
  std::vector<uint32_t> all(impl->nitems), pick;
  std::vector<uint64_t> at, cnt;
  hipError_t e = hipMemcpy(all.data(), impl->d_order, all.size() * 4, hipMemcpyDeviceToHost);
  for (const Tier& t : runs) {
    // at least as many cells as a launch of the tier keeps resident on the chip (4 per CU at
    // most), so every size is timed at the occupancy the full run has
    const uint64_t fill = 4ull * impl->num_cus;
    const uint64_t s = std::min<uint64_t>(
        t.count, std::max<uint64_t>(fill, std::max<uint64_t>(1, sample * t.count / impl->nitems)));
    const uint64_t stride = t.count / s;
    at.push_back(pick.size());
    cnt.push_back(s);
    for (uint64_t k = 0; k < s; ++k) pick.push_back(all[t.begin + k * stride]);
  }
  uint32_t* d_pick = nullptr;
  if (e == hipSuccess) e = hipMalloc(&d_pick, pick.size() * 4);
  if (e == hipSuccess) e = hipMemcpy(d_pick, pick.data(), pick.size() * 4, hipMemcpyHostToDevice);
  hipEvent_t a = nullptr, b = nullptr;
  if (e == hipSuccess) e = hipEventCreate(&a);
  if (e == hipSuccess) e = hipEventCreate(&b);
  const uint32_t cand[] = {64, 128, 256, 512, 1024};
  std::vector<float> best_ms(runs.size(), 0.0f);
  std::vector<uint32_t> best(runs.size(), 0);
  impl->P.timers = 0;
  const StateLog log = impl->P.log;
  impl->P.log = StateLog{};  // timing runs leave the state log alone
  for (int k = 0; k < 5 && e == hipSuccess; ++k) {
    float total = 0.0f;
    for (size_t r = 0; r < runs.size() && e == hipSuccess; ++r) {
      Tier t = runs[r];
      t.threads = cand[k];
      float ms = 0.0f;
      e = time_launch(stream, impl, t, d_items, d_cell_stats, d_pick + at[r], cnt[r], a, b, &ms);
      total += ms;
      if (e == hipSuccess && (best[r] == 0 || ms < best_ms[r])) {
        best_ms[r] = ms;
        best[r] = cand[k];
      }
    }
    if (ms_out) ms_out[k] = total;
  }
  impl->P.log = log;
  if (e == hipSuccess) {
    if (tiered(impl)) {
      for (size_t r = 0; r < runs.size(); ++r) impl->tiers[r].threads = best[r];
    } else {
      impl->threads = best[0];
    }
  }
  if (a) (void)hipEventDestroy(a);
  if (b) (void)hipEventDestroy(b);
  if (d_pick) (void)hipFree(d_pick);
  if (e != hipSuccess) {
    g_fast_err = std::string("autotune failed: ") + hipGetErrorString(e);
    return -1;
  }
  return 0;
}

void fast_job_set_log(FastJob* job, const StateLog& log) {
  if (auto* impl = static_cast<FastImpl*>(job->impl)) impl->P.log = log;
}

void fast_job_set_serial_fixups(FastJob* job, int serial) {
  if (auto* impl = static_cast<FastImpl*>(job->impl)) impl->P.serial_fixups = serial ? 1u : 0u;
}

hipError_t fast_job_prepare(FastJob* job) {
  if (!job->impl) return hipErrorInvalidValue;
  return set_lds_limits();
}

int fast_job_set_storage(FastJob* job, int global) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl) return -1;
  if (!global) {
    impl->global = false;
    return 0;
  }
  if (!impl->d_slab) {
    const hipError_t e = hipMalloc(&impl->d_slab_own, impl->slab_stride * impl->max_grid);
    if (e != hipSuccess) {
      g_fast_err = std::string("slab allocation failed: ") + hipGetErrorString(e);
      return -1;
    }
    impl->d_slab = static_cast<uint8_t*>(impl->d_slab_own);
  }
  impl->global = true;
  return 0;
}

int fast_job_global_storage(const FastJob* job) {
  const auto* impl = static_cast<const FastImpl*>(job->impl);
  return impl && impl->global ? 1 : 0;
}

static bool valid_threads(uint32_t t) {
  return t == 64 || t == 128 || t == 256 || t == 512 || t == 1024;
}

int fast_job_set_threads(FastJob* job, uint32_t threads) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl) return -1;
  if (threads != 0 && !valid_threads(threads)) {
    g_fast_err = "threads per cell must be 0, 64, 128, 256, 512 or 1024";
    return -1;
  }
  impl->threads = threads;
  return 0;
}

uint32_t fast_job_threads(const FastJob* job) {
  const auto* impl = static_cast<const FastImpl*>(job->impl);
  return impl ? impl->threads : 0;
}

uint32_t fast_job_tiers(const FastJob* job, uint32_t* threads, uint32_t* grid, uint64_t* cells,
                        uint32_t max, uint64_t* begin) {
  const auto* impl = static_cast<const FastImpl*>(job->impl);
  if (!impl) return 0;
  std::vector<Tier> runs = tiered(impl) ? impl->tiers : std::vector<Tier>{single_launch(impl)};
  uint32_t n = 0;
  for (const Tier& t : runs) {
    if (n == max) break;
    if (threads) threads[n] = t.threads;
    if (grid) grid[n] = launch_grid(impl, t.threads, t.per_cu, t.count);
    if (cells) cells[n] = t.count;
    if (begin) begin[n] = t.begin;
    ++n;
  }
  return n;
}

const uint32_t* fast_job_order(const FastJob* job) {
  const auto* impl = static_cast<const FastImpl*>(job->impl);
  return impl ? impl->order_host.data() : nullptr;
}

int fast_job_set_tier_threads(FastJob* job, uint32_t tier, uint32_t threads) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl || tier >= impl->tiers.size() || !valid_threads(threads)) {
    g_fast_err = "no such tier, or threads per cell not 64, 128, 256, 512 or 1024";
    return -1;
  }
  impl->tiers[tier].threads = threads;
  return 0;
}

uint32_t fast_job_grid(const FastJob* job) {
  uint32_t g[5] = {};
  const uint32_t n = fast_job_tiers(job, nullptr, g, nullptr, 5);
  uint32_t m = 0;
  for (uint32_t k = 0; k < n; ++k) m = std::max(m, g[k]);
  return m;
}

hipError_t fast_phase_times(uint64_t* out, int n) {
  unsigned long long h[kFastPhases] = {};
  hipError_t e = hipMemcpyFromSymbol(h, HIP_SYMBOL(g_fast_phase), sizeof(h), 0, hipMemcpyDeviceToHost);
  if (e != hipSuccess) return e;
  for (int k = 0; k < n && k < kFastPhases; ++k) out[k] = h[k];
  const unsigned long long z[kFastPhases] = {};
  return hipMemcpyToSymbol(HIP_SYMBOL(g_fast_phase), z, sizeof(z), 0, hipMemcpyHostToDevice);
}

void fast_job_destroy(FastJob* job) {
  auto* impl = static_cast<FastImpl*>(job->impl);
  if (!impl) return;
  for (size_t k = 0; k < 5; ++k) {
    if (impl->side[k]) {
      (void)hipStreamSynchronize(impl->side[k]);
      (void)hipStreamDestroy(impl->side[k]);
    }
    if (impl->join[k]) (void)hipEventDestroy(impl->join[k]);
  }
  if (impl->fork) (void)hipEventDestroy(impl->fork);
  if (impl->d_mem) (void)hipFree(impl->d_mem);
  if (impl->d_slab_own) (void)hipFree(impl->d_slab_own);
  delete impl;
  job->impl = nullptr;
}

const char* fast_last_error() { return g_fast_err.c_str(); }

hipError_t launch_zero_words(hipStream_t stream, void* p, uint64_t bytes) {
  const uint64_t n = bytes / 4;
  if (n == 0) return hipSuccess;
  const uint64_t blocks = std::min<uint64_t>((n + 255) / 256, 16384);
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, stream,
                     static_cast<uint32_t*>(p), n);
  return hipGetLastError();
}

hipError_t launch_philox_kat(hipStream_t stream, const uint32_t* d_in, uint32_t* d_out, int n) {
  hipLaunchKernelGGL(philox_kat_kernel, dim3((n + 63) / 64), dim3(64), 0, stream, d_in, d_out, n);
  return hipGetLastError();
}

hipError_t launch_fast_band(hipStream_t stream, uint32_t* d_band, unsigned long long* d_missed,
                            uint32_t nrows, const uint32_t* d_rows, const uint32_t* d_cols,
                            uint64_t n) {
  FastChrom ch{};
  ch.band = d_band;
  ch.missed = d_missed;
  ch.nrows = nrows;
  hipLaunchKernelGGL(fast_band_kernel, dim3(1), dim3(64), 0, stream, ch, d_rows, d_cols, n);
  return hipGetLastError();
}

hipError_t launch_fast_math(hipStream_t stream, int fn, const double* d_in, double* d_out, int n) {
  hipLaunchKernelGGL(fast_math_kernel, dim3((n + 255) / 256), dim3(256), 0, stream, fn, d_in,
                     d_out, n);
  return hipGetLastError();
}

// ---------------------------------------------------------------------------------------------
// fast mode without device allocation (mgpu_simulate_chrom_noalloc)
// ---------------------------------------------------------------------------------------------
// Everything a call keeps in device memory besides the caller's band: counters, the chromosome
// descriptor, the barrier tables, work items, work order and per-cell statistics. Each GPU holds
// its own copy until the process ends; every call on a GPU reuses it, so calls that overlap on
// one GPU race. In host code this name is a stand-in, not a device address: only the pointer
// hipGetSymbolAddress returns for the current device is used.
static __device__ __align__(256) uint8_t noalloc_scratch[kNoAllocScratchBytes];

namespace {

constexpr int kNoAllocMaxDevices = 64;
std::mutex g_noalloc_mtx;
void* g_noalloc_ptr[kNoAllocMaxDevices] = {};

uint64_t round256(uint64_t x) { return (x + 255) & ~255ULL; }

}  // namespace

hipError_t fast_noalloc_init(int device) {
  if (device < 0 || device >= kNoAllocMaxDevices) return hipErrorInvalidDevice;
  std::lock_guard<std::mutex> lock(g_noalloc_mtx);
  if (g_noalloc_ptr[device]) return hipSuccess;
  void* ptr = nullptr;
  hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(noalloc_scratch));
  if (e == hipSuccess && ptr == nullptr) e = hipErrorInvalidSymbol;
  if (e == hipSuccess) e = set_lds_limits();  // the launches below change no kernel attribute
  if (e == hipSuccess) g_noalloc_ptr[device] = ptr;
  return e;
}

int fast_simulate_noalloc(hipStream_t stream, int device, uint32_t num_cus, const mgpu_params& p,
                          const DetChrom& chrom, const uint64_t* pos, const uint8_t* dir,
                          const double* sa, const double* si, double release,
                          double release_burnin, const std::vector<WorkItem>& items,
                          hipEvent_t ev0, hipEvent_t ev1, uint64_t* missed, uint64_t* dropped,
                          uint64_t* sums) {
  void* base_ptr = nullptr;
  {
    std::lock_guard<std::mutex> lock(g_noalloc_mtx);
    if (device >= 0 && device < kNoAllocMaxDevices) base_ptr = g_noalloc_ptr[device];
  }
  if (!base_ptr) {
    g_fast_err = "no scratch memory for this device (mgpu_init sets it up)";
    return -1;
  }
  const uint64_t nbar = chrom.nbar, ncells = items.size();
  if (nbar > kNoAllocMaxBarriers || ncells == 0 || ncells > kNoAllocMaxCells) {
    g_fast_err = "the allocation-free entry point takes up to 65536 barriers and 1 to 16384 cells";
    return -1;
  }
  FastImpl impl;
  if (!fill_fast_params(p, impl.P)) return -1;
  // scratch layout: counters (work queue, missed, dropped x 2), descriptor, barrier tables,
  // directions, work items, order, per-cell statistics
  const uint64_t off_desc = 256;
  const uint64_t off_bars = off_desc + round256(sizeof(FastChrom));
  const uint64_t off_dir = off_bars + round256(16 * nbar);
  const uint64_t off_items = off_dir + round256(nbar);
  const uint64_t off_order = off_items + round256(sizeof(WorkItem) * ncells);
  const uint64_t off_stats = off_order + round256(4 * ncells);
  const uint64_t total = off_stats + round256(32 * ncells);
  if (total > kNoAllocScratchBytes) {
    g_fast_err = "inputs exceed the allocation-free entry point's scratch memory";
    return -1;
  }
  uint8_t* base = static_cast<uint8_t*>(base_ptr);
  DetChrom d = chrom;
  d.bar_dir = base + off_dir;
  d.missed = reinterpret_cast<unsigned long long*>(base + 8);
  d.dropped = reinterpret_cast<unsigned long long*>(base + 16);
  d.occupancy = nullptr;
  const FastChrom f = make_fast_chrom(impl.P, d, reinterpret_cast<uint32_t*>(base + off_bars),
                                      release, release_burnin);
  const uint32_t tier = tier_of(f.lds_bytes, hybrid_lds_bytes(f.nlefs, f.nbar, impl.P.hist_len));
  if (tier < 1 || tier > 3) {
    g_fast_err = "the allocation-free entry point takes cells whose working set fits in LDS";
    return -1;
  }
  // host image of the scratch up to the statistics (the counters start at zero with it)
  std::vector<uint8_t> stage(off_stats, 0);
  std::memcpy(stage.data() + off_desc, &f, sizeof(FastChrom));
  fill_barrier_tables(reinterpret_cast<uint32_t*>(stage.data() + off_bars), nbar, pos, sa, si);
  if (nbar) std::memcpy(stage.data() + off_dir, dir, nbar);
  std::memcpy(stage.data() + off_items, items.data(), sizeof(WorkItem) * ncells);
  uint32_t* order = reinterpret_cast<uint32_t*>(stage.data() + off_order);
  for (uint64_t i = 0; i < ncells; ++i) order[i] = (uint32_t)i;

  impl.d_chroms = reinterpret_cast<FastChrom*>(base + off_desc);
  impl.d_order = reinterpret_cast<uint32_t*>(base + off_order);
  impl.d_next = reinterpret_cast<unsigned long long*>(base);
  impl.num_cus = num_cus;
  impl.nitems = ncells;
  Tier t = tier_shape(tier);
  t.shmem = f.lds_bytes;
  t.begin = 0;
  t.count = ncells;
  const WorkItem* d_items = reinterpret_cast<const WorkItem*>(base + off_items);
  uint64_t* d_stats = reinterpret_cast<uint64_t*>(base + off_stats);
  std::vector<uint64_t> out(4 + 4 * ncells);
  hipError_t e = hipMemcpyAsync(base, stage.data(), stage.size(), hipMemcpyHostToDevice, stream);
  if (e == hipSuccess) e = hipEventRecord(ev0, stream);
  if (e == hipSuccess) e = launch_tier(stream, &impl, t, d_items, d_stats, impl.d_order, ncells);
  if (e == hipSuccess) e = hipEventRecord(ev1, stream);
  if (e == hipSuccess) e = hipMemcpyAsync(out.data(), base, 32, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(out.data() + 4, d_stats, 32 * ncells, hipMemcpyDeviceToHost, stream);
  if (e == hipSuccess) e = hipStreamSynchronize(stream);
  if (e != hipSuccess) {
    g_fast_err = std::string("allocation-free run failed: ") + hipGetErrorString(e);
    return -1;
  }
  *missed = out[1];
  dropped[0] = out[2];
  dropped[1] = out[3];
  sums[0] = sums[1] = sums[2] = 0;
  for (uint64_t i = 0; i < ncells; ++i)
    for (int k = 0; k < 3; ++k) sums[k] += out[4 + 4 * i + k];
  return 0;
}

}  // namespace mgpu

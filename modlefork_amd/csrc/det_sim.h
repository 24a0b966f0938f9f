// Deterministic-mode epoch loop (device).
//
// One wavefront owns one cell; its LEF, rank, move and collision arrays live in LDS when they fit
// (global scratch otherwise). The per-cell engine is a strictly serial dependency chain (every
// step consumes draws whose count depends on earlier outcomes), so lane 0 runs the chain in the
// reference's order:
//   run_burnin            simulation.cpp:1027-1054 (+ compute_loop_size_stats / evaluate_burnin
//                         :957-1025, stats::mean / standard_dev descriptive_impl.hpp:21-105)
//   select_and_bind_lefs  simulation.cpp:1154-1164, simulation_impl.hpp:28-89, rank_lefs :402-487
//   sample_and_register_contacts register_contacts.cpp:73-198
//   generate_moves        simulation.cpp:265-340 (adjust :342-399, clamp :325-340)
//   barrier next_state    extrusion_barriers.cpp:145-161
//   process_collisions    simulation.cpp:925-955 and simulation_detect_collisions.cpp,
//                         simulation_correct_moves.cpp
//   extrude / release     simulation.cpp:489-512, :544-592
// Contacts go straight to the HBM band with 32-bit atomics; integer addition commutes, so the
// band does not depend on how cells interleave.
#pragma once
#ifndef MGPU_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

#include "det_engine.h"
#include "det_sort.h"
#include "kernels.h"

namespace mgpu {

constexpr uint64_t kUnbound = ~0ULL;
constexpr unsigned kEvCollision = 0x10, kEvChrom = 0x08, kEvLefBar = 0x04, kEvPrimary = 0x02,
                   kEvSecondary = 0x01;
constexpr uint64_t kCollIndexMask = (~0ULL) >> 9;
constexpr uint8_t kDirRev = 1, kDirFwd = 2;

__device__ __host__ inline uint64_t cenc(uint64_t idx, unsigned ev) {
  return (idx & kCollIndexMask) | ((uint64_t)ev << 56);
}
__device__ __host__ inline uint64_t cidx(uint64_t c) { return c & kCollIndexMask; }
__device__ __host__ inline unsigned cev(uint64_t c) { return (unsigned)(c >> 56); }
__device__ __host__ inline bool c_any(uint64_t c) { return (cev(c) & kEvCollision) != 0; }
__device__ __host__ inline bool c_is(uint64_t c, unsigned ev) { return cev(c) == (ev | kEvCollision); }
__device__ __host__ inline bool c_avoided(uint64_t c, unsigned ev) { return !c_any(c) && cev(c) == ev; }
__device__ __host__ inline uint64_t umin64(uint64_t a, uint64_t b) { return a < b ? a : b; }
__device__ __host__ inline uint64_t umax64(uint64_t a, uint64_t b) { return a > b ? a : b; }

// Per-cell working set (pointers into LDS or global scratch).
struct DetCell {
  uint64_t *rev, *fwd, *ep;
  uint32_t *rr, *fr, *tmp;
  uint64_t *rm, *fm, *rc, *fc;
  uint8_t* bst;
  double *cfx, *avg;
  uint64_t hlen;
  uint64_t n, nact;
  uint64_t epoch, burnin_epochs, contacts;
  bool burnin_done;
  DetRng g;
};

__device__ inline uint64_t det_cell_bytes(uint64_t nlefs, uint64_t nbar, uint64_t hist) {
  // 7 u64 arrays + 3 u32 arrays per LEF, barrier states, two histories (+1 slack each)
  uint64_t b = nlefs * (7 * 8 + 3 * 4) + nbar + 2 * (hist + 1) * 8;
  return (b + 255) & ~255ULL;
}

__device__ inline void det_carve(DetCell& s, uint8_t* base, uint64_t nlefs, uint64_t nbar,
                                 uint64_t hist) {
  uint64_t* p64 = reinterpret_cast<uint64_t*>(base);
  s.rev = p64;
  s.fwd = p64 + nlefs;
  s.ep = p64 + 2 * nlefs;
  s.rm = p64 + 3 * nlefs;
  s.fm = p64 + 4 * nlefs;
  s.rc = p64 + 5 * nlefs;
  s.fc = p64 + 6 * nlefs;
  double* pd = reinterpret_cast<double*>(p64 + 7 * nlefs);
  s.cfx = pd;
  s.avg = pd + hist + 1;
  uint32_t* p32 = reinterpret_cast<uint32_t*>(pd + 2 * (hist + 1));
  s.rr = p32;
  s.fr = p32 + nlefs;
  s.tmp = p32 + 2 * nlefs;
  s.bst = reinterpret_cast<uint8_t*>(p32 + 3 * nlefs);
}

// ---------------------------------------------------------------------------------------------
// ranks
// ---------------------------------------------------------------------------------------------
__device__ inline void fix_ties(uint32_t* ranks, uint64_t n, const uint64_t* pos,
                                const uint64_t* ep, bool asc) {
  for (uint64_t i = 1; i < n; ++i) {
    if (pos[ranks[i - 1]] != pos[ranks[i]]) continue;
    const uint64_t begin = i - 1;
    for (; i < n; ++i) {
      if (pos[ranks[i - 1]] != pos[ranks[i]]) break;
    }
    // stable insertion sort of ranks[begin, i) by binding epoch
    for (uint64_t c = begin + 1; c < i; ++c) {
      const uint32_t tmp = ranks[c];
      uint64_t k = c;
      while (k > begin) {
        const uint32_t prev = ranks[k - 1];
        const bool lt = asc ? (ep[tmp] < ep[prev]) : (ep[prev] < ep[tmp]);
        if (!lt) break;
        ranks[k] = prev;
        --k;
      }
      ranks[k] = tmp;
    }
  }
}

__device__ inline void det_rank(DetCell& s, uint64_t n, bool partially_sorted) {
  if (partially_sorted) {
    split_sort(s.rr, (uint32_t)n, s.tmp, KeyLess{s.rev});
    split_sort(s.fr, (uint32_t)n, s.tmp, KeyLess{s.fwd});
  } else {
    pdq_sort(s.rr, (uint32_t)n, KeyLess{s.rev});
    pdq_sort(s.fr, (uint32_t)n, KeyLess{s.fwd});
  }
  fix_ties(s.rr, n, s.rev, s.ep, true);
  fix_ties(s.fr, n, s.fwd, s.ep, false);
}

// ---------------------------------------------------------------------------------------------
// moves
// ---------------------------------------------------------------------------------------------
__device__ inline void det_adjust_clamp(const DetChrom& ch, DetCell& s, uint64_t n, bool adjust) {
  if (adjust && n > 0) {
    for (uint64_t i = n - 1; i > 0; --i) {
      const uint32_t i1 = s.rr[i - 1], i2 = s.rr[i];
      if (s.ep[i1] == kUnbound || s.ep[i2] == kUnbound) continue;
      if (s.rev[i1] <= ch.start + s.rm[i1] || s.rev[i2] <= ch.start + s.rm[i2]) continue;
      const uint64_t p1 = s.rev[i1] - s.rm[i1];
      const uint64_t p2 = s.rev[i2] - s.rm[i2];
      if (p2 <= p1) s.rm[i1] += (p1 - p2) + 1;
    }
    for (uint64_t i = 1; i < n; ++i) {
      const uint32_t i1 = s.fr[i - 1], i2 = s.fr[i];
      if (s.ep[i1] == kUnbound || s.ep[i2] == kUnbound) continue;
      if (s.fwd[i1] + s.fm[i1] > ch.end - 1 || s.fwd[i2] + s.fm[i2] > ch.end - 1) continue;
      const uint64_t p1 = s.fwd[i1] + s.fm[i1];
      const uint64_t p2 = s.fwd[i2] + s.fm[i2];
      if (p1 >= p2) s.fm[i2] += (p1 - p2) + 1;
    }
  }
  for (uint64_t i = 0; i < n; ++i) {
    if (s.ep[i] == kUnbound) continue;
    s.rm[i] = umin64(s.rm[i], s.rev[i] - ch.start);
    s.fm[i] = umin64(s.fm[i], ch.end - s.fwd[i] - 1);
  }
}

__device__ inline uint64_t det_move(DetRng& g, uint64_t ep, double speed, double sd) {
  if (ep == kUnbound) return 0;
  if (sd == 0.0) return (uint64_t)round(speed);
  const double v = det_normal(g, speed, sd);
  return (uint64_t)round(0.0 < v ? v : 0.0);
}

__device__ inline void det_generate_moves(const DetParams& P, const DetChrom& ch, DetCell& s,
                                          uint64_t n) {
  const double rs = (double)(s.burnin_done ? P.rev_speed : P.rev_speed_burnin);
  const double fs = (double)(s.burnin_done ? P.fwd_speed : P.fwd_speed_burnin);
  for (uint64_t i = 0; i < n; ++i) s.rm[i] = det_move(s.g, s.ep[i], rs, P.rev_std);
  for (uint64_t i = 0; i < n; ++i) s.fm[i] = det_move(s.g, s.ep[i], fs, P.fwd_std);
  det_adjust_clamp(ch, s, n, true);
}

// ---------------------------------------------------------------------------------------------
// collisions
// ---------------------------------------------------------------------------------------------
__device__ inline void det_boundaries(const DetChrom& ch, DetCell& s, uint64_t n, uint64_t& n5,
                                      uint64_t& n3) {
  n5 = 0;
  n3 = 0;
  const uint64_t first_fwd = s.fwd[s.fr[0]];
  uint64_t k = n;
  while (k > 0 && s.ep[s.rr[k - 1]] == kUnbound) --k;
  const uint64_t last_rev = s.rev[s.rr[k > 0 ? k - 1 : 0]];
  for (uint64_t i = 0; i < n; ++i) {
    const uint32_t idx = s.rr[i];
    const uint64_t pos = s.rev[idx];
    if (pos == ch.start) {
      ++n5;
      s.rc[idx] = cenc(5, kEvCollision | kEvChrom);
    } else if (pos > first_fwd) {
      break;
    } else if (pos - s.rm[idx] == ch.start) {
      s.rc[idx] = cenc(5, kEvCollision | kEvChrom);
      ++n5;
      break;
    }
  }
  for (uint64_t i = n - 1; i > 0; --i) {
    const uint32_t idx = s.fr[i];
    if (s.ep[idx] == kUnbound) {
      ++n3;
      continue;
    }
    const uint64_t pos = s.fwd[idx];
    if (pos == ch.end - 1) {
      ++n3;
      s.fc[idx] = cenc(3, kEvCollision | kEvChrom);
    } else if (pos < last_rev) {
      break;
    } else if (pos + s.fm[idx] == ch.end - 1) {
      s.fc[idx] = cenc(3, kEvCollision | kEvChrom);
      ++n3;
      break;
    }
  }
}

__device__ inline bool lef_bar_trial(double pblock, DetRng& g) {
  return pblock == 1.0 || det_bernoulli(g, pblock);
}
__device__ inline bool lef_lef_trial(const DetParams& P, DetRng& g) {
  return P.p_bypass == 0.0 || det_bernoulli(g, 1.0 - P.p_bypass);
}

__device__ inline void det_lef_bar(const DetParams& P, const DetChrom& ch, DetCell& s, uint64_t n,
                                   uint64_t n5, uint64_t n3) {
  uint64_t j = umin64(n5, n5 - 1);
  uint32_t uidx = s.rr[j];
  uint64_t upos = s.rev[uidx];
  bool rev_done = false;
  for (uint64_t i = 0; i < ch.nbar && !rev_done; ++i) {
    if (s.bst[i] == 0) continue;
    const uint64_t bpos = ch.bar_pos[i];
    const double pblock = ch.bar_dir[i] == kDirRev ? P.pblock_major : P.pblock_minor;
    while (upos <= bpos) {
      if (++j == n) {
        rev_done = true;
        break;
      }
      uidx = s.rr[j];
      upos = s.rev[uidx];
    }
    if (rev_done) break;
    if (s.ep[uidx] != kUnbound) {
      const uint64_t delta = upos - bpos;
      if (delta > 0 && delta <= s.rm[uidx] && lef_bar_trial(pblock, s.g)) {
        s.rc[uidx] = cenc(i, kEvCollision | kEvLefBar);
      }
    }
  }
  j = n - umin64(n3, n3 - 1);
  uidx = s.fr[--j];
  upos = s.fwd[uidx];
  for (uint64_t i = ch.nbar - 1; i != ~0ULL; --i) {
    if (s.bst[i] == 0) continue;
    const uint64_t bpos = ch.bar_pos[i];
    const double pblock = ch.bar_dir[i] == kDirFwd ? P.pblock_major : P.pblock_minor;
    while (upos >= bpos) {
      if (--j == ~0ULL) return;
      uidx = s.fr[j];
      upos = s.fwd[uidx];
    }
    if (s.ep[uidx] != kUnbound) {
      const uint64_t delta = bpos - upos;
      if (delta > 0 && delta <= s.fm[uidx] && lef_bar_trial(pblock, s.g)) {
        s.fc[uidx] = cenc(i, kEvCollision | kEvLefBar);
      }
    }
  }
}

__device__ inline void collision_pos(uint64_t rpos, uint64_t fpos, uint64_t rmove, uint64_t fmove,
                                     uint64_t& prev, uint64_t& pfwd) {
  const uint64_t rel = rmove + fmove;
  const double ttc = (double)(rpos - fpos) / (double)rel;
  const uint64_t cpos = fpos + (uint64_t)round((double)fmove * ttc);
  if (cpos == fpos) {
    prev = cpos + 1;
    pfwd = cpos;
  } else {
    prev = cpos;
    pfwd = cpos - 1;
  }
}

__device__ inline void det_primary(const DetParams& P, const DetChrom& ch, DetCell& s, uint64_t n,
                                   uint64_t n5, uint64_t n3) {
  if (n5 == n || n3 == n) return;
  uint64_t i1 = 0, j1 = n5;
  const uint64_t i2 = n - umin64(n3, n3 - 1);
  for (;;) {
    uint32_t ridx = s.rr[j1];
    uint64_t rpos = s.rev[ridx];
    uint32_t fidx = s.fr[i1];
    uint64_t fpos = s.fwd[fidx];
    while (rpos <= fpos) {
      if (++j1 == n) return;
      ridx = s.rr[j1];
      rpos = s.rev[ridx];
    }
    while (fpos < rpos) {
      if (++i1 == i2) return;
      fidx = s.fr[i1];
      fpos = s.fwd[fidx];
    }
    fidx = s.fr[umin64(i1, i1 - 1)];
    fpos = s.fwd[fidx];
    const uint64_t delta = rpos - fpos;
    if (delta > 0 && delta < s.rm[ridx] + s.fm[fidx] && lef_lef_trial(P, s.g)) {
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
  }
}

__device__ inline void det_correct_lef_bar(const DetChrom& ch, DetCell& s, uint64_t n) {
  for (uint64_t i = 0; i < n; ++i) {
    if (c_is(s.rc[i], kEvLefBar)) s.rm[i] = (s.rev[i] - ch.bar_pos[cidx(s.rc[i])]) - 1;
    if (c_is(s.fc[i], kEvLefBar)) s.fm[i] = (ch.bar_pos[cidx(s.fc[i])] - s.fwd[i]) - 1;
  }
}

__device__ inline void det_correct_primary(DetCell& s, uint64_t n) {
  for (uint64_t k = 0; k < n; ++k) {
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
  for (uint64_t k = 0; k < n; ++k) {
    const uint32_t fidx = s.fr[k];
    if (!c_is(s.fc[fidx], kEvPrimary)) continue;
    const uint64_t ridx = cidx(s.fc[fidx]);
    if (c_is(s.rc[ridx], kEvLefBar)) s.fm[fidx] = (s.rev[ridx] - s.rm[ridx]) - s.fwd[fidx] - 1;
  }
}

__device__ inline void det_secondary(const DetParams& P, DetCell& s, uint64_t n, uint64_t n5,
                                     uint64_t n3) {
  for (uint64_t i = umax64(1, n5); i < n; ++i) {
    const uint32_t r1 = s.rr[i - 1];
    if (!c_any(s.rc[r1])) continue;
    const uint32_t r2 = s.rr[i];
    if (c_any(s.rc[r2])) continue;
    if (s.rev[r2] - s.rm[r2] <= s.rev[r1] - s.rm[r1]) {
      if (lef_lef_trial(P, s.g)) {
        s.rc[r2] = cenc(r1, kEvCollision | kEvSecondary);
        const uint64_t mv = s.rev[r2] - (s.rev[r1] - s.rm[r1]);
        s.rm[r2] = umin64(mv, mv - 1);
      } else {
        s.rc[r2] = cenc(r1, kEvSecondary);
      }
    }
  }
  for (uint64_t i = n - umin64(n3, n3 - 1) - 1; i > 0; --i) {
    const uint32_t f2 = s.fr[i];
    if (!c_any(s.fc[f2])) continue;
    const uint32_t f1 = s.fr[i - 1];
    if (c_any(s.fc[f1])) continue;
    if (s.fwd[f1] + s.fm[f1] >= s.fwd[f2] + s.fm[f2]) {
      if (lef_lef_trial(P, s.g)) {
        s.fc[f1] = cenc(f2, kEvCollision | kEvSecondary);
        const uint64_t mv = (s.fwd[f2] + s.fm[f2]) - s.fwd[f1];
        s.fm[f1] = umin64(mv, mv - 1);
      } else {
        s.fc[f1] = cenc(f2, kEvSecondary);
      }
    }
  }
}

__device__ inline void det_fix_secondary(const DetChrom& ch, DetCell& s, uint64_t n, uint64_t n5,
                                         uint64_t n3) {
  const uint64_t n_active_fwd = n - umin64(n3, n3 - 1);
  for (uint64_t i = umax64(1, n5); i < n; ++i) {
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
  for (uint64_t i = 0; i + 1 < n_active_fwd; ++i) {
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
}

// op as in mgpu_debug_step: 1 without fix_secondary, 2 full.
__device__ inline void det_process_collisions(const DetParams& P, const DetChrom& ch, DetCell& s,
                                              uint64_t n, bool fix) {
  uint64_t n5, n3;
  det_boundaries(ch, s, n, n5, n3);
  det_lef_bar(P, ch, s, n, n5, n3);
  det_primary(P, ch, s, n, n5, n3);
  det_correct_lef_bar(ch, s, n);
  det_correct_primary(s, n);
  det_secondary(P, s, n, n5, n3);
  if (fix) det_fix_secondary(ch, s, n, n5, n3);
}

// ---------------------------------------------------------------------------------------------
// burn-in, contacts, release
// ---------------------------------------------------------------------------------------------
__device__ inline double mean_of(const double* v, uint64_t n) {
  double acc = 0.0;
  for (uint64_t i = 0; i < n; ++i) acc = acc + v[i];
  return acc / (double)n;
}

__device__ inline bool dips_ratio_stable(const double* v, uint64_t cap, uint64_t w) {
  uint64_t cnt = 0;
  for (uint64_t k = 0; k + w + 1 <= cap - 1; ++k) cnt += mean_of(v + k, w) > mean_of(v + k + 1, w);
  const double r = (double)cnt / (double)(cap - w - cnt);
  return r >= 0.95 && r <= 1.05;
}

__device__ inline void det_burnin(const DetParams& P, const DetChrom& ch, DetCell& s) {
  do {
    ++s.burnin_epochs;
    if (s.nact != s.n) {
      const uint64_t k = det_poisson(s.g, ch.bind_rate);
      s.nact = umin64(s.nact + k, s.n);
    } else {
      const uint64_t n = s.nact;
      double acc = 0.0;
      for (uint64_t i = 0; i < n; ++i) acc = acc + (double)(s.fwd[i] - s.rev[i]);
      const double mean = acc / (double)n;
      double ssd = 0.0;
      for (uint64_t i = 0; i < n; ++i) {
        const double d = (double)(s.fwd[i] - s.rev[i]) - mean;
        ssd = ssd + (d * d);
      }
      const double sd = sqrt(ssd / (double)n);
      const uint64_t cap = P.hist_len;
      if (s.hlen == cap) {
        for (uint64_t i = 1; i < cap; ++i) {
          s.avg[i - 1] = s.avg[i];
          s.cfx[i - 1] = s.cfx[i];
        }
        --s.hlen;
      }
      s.avg[s.hlen] = mean;
      s.cfx[s.hlen] = sd / mean;
      ++s.hlen;
      bool done = s.hlen == cap && dips_ratio_stable(s.cfx, cap, P.window) &&
                  dips_ratio_stable(s.avg, cap, P.window);
      done = done && s.epoch > P.min_burnin_epochs;
      if (!done && s.epoch >= P.max_burnin_epochs) done = true;
      s.burnin_done = done;
    }
  } while (s.nact == 0);
}

__device__ inline void band_add(const DetChrom& ch, uint64_t row, uint64_t col) {
  const uint64_t i = row > col ? row - col : col - row;
  const uint64_t j = row > col ? row : col;
  if (i >= ch.nrows) {
    atomicAdd(ch.missed, 1ULL);
    return;
  }
  atomicAdd(&ch.band[j * ch.nrows + i], 1u);
}

__device__ inline void det_register(const DetParams& P, const DetChrom& ch, DetCell& s,
                                    uint64_t nreg, bool tad) {
  const uint64_t sp = ch.start + 1, epp = ch.end - 1;
  const double spd = (double)sp, epd = (double)epp;
  const uint64_t n = s.nact;
  const uint64_t bucket = bucket_of(n - 1);
  while (nreg != 0) {
    const uint64_t i = det_uniform_b(s.g, 0, n - 1, bucket);
    if (s.ep[i] == kUnbound) continue;
    const uint64_t r = s.rev[i], f = s.fwd[i];
    if (!(r > sp && r < epp && f > sp && f < epp)) continue;
    const double a = (double)r - (P.noisify ? det_gev(s.g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
    const double b = (double)f + (P.noisify ? det_gev(s.g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
    const double p1 = b < a ? b : a;
    const double p2 = a < b ? b : a;
    if (!(p1 >= spd && p2 >= spd && p1 < epd && p2 < epd)) continue;
    uint64_t q1 = (uint64_t)p1, q2 = (uint64_t)p2;
    if (tad) {
      const uint64_t lo = q1, hi = q2;
      q1 = det_uniform(s.g, lo, hi);
      q2 = det_uniform(s.g, lo, hi);
    }
    band_add(ch, (q1 - sp) / P.bin_size, (q2 - sp) / P.bin_size);
    --nreg;
  }
}

// register_1d_lef_occupancy (register_contacts.cpp:200-232): nevents LEFs sampled with
// replacement; each accepted one adds the bins of its two (noisified) units.
__device__ inline void det_register_occupancy(const DetParams& P, const DetChrom& ch, DetCell& s,
                                              uint64_t nevents) {
  const uint64_t sp = ch.start + 1, epp = ch.end - 1;
  const double spd = (double)sp, epd = (double)epp;
  const uint64_t n = s.nact;
  const uint64_t bucket = bucket_of(n - 1);
  while (nevents != 0) {
    const uint64_t i = det_uniform_b(s.g, 0, n - 1, bucket);
    if (s.ep[i] == kUnbound) continue;
    const uint64_t r = s.rev[i], f = s.fwd[i];
    if (!(r > sp && r < epp && f > sp && f < epp)) continue;
    const double a = (double)r - (P.noisify ? det_gev(s.g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
    const double b = (double)f + (P.noisify ? det_gev(s.g, P.gev_mu, P.gev_sigma, P.gev_xi) : 0.0);
    const double p1 = b < a ? b : a;
    const double p2 = a < b ? b : a;
    if (!(p1 >= spd && p2 >= spd && p1 < epd && p2 < epd)) continue;
    atomicAdd(&ch.occupancy[((uint64_t)p1 - sp) / P.bin_size], 1ULL);
    atomicAdd(&ch.occupancy[((uint64_t)p2 - sp) / P.bin_size], 1ULL);
    --nevents;
  }
}

__device__ inline void det_sample_contacts(const DetParams& P, const DetChrom& ch, DetCell& s,
                                           uint64_t target) {
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
    nloop = (uint64_t)det_binomial(s.g, (int64_t)n, 1.0 / (P.tad_to_loop_ratio + 1.0), q);
  }
  det_register(P, ch, s, nloop, false);
  det_register(P, ch, s, n - nloop, true);
  if (ch.occupancy) det_register_occupancy(P, ch, s, n);
  s.contacts += n;
}

__device__ inline void det_release(const DetParams& P, const DetChrom& ch, DetCell& s) {
  const double base = s.burnin_done ? P.prob_release : P.prob_release_burnin;
  for (uint64_t i = 0; i < s.nact; ++i) {
    if (s.ep[i] == kUnbound) continue;
    const int h = (c_is(s.rc[i], kEvLefBar) && ch.bar_dir[cidx(s.rc[i])] == kDirRev) +
                  (c_is(s.fc[i], kEvLefBar) && ch.bar_dir[cidx(s.fc[i])] == kDirFwd);
    const double aff = h == 0 ? 1.0 : (h == 1 ? P.inv_soft : P.inv_hard);
    if (det_bernoulli(s.g, aff * base)) {
      s.rev[i] = s.fwd[i] = s.ep[i] = kUnbound;
    }
  }
}

__device__ inline uint64_t det_epoch_hash(const DetCell& s) {
  uint64_t h = 0x84222325cbf29ce4ULL ^ s.g.draws;
  for (uint64_t i = 0; i < s.nact; ++i) {
    h = (h ^ s.rev[i]) * 0x100000001b3ULL;
    h = (h ^ s.fwd[i]) * 0x100000001b3ULL;
  }
  return h ^ (s.contacts << 1);
}

// dump_stats (simulation.cpp:1166-1224): stalls are collisions that occurred (COLLISION bit),
// the typed counts sum both sides, the average loop size is stats::mean over the active LEFs
// (descriptive_impl.hpp:21-30, a sequential double sum).
__device__ inline void det_log_state(const DetParams& P, const DetChrom& ch, const DetCell& s,
                                     uint32_t item) {
  const unsigned long long slot = atomicAdd(P.log.count, 1ULL);
  if (slot >= P.log.cap) return;
  StateRecord r{};
  r.item = item;
  r.epoch = (uint32_t)s.epoch;
  r.burnin = s.burnin_done ? 0u : 1u;
  r.nact = (uint32_t)s.nact;
  for (uint64_t i = 0; i < ch.nbar; ++i) r.active_barriers += s.bst[i] == 1;
  double acc = 0.0;
  for (uint64_t i = 0; i < s.nact; ++i) {
    const bool sr = c_any(s.rc[i]), sf = c_any(s.fc[i]);
    r.stalls_rev += sr;
    r.stalls_fwd += sf;
    r.stalls_both += sr && sf;
    r.lef_bar += (uint32_t)c_is(s.rc[i], kEvLefBar) + (uint32_t)c_is(s.fc[i], kEvLefBar);
    r.primary += (uint32_t)c_is(s.rc[i], kEvPrimary) + (uint32_t)c_is(s.fc[i], kEvPrimary);
    r.secondary += (uint32_t)c_is(s.rc[i], kEvSecondary) + (uint32_t)c_is(s.fc[i], kEvSecondary);
    acc = acc + (double)(s.fwd[i] - s.rev[i]);
  }
  r.avg_loop_size = s.nact == 0 ? 0.0 : acc / (double)s.nact;
  P.log.rec[slot] = r;
}

// simulate_one_cell (simulation.cpp:1056-1152) on lane 0.
__device__ inline void det_simulate_cell(const DetParams& P, const DetChrom& ch, DetCell& s,
                                         uint64_t seed, uint64_t target, uint64_t* trace,
                                         uint64_t max_trace, uint32_t item = 0) {
  const uint64_t N = ch.nlefs;
  s.n = N;
  s.nact = 0;
  s.hlen = 0;
  s.epoch = 0;
  s.burnin_epochs = 0;
  s.contacts = 0;
  s.burnin_done = false;
  for (uint64_t i = 0; i < N; ++i) {
    s.rev[i] = s.fwd[i] = s.ep[i] = kUnbound;
    s.rr[i] = s.fr[i] = (uint32_t)i;
  }
  det_seed(s.g, seed);
  for (uint64_t i = 0; i < ch.nbar; ++i) {
    const double sa = ch.bar_stp_active[i], si = ch.bar_stp_inactive[i];
    double occ = 0.0;
    if (sa + si != 0) {
      const double tia = 1.0 - si, tai = 1.0 - sa;
      occ = tia / (tia + tai);
      occ = occ < 0.0 ? 0.0 : (occ > 1.0 ? 1.0 : occ);
    }
    s.bst[i] = det_bernoulli(s.g, occ) ? 1 : 0;
  }
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
    if (!s.burnin_done) det_burnin(P, ch, s);
    const uint64_t n = s.nact;
    // select + bind: every unbound LEF of the active span draws a position in [start, end-1]
    for (uint64_t i = 0; i < n; ++i) {
      if (s.ep[i] != kUnbound) continue;
      const uint64_t pos = det_uniform_b(s.g, ch.start, bind_range, bind_bucket);
      s.rev[i] = s.fwd[i] = pos;
      s.ep[i] = s.epoch;
    }
    det_rank(s, n, s.epoch != 0);
    if (s.burnin_done) {
      det_sample_contacts(P, ch, s, target);
      if (target != 0 && s.contacts >= target) break;
    }
    det_generate_moves(P, ch, s, n);
    for (uint64_t i = 0; i < ch.nbar; ++i) {
      const double u = det_canonical(s.g);
      if (s.bst[i] == 0 && u > ch.bar_stp_inactive[i]) s.bst[i] = 1;
      else if (s.bst[i] == 1 && u > ch.bar_stp_active[i]) s.bst[i] = 0;
    }
    for (uint64_t i = 0; i < n; ++i) s.rc[i] = s.fc[i] = 0;
    det_process_collisions(P, ch, s, n, true);
    for (uint64_t i = 0; i < n; ++i) {
      if (s.ep[i] == kUnbound) continue;
      s.rev[i] -= s.rm[i];
      s.fwd[i] += s.fm[i];
    }
    if (P.log.rec) det_log_state(P, ch, s, item);
    det_release(P, ch, s);
    if (trace && s.epoch < max_trace) trace[s.epoch] = det_epoch_hash(s);
  }
}

}  // namespace mgpu

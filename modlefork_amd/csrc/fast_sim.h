// Fast-mode epoch loop (device): one workgroup of 64 to 1024 lanes per cell.
//
// Same model as the reference's simulate_one_cell (simulation.cpp:1056-1152), restated for a
// workgroup: every LEF-, unit- and barrier-level step is data-parallel, draws come from a
// counter-based generator (Philox4x32-10 keyed by the cell's seed, counters (element, epoch,
// stream)), so no step waits on a serial engine. The sequential sweeps of the reference become
// scans over rank order:
//   ranks           rank arrays are kept between epochs; each epoch only the LEFs bound in it are
//                   sorted (bitonic, LDS) and merged into the kept order (binary-search ranks);
//                   an order check falls back to a full bitonic sort if anything moved out of order.
//   adjust moves    target positions t = pos -/+ move become suffix-min / prefix-max scans of
//                   t -/+ rank (simulation.cpp:342-399);
//   LEF-BAR         one lane per active barrier finds its nearest unit by binary search; the
//                   closest blocking barrier wins through an LDS atomicMax
//                   (simulation_detect_collisions.cpp:128-252);
//   primary LEF-LEF each fwd unit that is the last one before a rev unit in position order pairs
//                   with it (:255-401); pairs are disjoint, so lanes resolve them independently;
//   secondary       a unit behind a stalled one stalls too unless it bypasses: the chains are
//                   prefix-max scans of segment starts and first failures (:404-518); the rare
//                   bypass fix-ups (:520-647) run on one lane over the compacted list.
// Contacts (register_contacts.cpp:94-198) are sampled one per lane and added to the HBM band
// with 32-bit atomics.
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include "kernels.h"

namespace mgpu {

constexpr int kFastMaxThreads = 1024;
constexpr int kFastMaxWaves = kFastMaxThreads / 64;
constexpr int kFastPhases = 16;  // epoch phases timed when FastParams::timers is set
constexpr uint32_t kFastDefaultThreads = 512;  // lanes per cell of launches from device memory
constexpr uint32_t kFreePos = 0xFFFFFFFFu;
constexpr uint32_t kFastBurninCap = 100000;  // burn-in epochs after which fast mode stops it
// Rejection sampling of a contact (or an occupancy sample) gives up after this many draws and
// counts the sample as dropped; the reference retries without bound
// (register_contacts.cpp:144-160,177-195).
constexpr uint32_t kFastMaxAttempts = 1u << 16;
// Positions and the per-epoch move arithmetic use 32-bit integers (signed where targets may lie
// before the chromosome start): chromosome ends stay at or below kFastMaxEnd (2^30) and raw
// moves are capped at 2^30, so every target, plus or minus a rank below 2^16, fits in int32.
constexpr uint32_t kFastMaxMove = 1u << 30;
// Barrier probabilities are 32-bit thresholds (an event of probability p happens when a draw is
// below p * 2^32); this value stands for p = 1, which no draw fails.
constexpr uint32_t kThresholdOne = 0xFFFFFFFFu;
static_assert(kFastMaxEnd == (1ull << 30), "fast-mode move arithmetic assumes ends <= 2^30");

// collision word: kind in bits 29..31, index in bits 0..28
constexpr uint32_t kCkNone = 0, kCkChrom = 1, kCkBar = 2, kCkPrimary = 3, kCkSecondary = 4,
                   kCkAvoided = 5, kCkBarFwdTmp = 6;
constexpr uint32_t kCkIdxMask = (1u << 29) - 1;
__device__ inline uint32_t ck(uint32_t kind, uint32_t idx) { return (kind << 29) | idx; }
__device__ inline uint32_t ck_kind(uint32_t c) { return c >> 29; }
__device__ inline uint32_t ck_idx(uint32_t c) { return c & kCkIdxMask; }
__device__ inline bool ck_stalled(uint32_t c) {
  const uint32_t k = c >> 29;
  return k != kCkNone && k != kCkAvoided;
}

// ---- Philox4x32-10 (Salmon et al., SC'11) ----
struct U4 {
  uint32_t x, y, z, w;
};
__device__ inline U4 philox(U4 c, uint32_t k0, uint32_t k1) {
#pragma unroll
  for (int r = 0; r < 10; ++r) {
    const uint32_t hi0 = __umulhi(0xD2511F53u, c.x), lo0 = 0xD2511F53u * c.x;
    const uint32_t hi1 = __umulhi(0xCD9E8D57u, c.z), lo1 = 0xCD9E8D57u * c.z;
    c = U4{hi1 ^ c.y ^ k0, lo1, hi0 ^ c.w ^ k1, lo0};
    k0 += 0x9E3779B9u;
    k1 += 0xBB67AE85u;
  }
  return c;
}

enum Stream : uint32_t {
  kSBarInit = 1,
  kSBind = 2,
  kSContact = 3,
  kSMove = 4,
  kSBarrier = 5,
  kSLefBar = 6,
  kSPrimary = 7,
  kSSecondary = 8,
  kSRelease = 9,
  kSPoisson = 10,
  kSOccupancy = 11,
};

struct Rng {
  uint32_t k0, k1;
  __device__ U4 draw(uint32_t elem, uint32_t epoch, uint32_t stream, uint32_t sub = 0) const {
    return philox(U4{elem, epoch, stream, sub}, k0, k1);
  }
};

// uniform doubles with 53 random bits taken from two words (27 from a, 26 from b), the
// resolution of the reference's generate_canonical<double, 53>: in [0, 1), and in (0, 1] for logs
__device__ inline double u01d(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6)) * (1.0 / 9007199254740992.0);
}
__device__ inline double u01d_open(uint32_t a, uint32_t b) {
  return ((double)(a >> 5) * 67108864.0 + (double)(b >> 6) + 1.0) * (1.0 / 9007199254740992.0);
}
// integer in [0, range) by multiply-shift (Lemire); bias < range / 2^32, negligible here
__device__ inline uint32_t ubounded(uint32_t x, uint32_t range) {
  return (uint32_t)(((uint64_t)x * range) >> 32);
}
// Bernoulli with probability threshold t = p * 2^32 (t = 2^32 never fires false)
__device__ inline bool bern(uint32_t x, uint64_t t) { return (uint64_t)x < t; }
__device__ __host__ inline uint64_t thresh(double p) {
  if (p <= 0.0) return 0;
  if (p >= 1.0) return 1ULL << 32;
  return (uint64_t)(p * 4294967296.0);
}

// ---- per-job constants ----
struct FastParams {
  uint32_t bin_size;
  double rev_speed, fwd_speed, rev_speed_burnin, fwd_speed_burnin;
  double rev_std, fwd_std;
  uint64_t t_bypass_keep;    // threshold of (1 - p_bypass): collision happens
  uint64_t t_pblock_major, t_pblock_minor;
  double inv_soft, inv_hard;
  uint64_t t_loop;           // threshold of 1 / (ratio + 1): a contact is a loop contact
  double gev_mu, gev_sigma, gev_xi;
  int32_t noisify, skip_burnin;
  uint32_t hist_len, window;
  uint64_t min_burnin_epochs, max_burnin_epochs;
  double target_contact_density;
  uint32_t timers;  // nonzero: lane 0 accumulates wall-clock ticks per epoch phase
  StateLog log;     // per-epoch state records (rec == nullptr: off)
  uint32_t serial_fixups;  // nonzero: bypass fix-ups walk their list on one lane (check only)
};

struct FastChrom {
  uint32_t start, end;
  uint32_t nlefs, nbar;
  const uint32_t* bar_pos;
  const uint8_t* bar_dir;
  const uint32_t* bar_t_active;    // threshold of stp_active
  const uint32_t* bar_t_inactive;  // threshold of stp_inactive
  const uint32_t* bar_t_occ;       // threshold of the stationary occupancy
  uint32_t* band;
  unsigned long long* missed;
  // samples abandoned after kFastMaxAttempts rejected draws: [0] contacts, [1] 1D occupancy
  unsigned long long* dropped;
  uint32_t nrows;
  uint32_t cpe;
  uint64_t target_epochs;
  double bind_rate;                   // mean LEF activations per burn-in epoch
  double bind_exp_neg, bind_sd;       // e^-bind_rate and sqrt(bind_rate)
  double prob_release, prob_release_burnin;  // per chromosome: LEF processivity may vary
  uint64_t lds_bytes;  // working-set bytes of one cell
  unsigned long long* occupancy;  // 1D LEF occupancy, ceil(size / bin) bins, or null
};

}  // namespace mgpu

// C-ABI of the engine (include/modle_gpu.h): parameters, planning, device jobs and the debug
// entry points. Host code only; kernels live in det_kernels.hip and fast_kernels.hip.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstring>
#include <limits>
#include <string>
#include <vector>

#include "../../include/modle_gpu.h"
#include "det_sort.h"
#include "fast_kernels.h"
#include "kernels.h"
#include "xxh3_host.h"

using namespace mgpu;

struct mgpu_ctx {
  int device = 0;
  hipStream_t stream = nullptr;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  hipEvent_t lev[10] = {};  // start and end of each launch of a run (fast mode: one per tier)
  std::string err;
  int num_cus = 0;
  int phase_timers = 0;
  int det_serial = 0;  // deterministic mode: 1 runs every cell on the one-lane kernel
  uint64_t det_slab_slots = 0;  // deterministic mode: cap on device-memory cell slots (0: none)
  std::string noalloc_err;  // why mgpu_simulate_chrom_noalloc cannot run on this device, if so
};

namespace {

thread_local std::string g_err;

int fail(const mgpu_ctx* ctx, int code, const char* fmt, ...) {
  char buf[1024];
  va_list ap;
  va_start(ap, fmt);
  vsnprintf(buf, sizeof(buf), fmt, ap);
  va_end(ap);
  if (ctx) const_cast<mgpu_ctx*>(ctx)->err = buf;
  g_err = buf;
  return code;
}

#define HIPCHK(ctx, expr)                                                                      \
  do {                                                                                         \
    hipError_t e_ = (expr);                                                                    \
    if (e_ != hipSuccess) return fail(ctx, MGPU_EDEVICE, "%s: %s", #expr, hipGetErrorString(e_)); \
  } while (0)

double clamp01(double v) { return v < 0.0 ? 0.0 : (v > 1.0 ? 1.0 : v); }

uint64_t ceil_div(uint64_t a, uint64_t b) { return (a + b - 1) / b; }

uint64_t contacts_per_epoch(const mgpu_params* p, uint64_t nlefs) {
  const double speed = (double)(p->rev_speed + p->fwd_speed);
  const double prob = speed / (double)p->contact_sampling_interval;
  return (uint64_t)std::max(1.0, std::round((double)nlefs * prob));
}

DetParams det_params(const mgpu_params* p) {
  DetParams P{};
  P.bin_size = p->bin_size;
  P.rev_speed = p->rev_speed;
  P.fwd_speed = p->fwd_speed;
  P.rev_speed_burnin = p->rev_speed_burnin;
  P.fwd_speed_burnin = p->fwd_speed_burnin;
  P.rev_std = p->rev_speed_std;
  P.fwd_std = p->fwd_speed_std;
  P.prob_release = p->prob_release;
  P.prob_release_burnin = p->prob_release_burnin;
  P.p_bypass = p->p_bypass;
  P.pblock_major = p->pblock_major;
  P.pblock_minor = p->pblock_minor;
  P.inv_soft = 1.0 / p->soft_stall_mult;
  P.inv_hard = 1.0 / p->hard_stall_mult;
  P.tad_to_loop_ratio = p->tad_to_loop_ratio;
  P.gev_mu = p->gev_mu;
  P.gev_sigma = p->gev_sigma;
  P.gev_xi = p->gev_xi;
  P.noisify = p->noisify;
  P.skip_burnin = p->skip_burnin;
  P.hist_len = p->burnin_history_length;
  P.window = p->burnin_smoothing_window;
  P.min_burnin_epochs = p->min_burnin_epochs;
  P.max_burnin_epochs = p->max_burnin_epochs;
  P.target_contact_density = p->target_contact_density;
  return P;
}

int check_params(const mgpu_params* p) {
  if (!p) return fail(nullptr, MGPU_EINVAL, "null parameter block");
  if (p->bin_size == 0) return fail(nullptr, MGPU_EINVAL, "bin_size must be > 0");
  if (p->contact_sampling_interval == 0)
    return fail(nullptr, MGPU_EINVAL, "contact_sampling_interval must be > 0");
  if (p->burnin_history_length <= p->burnin_smoothing_window)
    return fail(nullptr, MGPU_EINVAL, "burn-in history must be longer than its smoothing window");
  if (p->burnin_target_epochs_for_lef_activation == 0)
    return fail(nullptr, MGPU_EINVAL, "burn-in LEF activation epochs must be > 0");
  return MGPU_OK;
}

int check_chrom(const mgpu_chrom* c) {
  if (!c || !c->name) return fail(nullptr, MGPU_EINVAL, "null chromosome");
  if (c->end <= c->start + 2)
    return fail(nullptr, MGPU_EINVAL, "chromosome %.*s: simulated range is empty",
                (int)c->name_len, c->name);
  if (c->end > c->size) return fail(nullptr, MGPU_EINVAL, "chromosome range beyond its size");
  if (c->name_len > 224) return fail(nullptr, MGPU_ERANGE, "chromosome name longer than 224 bytes");
  return MGPU_OK;
}

}  // namespace

extern "C" {

const char* mgpu_version(void) { return "modle-mi355x 0.1.0"; }

const char* mgpu_last_error(const mgpu_ctx* ctx) { return ctx ? ctx->err.c_str() : g_err.c_str(); }

int mgpu_init(int device, mgpu_ctx** out) {
  if (!out) return fail(nullptr, MGPU_EINVAL, "null output pointer");
  int ndev = 0;
  hipError_t e = hipGetDeviceCount(&ndev);
  if (e != hipSuccess || ndev == 0) return fail(nullptr, MGPU_EDEVICE, "no HIP device available");
  if (device < 0 || device >= ndev) return fail(nullptr, MGPU_EINVAL, "device %d out of range", device);
  auto* ctx = new mgpu_ctx();
  ctx->device = device;
  HIPCHK(ctx, hipSetDevice(device));
  HIPCHK(ctx, hipStreamCreateWithFlags(&ctx->stream, hipStreamNonBlocking));
  HIPCHK(ctx, hipEventCreate(&ctx->ev0));
  HIPCHK(ctx, hipEventCreate(&ctx->ev1));
  for (auto& e : ctx->lev) HIPCHK(ctx, hipEventCreate(&e));
  hipDeviceProp_t prop;
  HIPCHK(ctx, hipGetDeviceProperties(&prop, device));
  ctx->num_cus = prop.multiProcessorCount;
  // this device's scratch memory for mgpu_simulate_chrom_noalloc (resolved once per device); a
  // failure is kept and reported by that entry point, the others do not need it
  const hipError_t ne = fast_noalloc_init(device);
  if (ne != hipSuccess)
    ctx->noalloc_err = std::string("scratch memory set-up failed: ") + hipGetErrorString(ne);
  *out = ctx;
  return MGPU_OK;
}

void mgpu_destroy(mgpu_ctx* ctx) {
  if (!ctx) return;
  (void)hipSetDevice(ctx->device);
  if (ctx->ev0) (void)hipEventDestroy(ctx->ev0);
  if (ctx->ev1) (void)hipEventDestroy(ctx->ev1);
  for (auto e : ctx->lev)
    if (e) (void)hipEventDestroy(e);
  if (ctx->stream) (void)hipStreamDestroy(ctx->stream);
  delete ctx;
}

// ---- parameters --------------------------------------------------------------------------------

int mgpu_config_defaults(mgpu_config* c) {
  if (!c) return fail(nullptr, MGPU_EINVAL, "null config");
  *c = mgpu_config{};
  c->bin_size = 5000;
  c->diagonal_width = 3000000;
  c->sample_loop_contacts = 1;
  c->sample_tad_contacts = 1;
  c->noisify_contacts = 1;
  c->stopping_criterion = MGPU_STOP_CONTACT_DENSITY;
  c->tad_to_loop_contact_ratio = 5.0;
  c->genextreme_mu = 0.0;
  c->genextreme_sigma = 5000.0;
  c->genextreme_xi = 0.001;
  c->rev_extrusion_speed = c->fwd_extrusion_speed = c->bin_size * 8 / 10;
  c->rev_extrusion_speed_std = c->fwd_extrusion_speed_std = 0.05;
  c->number_of_lefs_per_mbp = 20.0;
  c->avg_lef_processivity = 300000;
  c->contact_sampling_interval = 50000;
  c->extrusion_barrier_occupancy = 0.825;
  c->barrier_occupied_stp = 0.0;
  c->barrier_not_occupied_stp = 0.70;
  c->hard_stall_lef_stability_multiplier = 5.0;
  c->soft_stall_lef_stability_multiplier = 1.0;
  c->probability_of_extrusion_unit_bypass = 0.1;
  c->lef_bar_major_collision_pblock = 1.0;
  c->lef_bar_minor_collision_pblock = 0.0;
  c->target_simulation_epochs = 2000;
  c->target_contact_density = 1.0;
  c->num_cells = 512;
  c->seed = 0;
  c->probability_normalization_factor = 8000;
  c->normalize_probabilities = 1;
  c->skip_burnin = 0;
  c->burnin_history_length = 100;
  c->burnin_smoothing_window_size = 5;
  c->min_burnin_epochs = 0;
  c->max_burnin_epochs = std::numeric_limits<uint64_t>::max();
  c->burnin_speed_coefficient = 1.0;
  return MGPU_OK;
}

double mgpu_stp_active_from_occupancy(double stp_inactive, double occupancy) {
  if (occupancy == 0) return 0.0;
  const double tp_inactive_to_active = 1.0 - stp_inactive;
  const double tp_active_to_inactive =
      (tp_inactive_to_active - (occupancy * tp_inactive_to_active)) / occupancy;
  return clamp01(1.0 - tp_active_to_inactive);
}

double mgpu_occupancy_from_stp(double stp_active, double stp_inactive) {
  if (stp_active + stp_inactive == 0) return 0.0;
  const double tp_inactive_to_active = 1.0 - stp_inactive;
  const double tp_active_to_inactive = 1.0 - stp_active;
  return clamp01(tp_inactive_to_active / (tp_inactive_to_active + tp_active_to_inactive));
}

// Cli::transform_args (cli.cpp:1192-1215) with its helpers cli_update_extr_speed (:1085-1110),
// cli_compute_prob_of_lef_release (:1112-1119), cli_update_barrier_stp_and_occupancy (:1122-1135),
// cli_update_tad_to_loop_contact_ratio (:1169-1182), cli_update_burnin_params (:1184-1190) and
// cli_normalize_probabilities (:1138-1167).
int mgpu_config_transform(const mgpu_config* in, uint32_t set_mask, mgpu_params* out) {
  if (!in || !out) return fail(nullptr, MGPU_EINVAL, "null argument");
  if (in->bin_size == 0) return fail(nullptr, MGPU_EINVAL, "bin_size must be > 0");
  if (!in->sample_loop_contacts && !in->sample_tad_contacts)
    return fail(nullptr, MGPU_EINVAL, "at least one of loop/TAD contact sampling must be enabled");
  mgpu_config c = *in;
  if (!(set_mask & MGPU_SET_REV_SPEED)) c.rev_extrusion_speed = c.bin_size * 8 / 10;
  if (!(set_mask & MGPU_SET_FWD_SPEED)) c.fwd_extrusion_speed = c.bin_size * 8 / 10;
  if (c.fwd_extrusion_speed_std > 0 && c.fwd_extrusion_speed_std < 1)
    c.fwd_extrusion_speed_std *= (double)c.fwd_extrusion_speed;
  if (c.rev_extrusion_speed_std > 0 && c.rev_extrusion_speed_std < 1)
    c.rev_extrusion_speed_std *= (double)c.rev_extrusion_speed;
  const uint64_t rev_burnin =
      (uint64_t)std::round(c.burnin_speed_coefficient * (double)c.rev_extrusion_speed);
  const uint64_t fwd_burnin =
      (uint64_t)std::round(c.burnin_speed_coefficient * (double)c.fwd_extrusion_speed);
  if (c.avg_lef_processivity == 0) return fail(nullptr, MGPU_EINVAL, "processivity must be > 0");
  if (rev_burnin + fwd_burnin == 0) return fail(nullptr, MGPU_EINVAL, "extrusion speeds are zero");

  mgpu_params p{};
  p.bin_size = c.bin_size;
  p.diagonal_width = c.diagonal_width;
  p.rev_speed = c.rev_extrusion_speed;
  p.fwd_speed = c.fwd_extrusion_speed;
  p.rev_speed_burnin = rev_burnin;
  p.fwd_speed_burnin = fwd_burnin;
  p.rev_speed_std = c.rev_extrusion_speed_std;
  p.fwd_speed_std = c.fwd_extrusion_speed_std;
  p.prob_release =
      (double)(c.rev_extrusion_speed + c.fwd_extrusion_speed) / (double)c.avg_lef_processivity;
  p.prob_release_burnin = (double)(rev_burnin + fwd_burnin) / (double)c.avg_lef_processivity;

  double occupancy = c.extrusion_barrier_occupancy;
  double stp_active = c.barrier_occupied_stp;
  double stp_inactive = c.barrier_not_occupied_stp;
  if (set_mask & MGPU_SET_BARRIER_OCCUPANCY) {
    stp_active = mgpu_stp_active_from_occupancy(stp_inactive, occupancy);
  } else {
    occupancy = mgpu_occupancy_from_stp(stp_active, stp_inactive);
  }

  double ratio_tl = c.tad_to_loop_contact_ratio;
  if (c.sample_loop_contacts && !c.sample_tad_contacts) ratio_tl = 0.0;
  if (!c.sample_loop_contacts && c.sample_tad_contacts)
    ratio_tl = std::numeric_limits<double>::infinity();

  const uint64_t act_bp = 5 * c.avg_lef_processivity;
  p.burnin_target_epochs_for_lef_activation =
      std::min(c.max_burnin_epochs, act_bp / (rev_burnin + fwd_burnin));

  double p_bypass = c.probability_of_extrusion_unit_bypass;
  double pblock_major = c.lef_bar_major_collision_pblock;
  double pblock_minor = c.lef_bar_minor_collision_pblock;
  if (c.normalize_probabilities) {
    const double ratio = (double)(c.rev_extrusion_speed + c.fwd_extrusion_speed) /
                         (double)c.probability_normalization_factor;
    if (ratio != 1.0) {
      auto stable_pow = [](double base, double e) {
        if (base == 0.0) return 0.0;
        if (base == 1.0) return 1.0;
        return std::exp(std::log(base) * e);
      };
      stp_inactive = stable_pow(stp_inactive, ratio);
      stp_active = mgpu_stp_active_from_occupancy(stp_inactive, occupancy);
      if (p_bypass != 0.0 && p_bypass != 1.0) p_bypass = std::min(p_bypass * ratio, 1.0);
      pblock_major = stable_pow(pblock_major, ratio);
      pblock_minor = stable_pow(pblock_minor, ratio);
    }
  }
  p.p_bypass = p_bypass;
  p.pblock_major = pblock_major;
  p.pblock_minor = pblock_minor;
  p.hard_stall_mult = c.hard_stall_lef_stability_multiplier;
  p.soft_stall_mult = c.soft_stall_lef_stability_multiplier;
  p.tad_to_loop_ratio = ratio_tl;
  p.gev_mu = c.genextreme_mu;
  p.gev_sigma = c.genextreme_sigma;
  p.gev_xi = c.genextreme_xi;
  p.noisify = c.noisify_contacts ? 1 : 0;
  p.skip_burnin = c.skip_burnin ? 1 : 0;
  p.contact_sampling_interval = c.contact_sampling_interval;
  p.burnin_history_length = c.burnin_history_length;
  p.burnin_smoothing_window = c.burnin_smoothing_window_size;
  p.min_burnin_epochs = c.min_burnin_epochs;
  p.max_burnin_epochs = c.max_burnin_epochs;
  p.target_contact_density =
      c.stopping_criterion == MGPU_STOP_EPOCHS ? -1.0 : c.target_contact_density;
  p.target_simulation_epochs = c.target_simulation_epochs;
  p.seed = c.seed;
  p.lefs_per_mbp = c.number_of_lefs_per_mbp;
  p.barrier_occupied_stp = stp_active;
  p.barrier_not_occupied_stp = stp_inactive;
  p.override_barrier_occupancy = (set_mask & MGPU_SET_BARRIER_OCCUPANCY) ? 1 : 0;
  p.track_1d_lef_position = c.track_1d_lef_position ? 1 : 0;
  p.log_model_internal_state = c.log_model_internal_state ? 1 : 0;
  *out = p;
  return MGPU_OK;
}

namespace {
// The permutation cpp-sort's pdq_sort makes of indices [0, n) ordered by key (unstable: equal
// keys end up in pdq_sort's order).
std::vector<uint32_t> pdq_order(const std::vector<uint64_t>& key) {
  std::vector<uint32_t> idx(key.size());
  for (uint32_t i = 0; i < idx.size(); ++i) idx[i] = i;
  pdq_sort(idx.data(), (uint32_t)idx.size(), KeyLess{key.data()});
  return idx;
}
}  // namespace

int mgpu_import_barriers(const mgpu_params* p, const mgpu_bed_record* recs,
                         const double* name_puu, uint64_t n, uint64_t chrom_start,
                         uint64_t chrom_end, uint64_t* bed_start, uint64_t* bed_end,
                         uint64_t* pos, uint8_t* dir, double* stp_active, double* stp_inactive,
                         uint64_t* nout) {
  if (!p || (n && !recs) || !nout) return fail(nullptr, MGPU_EINVAL, "null argument");
  if (n > 0xFFFFFFFFull) return fail(nullptr, MGPU_ERANGE, "more than 2^32 records");
  for (uint64_t i = 0; i < n; ++i) {
    const mgpu_bed_record& r = recs[i];
    if (r.strand != '+' && r.strand != '-' && r.strand != '.')
      return fail(nullptr, MGPU_EINVAL, "record %llu: invalid strand", (unsigned long long)i);
    if (r.score < 0.0 || r.score > 1.0)
      return fail(nullptr, MGPU_EINVAL, "record %llu: score must be in [0, 1]",
                  (unsigned long long)i);
    // the name field as the record's not-bound stp (genome.cpp:459-468), range-checked for
    // every record before the position test drops any
    if (name_puu && !(name_puu[i] >= 0.0 && name_puu[i] <= 1.0))
      return fail(nullptr, MGPU_EINVAL, "record %llu: name field must be in [0, 1]",
                  (unsigned long long)i);
  }
  // The parser's per-chromosome interval tree orders the records by start (IITree::make_BST,
  // interval_tree_impl.hpp:58-68, a pdq_sort of the indices), Genome::import_barriers walks
  // them in that order and the chromosome's own tree keeps those in range
  // (Chromosome::add_extrusion_barrier, genome.cpp:194-210), ordered by start once more when
  // import_barriers indexes it (genome.cpp:474).
  std::vector<uint64_t> key(n);
  for (uint64_t i = 0; i < n; ++i) key[i] = recs[i].chrom_start;
  std::vector<uint32_t> kept;
  kept.reserve(n);
  for (const uint32_t i : pdq_order(key)) {
    const uint64_t bp = (recs[i].chrom_start + recs[i].chrom_end + 1) / 2;
    if (bp >= chrom_start && bp < chrom_end) kept.push_back(i);
  }
  std::vector<uint64_t> key2(kept.size());
  for (uint64_t k = 0; k < kept.size(); ++k) key2[k] = recs[kept[k]].chrom_start;
  const std::vector<uint32_t> order = pdq_order(key2);
  for (uint64_t k = 0; k < kept.size(); ++k) {
    const uint32_t i = kept[order[k]];
    const mgpu_bed_record& r = recs[i];
    // compute_barrier_stp (genome.cpp:181-192) with the default or the record's not-bound stp,
    // then the occupancy override of Simulation::Simulation (simulation.cpp:68-78), which
    // restores the configured pair
    double sa = p->barrier_occupied_stp, si = name_puu ? name_puu[i] : p->barrier_not_occupied_stp;
    if (p->override_barrier_occupancy) {
      si = p->barrier_not_occupied_stp;
    } else if (r.score != 0.0) {
      sa = mgpu_stp_active_from_occupancy(si, r.score);
    }
    if (bed_start) bed_start[k] = r.chrom_start;
    if (bed_end) bed_end[k] = r.chrom_end;
    pos[k] = (r.chrom_start + r.chrom_end + 1) / 2;
    // motif '+' blocks extrusion in the rev direction, anything else in the fwd direction
    dir[k] = (uint8_t)(r.strand == '+' ? 1 : 2);
    stp_active[k] = sa;
    stp_inactive[k] = si;
  }
  *nout = kept.size();
  return MGPU_OK;
}

int mgpu_pdq_order(const uint64_t* keys, uint64_t n, uint32_t* order) {
  if (n && (!keys || !order)) return fail(nullptr, MGPU_EINVAL, "null argument");
  if (n > 0xFFFFFFFFull) return fail(nullptr, MGPU_ERANGE, "more than 2^32 keys");
  const std::vector<uint32_t> idx = pdq_order(std::vector<uint64_t>(keys, keys + n));
  std::copy(idx.begin(), idx.end(), order);
  return MGPU_OK;
}

int mgpu_sort_barriers(uint64_t n, uint64_t* pos, uint8_t* dir, double* stp_active,
                       double* stp_inactive) {
  if (n && (!pos || !dir || !stp_active || !stp_inactive))
    return fail(nullptr, MGPU_EINVAL, "null argument");
  if (n > 0xFFFFFFFFull) return fail(nullptr, MGPU_ERANGE, "more than 2^32 barriers");
  // ExtrusionBarriers::sort (extrusion_barriers.cpp:239-259): pdq_sort of the indices by
  // position, then the permutation applied to every field
  const std::vector<uint64_t> key(pos, pos + n);
  const std::vector<uint32_t> idx = pdq_order(key);
  const std::vector<uint8_t> d0(dir, dir + n);
  const std::vector<double> a0(stp_active, stp_active + n), i0(stp_inactive, stp_inactive + n);
  for (uint64_t k = 0; k < n; ++k) {
    pos[k] = key[idx[k]];
    dir[k] = d0[idx[k]];
    stp_active[k] = a0[idx[k]];
    stp_inactive[k] = i0[idx[k]];
  }
  return MGPU_OK;
}

int mgpu_prepare_barriers(const mgpu_params* p, const mgpu_bed_record* recs, uint64_t n,
                          uint64_t chrom_start, uint64_t chrom_end, uint64_t* pos, uint8_t* dir,
                          double* stp_active, double* stp_inactive, uint64_t* nout) {
  return mgpu_prepare_barriers_name_puu(p, recs, nullptr, n, chrom_start, chrom_end, pos, dir,
                                        stp_active, stp_inactive, nout);
}

int mgpu_prepare_barriers_name_puu(const mgpu_params* p, const mgpu_bed_record* recs,
                                   const double* name_puu, uint64_t n, uint64_t chrom_start,
                                   uint64_t chrom_end, uint64_t* pos, uint8_t* dir,
                                   double* stp_active, double* stp_inactive, uint64_t* nout) {
  int rc = mgpu_import_barriers(p, recs, name_puu, n, chrom_start, chrom_end, nullptr, nullptr,
                                pos, dir, stp_active, stp_inactive, nout);
  if (rc) return rc;
  return mgpu_sort_barriers(*nout, pos, dir, stp_active, stp_inactive);
}

int mgpu_plan_chrom(const mgpu_params* p, const mgpu_chrom* chrom, uint64_t ncells,
                    uint64_t* nlefs, uint64_t* nrows, uint64_t* ncols, uint64_t* target_contacts,
                    uint64_t* target_epochs) {
  if (int rc = check_params(p)) return rc;
  if (int rc = check_chrom(chrom)) return rc;
  if (ncells == 0) return fail(nullptr, MGPU_EINVAL, "ncells must be > 0");
  const uint64_t sim_size = chrom->end - chrom->start;
  const uint64_t nl = std::max<uint64_t>(
      1, (uint64_t)std::round(p->lefs_per_mbp * ((double)sim_size / 1.0e6)));
  const uint64_t nc = ceil_div(sim_size, p->bin_size);
  const uint64_t nr = std::min(ceil_div(p->diagonal_width, p->bin_size), nc);
  const uint64_t npixels = nr * nc;
  if (nlefs) *nlefs = nl;
  if (nrows) *nrows = nr;
  if (ncols) *ncols = nc;
  const uint64_t tot = (uint64_t)std::round((double)npixels * p->target_contact_density);
  if (target_contacts) {
    const uint64_t per_cell = (tot + ncells - 1) / ncells;
    uint64_t rolling = 0;
    for (uint64_t c = 0; c < ncells; ++c) {
      const uint64_t eff = std::min(per_cell, tot - rolling);
      rolling += eff;
      target_contacts[c] = eff;
    }
  }
  if (target_epochs) {
    if (p->target_contact_density < 0) {
      *target_epochs = ncells * p->target_simulation_epochs;
    } else {
      const double tot_d = std::max(1.0, std::round(p->target_contact_density * (double)npixels));
      *target_epochs = (uint64_t)std::round(tot_d / (double)contacts_per_epoch(p, nl));
    }
  }
  return MGPU_OK;
}

uint64_t mgpu_cell_seed(const mgpu_chrom* chrom, uint64_t cell_id, uint64_t seed) {
  if (!chrom || chrom->name_len > 224) return 0;
  uint8_t buf[240];
  memcpy(buf, chrom->name, chrom->name_len);
  for (int i = 0; i < 8; ++i) buf[chrom->name_len + i] = (uint8_t)(chrom->size >> (8 * i));
  for (int i = 0; i < 8; ++i) buf[chrom->name_len + 8 + i] = (uint8_t)(cell_id >> (8 * i));
  uint64_t h = 0;
  xxh3::hash64(buf, chrom->name_len + 16, seed, &h);
  return h;
}

// ---- jobs -----------------------------------------------------------------------------------

}  // extern "C"

struct mgpu_job {
  mgpu_ctx* ctx = nullptr;
  int mode = 0;
  mgpu_params p{};
  uint64_t nchroms = 0;
  std::vector<DetChrom> chroms;          // device pointers inside
  std::vector<uint64_t> band_off, band_len, nrows;
  std::vector<uint64_t> items_per_chrom;
  uint64_t nitems = 0;
  uint64_t band_total = 0;
  void* d_pool = nullptr;                // one allocation for all inputs and counters
  uint32_t* d_bands = nullptr;
  unsigned long long* d_missed = nullptr;
  unsigned long long* d_dropped = nullptr;  // 2 per chromosome: contacts, occupancy samples
  unsigned long long* d_next = nullptr;
  DetChrom* d_chroms = nullptr;
  WorkItem* d_items = nullptr;
  uint64_t* d_cell_stats = nullptr;
  uint8_t* d_scratch = nullptr;
  uint64_t stride = 0;
  bool use_lds = false;
  uint32_t grid = 0;
  // deterministic mode, wave-parallel kernel: LDS per block, cells whose run failed; for cells
  // too large for LDS, the bytes of one cell's arrays in device memory and the number of slots
  // (cells resident at once; launches take the work items in chunks of that many)
  uint64_t wave_bytes = 0;
  uint64_t wave_stride = 0;
  uint64_t wave_slots = 0;
  uint8_t* d_wave_slab = nullptr;
  uint32_t* d_det_order = nullptr;  // work items by expected cost, largest first
  unsigned long long* d_err = nullptr;
  FastJob fast{};
  hipGraph_t graph = nullptr;       // captured run (mgpu_job_capture)
  hipGraphExec_t graph_exec = nullptr;
  // optional outputs: 1D LEF occupancy (in the pool) and the state log (own allocation)
  unsigned long long* d_occ = nullptr;
  std::vector<uint64_t> occ_off, occ_len;
  uint64_t occ_total = 0;
  StateLog log{};
  void* d_log = nullptr;
  uint64_t log_default = 0;  // records a job created with log_model_internal_state starts with
  std::vector<float> launch_ms;  // device time of each launch of the last uncaptured run
  std::vector<uint32_t> item_chrom;  // chromosome of each work item
};

namespace {

struct PoolBuilder {
  std::vector<std::pair<uint64_t, uint64_t>> spans;  // offset, size
  uint64_t total = 0;
  uint64_t add(uint64_t bytes) {
    const uint64_t off = total;
    total += (bytes + 255) & ~255ULL;
    return off;
  }
};

}  // namespace

namespace {

// Parameter blocks of one multi-parameter job may differ only in the fields that describe LEF
// density and processivity (the release probabilities and the burn-in activation horizon that
// Cli::transform_args derives from it); everything the kernels read job-wide must match the
// first block.
bool same_job_wide(const mgpu_params& a, const mgpu_params& b) {
  mgpu_params x = a, y = b;
  x.prob_release = y.prob_release = 0;
  x.prob_release_burnin = y.prob_release_burnin = 0;
  x.lefs_per_mbp = y.lefs_per_mbp = 0;
  x.burnin_target_epochs_for_lef_activation = y.burnin_target_epochs_for_lef_activation = 0;
  return std::memcmp(&x, &y, sizeof(mgpu_params)) == 0;
}

// params has nchroms entries when per_chrom, else one. Chromosome c's cells are cell ids
// [cell_begin[c], cell_begin[c] + ncells[c]) (from 0 when cell_begin is null).
int job_create(mgpu_ctx* ctx, const mgpu_params* params, bool per_chrom, uint64_t nchroms,
               const mgpu_chrom* chroms, const mgpu_barriers* barriers, const uint64_t* nlefs,
               const uint64_t* ncells, const uint64_t* const* target_contacts,
               const uint64_t* target_epochs, int mode, mgpu_job** out,
               const uint64_t* cell_begin = nullptr) {
  if (!ctx || !out || !params) return fail(ctx, MGPU_EINVAL, "null argument");
  if (nchroms == 0 || !chroms || !barriers || !nlefs || !ncells || !target_contacts ||
      !target_epochs)
    return fail(ctx, MGPU_EINVAL, "empty or null chromosome arrays");
  if (mode != MGPU_MODE_DETERMINISTIC && mode != MGPU_MODE_FAST)
    return fail(ctx, MGPU_EINVAL, "unknown mode %d", mode);
  const mgpu_params* p = params;  // job-wide fields
  for (uint64_t c = 0; c < (per_chrom ? nchroms : 1); ++c) {
    if (int rc = check_params(&params[c]))
      return fail(ctx, rc, "parameters of chromosome %llu: %s", (unsigned long long)c,
                  g_err.c_str());
    if (per_chrom && c > 0) {
      if (mode == MGPU_MODE_DETERMINISTIC ? std::memcmp(&params[c], p, sizeof(mgpu_params)) != 0
                                          : !same_job_wide(params[c], *p))
        return fail(ctx, MGPU_EINVAL,
                    "chromosome %llu: parameter blocks of one job may differ only in LEF "
                    "density and release probabilities (and not at all in deterministic mode)",
                    (unsigned long long)c);
    }
  }
  auto par = [&](uint64_t c) -> const mgpu_params& { return per_chrom ? params[c] : *p; };
  HIPCHK(ctx, hipSetDevice(ctx->device));

  auto* job = new mgpu_job();
  job->ctx = ctx;
  job->mode = mode;
  job->p = *p;
  job->nchroms = nchroms;

  // Inputs (barrier arrays, q_n tables, chromosome descriptors, work items) are laid out first
  // in one pool and uploaded with a single copy; counters, statistics and bands follow.
  PoolBuilder pool;
  std::vector<uint64_t> off_pos(nchroms), off_dir(nchroms), off_sa(nchroms), off_si(nchroms),
      off_qn(nchroms), cpe_of(nchroms);
  uint64_t max_lefs = 0, max_bar = 0;
  const double p_loop = 1.0 / (p->tad_to_loop_ratio + 1.0);
  for (uint64_t c = 0; c < nchroms; ++c) {
    if (int rc = check_chrom(&chroms[c])) {
      delete job;
      return fail(ctx, rc, "%s", g_err.c_str());
    }
    if (chroms[c].end > 0xFFFFFFF0ULL) {
      delete job;
      return fail(ctx, MGPU_ERANGE, "chromosome end %llu beyond the 32-bit position range",
                  (unsigned long long)chroms[c].end);
    }
    if (mode == MGPU_MODE_FAST && chroms[c].end > kFastMaxEnd) {
      delete job;
      return fail(ctx, MGPU_ERANGE,
                  "chromosome end %llu: fast mode takes chromosome ends up to %llu bp (its "
                  "per-epoch move arithmetic is 32-bit); use deterministic mode",
                  (unsigned long long)chroms[c].end, (unsigned long long)kFastMaxEnd);
    }
    if (nlefs[c] == 0 || nlefs[c] > (1u << 24) || ncells[c] == 0 || ncells[c] > 0x7FFFFFFFull) {
      delete job;
      return fail(ctx, MGPU_EINVAL, "chromosome %llu: invalid LEF or cell count",
                  (unsigned long long)c);
    }
    const uint64_t nb = barriers[c].n;
    if (nb && (!barriers[c].pos || !barriers[c].dir || !barriers[c].stp_active ||
               !barriers[c].stp_inactive)) {
      delete job;
      return fail(ctx, MGPU_EINVAL, "chromosome %llu: null barrier arrays", (unsigned long long)c);
    }
    for (uint64_t i = 1; i < nb; ++i) {
      if (barriers[c].pos[i] < barriers[c].pos[i - 1]) {
        delete job;
        return fail(ctx, MGPU_EINVAL, "chromosome %llu: barriers are not sorted by position",
                    (unsigned long long)c);
      }
    }
    for (uint64_t i = 0; i < nb; ++i) {
      if (barriers[c].pos[i] < chroms[c].start || barriers[c].pos[i] >= chroms[c].end ||
          (barriers[c].dir[i] != 1 && barriers[c].dir[i] != 2)) {
        delete job;
        return fail(ctx, MGPU_EINVAL, "chromosome %llu: barrier %llu outside the range or with "
                    "an invalid direction", (unsigned long long)c, (unsigned long long)i);
      }
    }
    max_lefs = std::max(max_lefs, nlefs[c]);
    max_bar = std::max(max_bar, nb);
    off_pos[c] = pool.add(std::max<uint64_t>(nb, 1) * 8);
    off_dir[c] = pool.add(std::max<uint64_t>(nb, 1));
    off_sa[c] = pool.add(std::max<uint64_t>(nb, 1) * 8);
    off_si[c] = pool.add(std::max<uint64_t>(nb, 1) * 8);
    cpe_of[c] = contacts_per_epoch(&par(c), nlefs[c]);
    off_qn[c] = pool.add((cpe_of[c] + 1) * 8);
    job->nrows.push_back(0);
    job->items_per_chrom.push_back(ncells[c]);
    job->nitems += ncells[c];
  }
  if (job->nitems > 0x7FFFFFFFull) {
    delete job;
    return fail(ctx, MGPU_EINVAL, "more than 2^31 - 1 cells in one job");
  }
  const uint64_t off_chroms = pool.add(nchroms * sizeof(DetChrom));
  const uint64_t off_items = pool.add(job->nitems * sizeof(WorkItem));
  const uint64_t off_order = pool.add(job->nitems * 4);
  const uint64_t inputs_end = pool.total;
  // bands
  for (uint64_t c = 0; c < nchroms; ++c) {
    const uint64_t sim = chroms[c].end - chroms[c].start;
    const uint64_t nc = ceil_div(sim, p->bin_size);
    const uint64_t nr = std::min(ceil_div(p->diagonal_width, p->bin_size), nc);
    job->nrows[c] = nr;
    job->band_off.push_back(job->band_total);
    job->band_len.push_back(nr * nc + 1);
    job->band_total += ((nr * nc + 1) + 63) & ~63ULL;
  }
  const uint64_t off_missed = pool.add(nchroms * 8);
  const uint64_t off_dropped = pool.add(nchroms * 16);
  const uint64_t off_next = pool.add(64);
  const uint64_t off_err = pool.add(64);
  const uint64_t off_stats = pool.add(job->nitems * 4 * 8);
  const uint64_t off_bands = pool.add(job->band_total * 4);
  // 1D LEF occupancy: ceil(size / bin) u64 bins per chromosome (Chromosome::
  // allocate_lef_occupancy_buffer, genome.cpp:236-245)
  uint64_t off_occ = 0;
  if (p->track_1d_lef_position) {
    for (uint64_t c = 0; c < nchroms; ++c) {
      job->occ_off.push_back(job->occ_total);
      job->occ_len.push_back(ceil_div(chroms[c].size, p->bin_size));
      job->occ_total += (job->occ_len.back() + 31) & ~31ULL;
    }
    off_occ = pool.add(job->occ_total * 8);
  }

  // deterministic-mode working set per block
  if (mode == MGPU_MODE_DETERMINISTIC) {
    job->stride = det_cell_bytes_host(max_lefs, max_bar, p->burnin_history_length);
    job->use_lds = job->stride <= 64 * 1024;
    const uint64_t per_cu = job->use_lds ? std::max<uint64_t>(1, (160 * 1024) / job->stride) : 4;
    job->grid = (uint32_t)std::min<uint64_t>(job->nitems,
                                             std::min<uint64_t>(per_cu, 8) * ctx->num_cus);
    // the wave-parallel kernel keeps a cell's whole working set and its engine-word ring in
    // LDS when the largest cell fits a CU's 160 KiB; otherwise the ring stays in LDS and each
    // cell's arrays take a slot of a device-memory slab (at most kWaveSlabBytes, so a job with
    // more cells than slots runs as several launches)
    const uint64_t wb = det_wave_bytes_host(max_lefs, max_bar, p->burnin_history_length);
    if (wb <= 160 * 1024) {
      job->wave_bytes = wb;
    } else {
      job->wave_bytes = det_wave_fixed_bytes_host();
      job->wave_stride = det_wave_cell_bytes_host(max_lefs, max_bar, p->burnin_history_length);
      job->wave_slots = std::max<uint64_t>(1, std::min<uint64_t>(job->nitems,
                                                               kWaveSlabBytes / job->wave_stride));
      if (ctx->det_slab_slots) job->wave_slots = std::min(job->wave_slots, ctx->det_slab_slots);
    }
  }
  // the one-lane kernel's device-memory working sets and the wave kernel's slab are used by
  // different runs of the job, never by one, so they share a span
  uint64_t off_scratch = 0;
  const uint64_t lane_scratch =
      mode == MGPU_MODE_DETERMINISTIC && !job->use_lds ? job->stride * job->grid : 0;
  const uint64_t slab_bytes = job->wave_stride * job->wave_slots;
  if (lane_scratch || slab_bytes) off_scratch = pool.add(std::max(lane_scratch, slab_bytes));

  hipError_t e = hipMalloc(&job->d_pool, pool.total);
  if (e != hipSuccess) {
    delete job;
    return fail(ctx, MGPU_ENOMEM, "device allocation of %llu bytes failed: %s",
                (unsigned long long)pool.total, hipGetErrorString(e));
  }
  uint8_t* base = static_cast<uint8_t*>(job->d_pool);
  job->d_missed = reinterpret_cast<unsigned long long*>(base + off_missed);
  job->d_dropped = reinterpret_cast<unsigned long long*>(base + off_dropped);
  job->d_next = reinterpret_cast<unsigned long long*>(base + off_next);
  job->d_err = reinterpret_cast<unsigned long long*>(base + off_err);
  job->d_chroms = reinterpret_cast<DetChrom*>(base + off_chroms);
  job->d_items = reinterpret_cast<WorkItem*>(base + off_items);
  job->d_det_order = reinterpret_cast<uint32_t*>(base + off_order);
  job->d_cell_stats = reinterpret_cast<uint64_t*>(base + off_stats);
  job->d_bands = reinterpret_cast<uint32_t*>(base + off_bands);
  if (p->track_1d_lef_position) job->d_occ = reinterpret_cast<unsigned long long*>(base + off_occ);
  if (lane_scratch) job->d_scratch = base + off_scratch;
  if (slab_bytes) job->d_wave_slab = base + off_scratch;

  std::vector<uint8_t> stage(inputs_end, 0);
  std::vector<WorkItem> items;
  std::vector<double> release(2 * nchroms);
  items.reserve(job->nitems);
  for (uint64_t c = 0; c < nchroms; ++c) {
    const mgpu_params& pc = par(c);
    const uint64_t nb = barriers[c].n;
    if (nb) {
      std::memcpy(stage.data() + off_pos[c], barriers[c].pos, nb * 8);
      std::memcpy(stage.data() + off_dir[c], barriers[c].dir, nb);
      std::memcpy(stage.data() + off_sa[c], barriers[c].stp_active, nb * 8);
      std::memcpy(stage.data() + off_si[c], barriers[c].stp_inactive, nb * 8);
    }
    // binomial inversion's q_n = (1 - p)^t, evaluated on the host libm as the reference does
    const double pr = p_loop > 0.5 ? 1 - p_loop : p_loop;
    double* qn = reinterpret_cast<double*>(stage.data() + off_qn[c]);
    for (uint64_t t = 0; t <= cpe_of[c]; ++t) qn[t] = std::pow(1 - pr, (double)t);
    DetChrom ch{};
    ch.start = chroms[c].start;
    ch.end = chroms[c].end;
    ch.nlefs = nlefs[c];
    ch.nbar = nb;
    ch.bar_pos = reinterpret_cast<const uint64_t*>(base + off_pos[c]);
    ch.bar_dir = reinterpret_cast<const uint8_t*>(base + off_dir[c]);
    ch.bar_stp_active = reinterpret_cast<const double*>(base + off_sa[c]);
    ch.bar_stp_inactive = reinterpret_cast<const double*>(base + off_si[c]);
    ch.band = job->d_bands + job->band_off[c];
    ch.missed = job->d_missed + c;
    ch.dropped = job->d_dropped + 2 * c;
    ch.nrows = job->nrows[c];
    ch.cpe = cpe_of[c];
    ch.target_epochs = target_epochs[c];
    ch.bind_rate = (double)nlefs[c] / (double)pc.burnin_target_epochs_for_lef_activation;
    ch.qn = reinterpret_cast<const double*>(base + off_qn[c]);
    ch.occupancy = job->d_occ ? job->d_occ + job->occ_off[c] : nullptr;
    job->chroms.push_back(ch);
    release[2 * c] = pc.prob_release;
    release[2 * c + 1] = pc.prob_release_burnin;
    const uint64_t cb = cell_begin ? cell_begin[c] : 0;
    for (uint64_t k = 0; k < ncells[c]; ++k) {
      WorkItem w{};
      w.chrom = (uint32_t)c;
      w.cell_id = cb + k;
      w.seed = mgpu_cell_seed(&chroms[c], cb + k, p->seed);
      w.target = target_contacts[c][k];
      items.push_back(w);
      job->item_chrom.push_back((uint32_t)c);
    }
  }
  std::memcpy(stage.data() + off_chroms, job->chroms.data(), nchroms * sizeof(DetChrom));
  std::memcpy(stage.data() + off_items, items.data(), items.size() * sizeof(WorkItem));
  {
    // A wave-parallel cell's time grows with its LEFs times its epochs (burn-in plus the epochs
    // its contacts take at the chromosome's contacts per epoch); blocks are dispatched in order,
    // so the costliest cells go first and the last ones to start are short.
    std::vector<double> cost(items.size());
    for (uint64_t i = 0; i < items.size(); ++i) {
      const uint32_t c = items[i].chrom;
      cost[i] = (double)nlefs[c] * ((double)items[i].target / (double)std::max<uint64_t>(1, cpe_of[c]) +
                                    (double)p->burnin_history_length);
    }
    std::vector<uint32_t> order(items.size());
    for (uint64_t i = 0; i < order.size(); ++i) order[i] = (uint32_t)i;
    std::stable_sort(order.begin(), order.end(),
                     [&](uint32_t a, uint32_t b) { return cost[a] > cost[b]; });
    std::memcpy(stage.data() + off_order, order.data(), order.size() * 4);
  }
  e = hipMemcpy(base, stage.data(), stage.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) {
    (void)hipFree(job->d_pool);
    delete job;
    return fail(ctx, MGPU_EDEVICE, "upload failed: %s", hipGetErrorString(e));
  }
  if (mode == MGPU_MODE_FAST) {
    int rc = fast_job_create(*p, job->chroms, barriers, release, items, &job->fast);
    if (rc != 0) {
      (void)hipFree(job->d_pool);
      delete job;
      return fail(ctx, MGPU_EDEVICE, "fast-mode setup failed: %s", fast_last_error());
    }
  }
  if (p->log_model_internal_state) {
    // about twice the epochs a cell is expected to need: its sampling epochs plus a burn-in of
    // LEF activation, the history window and the minimum burn-in
    uint64_t want = 0;
    for (uint64_t c = 0; c < nchroms; ++c) {
      const mgpu_params& pc = par(c);
      const uint64_t burn = pc.burnin_target_epochs_for_lef_activation + pc.burnin_history_length +
                            std::min<uint64_t>(pc.min_burnin_epochs, 1u << 20);
      for (uint64_t k = 0; k < ncells[c]; ++k) {
        const uint64_t sampling = pc.target_contact_density >= 0
                                      ? ceil_div(target_contacts[c][k], std::max<uint64_t>(1, cpe_of[c]))
                                      : target_epochs[c];
        want += 2 * (sampling + burn) + 16;
      }
    }
    job->log_default = want;
    if (int rc = mgpu_job_set_state_log(job, want)) {
      fast_job_destroy(&job->fast);
      (void)hipFree(job->d_pool);
      delete job;
      return rc;
    }
  }
  *out = job;
  return MGPU_OK;
}

}  // namespace

extern "C" {

int mgpu_job_create(mgpu_ctx* ctx, const mgpu_params* p, uint64_t nchroms, const mgpu_chrom* chroms,
                    const mgpu_barriers* barriers, const uint64_t* nlefs, const uint64_t* ncells,
                    const uint64_t* const* target_contacts, const uint64_t* target_epochs, int mode,
                    mgpu_job** out) {
  return job_create(ctx, p, false, nchroms, chroms, barriers, nlefs, ncells, target_contacts,
                    target_epochs, mode, out);
}

int mgpu_job_create_windows(mgpu_ctx* ctx, const mgpu_params* p, uint64_t ntasks,
                            const mgpu_window_task* tasks, int mode, mgpu_job** out) {
  if (!ctx || !p || !out || (ntasks && !tasks)) return fail(ctx, MGPU_EINVAL, "null argument");
  if (ntasks == 0) return fail(ctx, MGPU_EINVAL, "no window tasks");
  std::vector<mgpu_chrom> chroms(ntasks);
  std::vector<mgpu_barriers> bars(ntasks);
  std::vector<uint64_t> nlefs(ntasks), ncells(ntasks, 1), cell(ntasks), epochs(ntasks);
  std::vector<const uint64_t*> targets(ntasks);
  for (uint64_t t = 0; t < ntasks; ++t) {
    chroms[t] = tasks[t].chrom;
    bars[t] = tasks[t].barriers;
    nlefs[t] = tasks[t].nlefs;
    cell[t] = tasks[t].cell_id;
    epochs[t] = tasks[t].target_epochs;
    targets[t] = &tasks[t].target_contacts;
  }
  return job_create(ctx, p, false, ntasks, chroms.data(), bars.data(), nlefs.data(),
                    ncells.data(), targets.data(), epochs.data(), mode, out, cell.data());
}

int mgpu_apply_deletion(const uint64_t* pos, uint64_t n, uint64_t deletion_begin,
                        uint64_t deletion_size, double* stp_active, double* stp_inactive,
                        uint64_t* num_active) {
  if ((n && (!pos || !stp_active || !stp_inactive)) || !num_active)
    return fail(nullptr, MGPU_EINVAL, "null argument");
  uint64_t active = n;
  for (uint64_t i = 0; i < n; ++i) {
    if (pos[i] >= deletion_begin && pos[i] < deletion_begin + deletion_size) {
      stp_active[i] = 0.0;
      stp_inactive[i] = 1.0;
      --active;
    }
  }
  *num_active = active;
  return MGPU_OK;
}

namespace {
// unsafe_get_block of n queries on the bands of tab (device pointers): queries are checked on
// the host, run on the context's stream and copied back.
int run_block_sums(mgpu_ctx* ctx, const std::vector<BlockBand>& tab, uint64_t n,
                   const uint64_t* chrom, const uint64_t* row, const uint64_t* col,
                   uint64_t block_size, uint32_t* out) {
  if (block_size == 0 || block_size % 2 == 0 || block_size > 0xFFFFFFFFull)
    return fail(ctx, MGPU_EINVAL, "block size must be odd");
  for (uint64_t i = 0; i < n; ++i) {
    const uint64_t c = chrom ? chrom[i] : 0;
    if (c >= tab.size())
      return fail(ctx, MGPU_EINVAL, "query %llu: no band %llu", (unsigned long long)i,
                  (unsigned long long)c);
    const uint64_t nc = tab[c].ncols;
    if (row[i] >= nc || col[i] >= nc)
      return fail(ctx, MGPU_EINVAL, "query %llu: pixel (%llu, %llu) outside the %llu bins",
                  (unsigned long long)i, (unsigned long long)row[i], (unsigned long long)col[i],
                  (unsigned long long)nc);
  }
  if (n == 0) return MGPU_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t tab_bytes = (tab.size() * sizeof(BlockBand) + 255) & ~255ULL;
  const uint64_t q_bytes = (n * 8 + 255) & ~255ULL;
  uint8_t* d = nullptr;
  hipError_t e = hipMalloc(&d, tab_bytes + 3 * q_bytes + n * 4);
  if (e != hipSuccess)
    return fail(ctx, MGPU_ENOMEM, "device allocation failed: %s", hipGetErrorString(e));
  auto* d_tab = reinterpret_cast<BlockBand*>(d);
  auto* d_chrom = reinterpret_cast<uint64_t*>(d + tab_bytes);
  auto* d_row = reinterpret_cast<uint64_t*>(d + tab_bytes + q_bytes);
  auto* d_col = reinterpret_cast<uint64_t*>(d + tab_bytes + 2 * q_bytes);
  auto* d_out = reinterpret_cast<uint32_t*>(d + tab_bytes + 3 * q_bytes);
  std::vector<uint64_t> zeros;
  if (!chrom) zeros.assign(n, 0);
  e = hipMemcpyAsync(d_tab, tab.data(), tab.size() * sizeof(BlockBand), hipMemcpyHostToDevice,
                     ctx->stream);
  if (e == hipSuccess)
    e = hipMemcpyAsync(d_chrom, chrom ? chrom : zeros.data(), n * 8, hipMemcpyHostToDevice,
                       ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_row, row, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess) e = hipMemcpyAsync(d_col, col, n * 8, hipMemcpyHostToDevice, ctx->stream);
  if (e == hipSuccess)
    e = launch_block_sums(ctx->stream, d_tab, d_chrom, d_row, d_col, n, (uint32_t)block_size,
                          d_out);
  if (e == hipSuccess) e = hipMemcpyAsync(out, d_out, n * 4, hipMemcpyDeviceToHost, ctx->stream);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "block sums failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}
}  // namespace

int mgpu_job_block_sums(mgpu_job* job, uint64_t n, const uint64_t* chrom, const uint64_t* row,
                        const uint64_t* col, uint64_t block_size, uint32_t* out) {
  if (!job || (n && (!chrom || !row || !col || !out)))
    return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  std::vector<BlockBand> tab(job->nchroms);
  for (uint64_t c = 0; c < job->nchroms; ++c) {
    tab[c].band = job->d_bands + job->band_off[c];
    tab[c].nrows = job->nrows[c];
    tab[c].ncols = (job->band_len[c] - 1) / job->nrows[c];
  }
  return run_block_sums(job->ctx, tab, n, chrom, row, col, block_size, out);
}

int mgpu_band_block_sums(mgpu_ctx* ctx, const uint32_t* band, uint64_t nrows, uint64_t ncols,
                         uint64_t n, const uint64_t* row, const uint64_t* col,
                         uint64_t block_size, uint32_t* out) {
  if (!ctx || !band || (n && (!row || !col || !out)))
    return fail(ctx, MGPU_EINVAL, "null argument");
  if (nrows == 0 || ncols == 0 || nrows > ncols)
    return fail(ctx, MGPU_EINVAL, "band shape %llu x %llu", (unsigned long long)nrows,
                (unsigned long long)ncols);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t belems = nrows * ncols + 1;
  uint32_t* d_band = nullptr;
  HIPCHK(ctx, hipMalloc(&d_band, belems * 4));
  hipError_t e = hipMemcpy(d_band, band, belems * 4, hipMemcpyHostToDevice);
  int rc = MGPU_OK;
  if (e != hipSuccess) {
    rc = fail(ctx, MGPU_EDEVICE, "band upload failed: %s", hipGetErrorString(e));
  } else {
    const std::vector<BlockBand> tab{BlockBand{d_band, nrows, ncols}};
    rc = run_block_sums(ctx, tab, n, nullptr, row, col, block_size, out);
  }
  (void)hipFree(d_band);
  return rc;
}

int mgpu_job_create_multi(mgpu_ctx* ctx, const mgpu_params* params, uint64_t nchroms,
                          const mgpu_chrom* chroms, const mgpu_barriers* barriers,
                          const uint64_t* nlefs, const uint64_t* ncells,
                          const uint64_t* const* target_contacts, const uint64_t* target_epochs,
                          int mode, mgpu_job** out) {
  return job_create(ctx, params, true, nchroms, chroms, barriers, nlefs, ncells, target_contacts,
                    target_epochs, mode, out);
}

}  // extern "C"

namespace {

// The device work of one run: clear the bands, missed counts and queue counter (with a kernel,
// so the sequence can be captured into a graph), then the simulation launch.
hipError_t enqueue_clears(mgpu_job* job, hipStream_t s) {
  hipError_t e = launch_zero_words(s, job->d_bands, job->band_total * 4);
  if (e == hipSuccess) e = launch_zero_words(s, job->d_missed, job->nchroms * 8);
  if (e == hipSuccess) e = launch_zero_words(s, job->d_dropped, job->nchroms * 16);
  if (e == hipSuccess) e = launch_zero_words(s, job->d_next, 64);
  if (e == hipSuccess) e = launch_zero_words(s, job->d_err, 64);
  if (e == hipSuccess && job->d_occ) e = launch_zero_words(s, job->d_occ, job->occ_total * 8);
  if (e == hipSuccess && job->log.rec) e = launch_zero_words(s, job->log.count, 8);
  return e;
}

// Whether a deterministic job's cells run on one wavefront each (det_wave.h), as they do unless
// the context asks for the one-lane kernel.
bool det_wave_fits(const mgpu_job* job) {
  return job->mode == MGPU_MODE_DETERMINISTIC && job->wave_bytes != 0 && !job->ctx->det_serial;
}

bool det_uses_wave(const mgpu_job* job) { return det_wave_fits(job); }

hipError_t enqueue_simulation(mgpu_job* job, hipStream_t s, hipEvent_t* ev = nullptr,
                              uint32_t nev = 0) {
  if (job->mode == MGPU_MODE_DETERMINISTIC) {
    DetParams P = det_params(&job->p);
    P.log = job->log;
    if (job->ctx->phase_timers && det_uses_wave(job)) P.phase = det_wave_phase_ptr();
    hipError_t e = ev && nev >= 2 ? hipEventRecord(ev[0], s) : hipSuccess;
    if (e == hipSuccess && det_uses_wave(job) && job->wave_stride) {
      for (uint64_t first = 0; e == hipSuccess && first < job->nitems; first += job->wave_slots)
        e = launch_det_wave_slab(s, P, job->d_chroms, job->d_items, job->d_det_order + first,
                                 std::min(job->wave_slots, job->nitems - first), job->d_wave_slab,
                                 job->wave_stride, job->d_cell_stats, job->d_err);
    } else if (e == hipSuccess && det_uses_wave(job)) {
      e = launch_det_wave(s, P, job->d_chroms, job->d_items, job->d_det_order, job->nitems,
                          job->wave_bytes, job->d_cell_stats, job->d_err);
    } else if (e == hipSuccess)
      e = launch_det_cells(s, P, job->d_chroms, job->d_items, job->nitems, job->d_next,
                           job->d_scratch, job->stride, job->use_lds, job->grid,
                           job->d_cell_stats);
    if (e == hipSuccess && ev && nev >= 2) e = hipEventRecord(ev[1], s);
    return e;
  }
  return fast_job_run(s, &job->fast, job->d_items, job->d_cell_stats, job->ctx->phase_timers, ev,
                      nev);
}

void drop_graph(mgpu_job* job) {
  if (job->graph_exec) (void)hipGraphExecDestroy(job->graph_exec);
  if (job->graph) (void)hipGraphDestroy(job->graph);
  job->graph_exec = nullptr;
  job->graph = nullptr;
}

}  // namespace

extern "C" {

int mgpu_job_capture(mgpu_job* job) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  drop_graph(job);
  if (job->mode == MGPU_MODE_FAST) HIPCHK(ctx, fast_job_prepare(&job->fast));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  HIPCHK(ctx, hipStreamBeginCapture(ctx->stream, hipStreamCaptureModeThreadLocal));
  hipError_t e = enqueue_clears(job, ctx->stream);
  if (e == hipSuccess) e = enqueue_simulation(job, ctx->stream);
  hipGraph_t g = nullptr;
  const hipError_t e2 = hipStreamEndCapture(ctx->stream, &g);
  if (e == hipSuccess) e = e2;
  if (e == hipSuccess) e = hipGraphInstantiate(&job->graph_exec, g, nullptr, nullptr, 0);
  job->graph = g;
  if (e != hipSuccess) {
    drop_graph(job);
    return fail(ctx, MGPU_EDEVICE, "graph capture failed: %s", hipGetErrorString(e));
  }
  return MGPU_OK;
}

int mgpu_job_release_graph(mgpu_job* job) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  drop_graph(job);
  return MGPU_OK;
}

int mgpu_job_run(mgpu_job* job, mgpu_stats* stats) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const auto t0 = std::chrono::steady_clock::now();
  uint32_t nlaunch = 0;
  job->launch_ms.clear();
  if (job->graph_exec) {
    // the captured clears and launch; the events bracket the whole graph
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(ctx, hipGraphLaunch(job->graph_exec, ctx->stream));
  } else {
    nlaunch = job->mode == MGPU_MODE_FAST
                  ? fast_job_tiers(&job->fast, nullptr, nullptr, nullptr, 5)
                  : 1;
    HIPCHK(ctx, enqueue_clears(job, ctx->stream));
    HIPCHK(ctx, hipEventRecord(ctx->ev0, ctx->stream));
    HIPCHK(ctx, enqueue_simulation(job, ctx->stream, ctx->lev, 10));
  }
  HIPCHK(ctx, hipEventRecord(ctx->ev1, ctx->stream));
  HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
  const auto t1 = std::chrono::steady_clock::now();
  if (det_uses_wave(job)) {
    unsigned long long nerr = 0;
    HIPCHK(ctx, hipMemcpy(&nerr, job->d_err, 8, hipMemcpyDeviceToHost));
    if (nerr)
      return fail(ctx, MGPU_EDEVICE, "deterministic mode: %llu cells exceeded an internal bound "
                  "of the wave-parallel kernel (engine words beyond the ring's reach or a "
                  "contact loop that never completes); rerun with mgpu_debug_det_kernel(ctx, 1)",
                  nerr);
  }
  for (uint32_t k = 0; k < nlaunch; ++k) {
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->lev[2 * k], ctx->lev[2 * k + 1]));
    job->launch_ms.push_back(ms);
  }
  if (stats) {
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    std::vector<uint64_t> cs(job->nitems * 4);
    HIPCHK(ctx, hipMemcpy(cs.data(), job->d_cell_stats, cs.size() * 8, hipMemcpyDeviceToHost));
    std::vector<unsigned long long> missed(job->nchroms), dropped(2 * job->nchroms);
    HIPCHK(ctx, hipMemcpy(missed.data(), job->d_missed, job->nchroms * 8, hipMemcpyDeviceToHost));
    HIPCHK(ctx, hipMemcpy(dropped.data(), job->d_dropped, job->nchroms * 16,
                          hipMemcpyDeviceToHost));
    *stats = mgpu_stats{};
    stats->cells = job->nitems;
    for (uint64_t i = 0; i < job->nitems; ++i) {
      stats->epochs += cs[4 * i];
      stats->burnin_epochs += cs[4 * i + 1];
      stats->contacts += cs[4 * i + 2];
    }
    for (auto m : missed) stats->missed += m;
    for (uint64_t c = 0; c < job->nchroms; ++c) {
      stats->dropped += dropped[2 * c];
      stats->dropped_occupancy += dropped[2 * c + 1];
    }
    stats->kernel_ms = ms;
    stats->total_ms = std::chrono::duration<double, std::milli>(t1 - t0).count();
  }
  return MGPU_OK;
}

int mgpu_job_set_tuning(mgpu_job* job, const mgpu_tuning* t) {
  if (!job || !t) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  drop_graph(job);  // a captured run keeps the launch shape it was recorded with
  if (job->mode != MGPU_MODE_FAST) return MGPU_OK;  // deterministic mode: one lane per cell
  if (fast_job_set_threads(&job->fast, t->threads_per_cell) != 0)
    return fail(job->ctx, MGPU_EINVAL, "%s", fast_last_error());
  if (t->threads_per_cell == 0) {
    const uint32_t n = fast_job_tiers(&job->fast, nullptr, nullptr, nullptr, 5);
    for (uint32_t k = 0; k < n && k < 5; ++k)
      if (t->tier_threads[k] && fast_job_set_tier_threads(&job->fast, k, t->tier_threads[k]) != 0)
        return fail(job->ctx, MGPU_EINVAL, "%s", fast_last_error());
  }
  if (t->global_working_set != (uint32_t)fast_job_global_storage(&job->fast) &&
      fast_job_set_storage(&job->fast, t->global_working_set ? 1 : 0) != 0)
    return fail(job->ctx, MGPU_EINVAL, "%s", fast_last_error());
  return MGPU_OK;
}

int mgpu_job_autotune(mgpu_job* job, uint64_t sample_cells, mgpu_tuning* chosen, float* ms) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  drop_graph(job);
  if (job->mode == MGPU_MODE_FAST) {
    if (fast_job_autotune(ctx->stream, &job->fast, job->d_items, job->d_cell_stats, sample_cells,
                          ms) != 0)
      return fail(ctx, MGPU_EDEVICE, "%s", fast_last_error());
  }
  return chosen ? mgpu_job_get_tuning(job, chosen) : MGPU_OK;
}

int mgpu_job_get_tuning(mgpu_job* job, mgpu_tuning* t) {
  if (!job || !t) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  *t = mgpu_tuning{};
  if (job->mode == MGPU_MODE_FAST) {
    t->threads_per_cell = fast_job_threads(&job->fast);
    t->grid = fast_job_grid(&job->fast);
    t->global_working_set = (uint32_t)fast_job_global_storage(&job->fast);
    t->launches = fast_job_tiers(&job->fast, t->tier_threads, t->tier_grid, nullptr, 5);
  } else if (det_uses_wave(job)) {
    // one block per cell; with the cells' arrays in the slab, launches of at most wave_slots
    t->threads_per_cell = 64;
    t->grid = (uint32_t)(job->wave_stride ? job->wave_slots : job->nitems);
    t->global_working_set = job->wave_stride ? 1u : 0u;
    t->launches = (uint32_t)(job->wave_stride ? ceil_div(job->nitems, job->wave_slots) : 1);
  } else {
    t->threads_per_cell = 64;
    t->grid = job->grid;
    t->global_working_set = job->use_lds ? 0u : 1u;
    t->launches = 1;
  }
  return MGPU_OK;
}

int mgpu_job_read_band(mgpu_job* job, uint64_t c, uint32_t* out_band, uint64_t* out_missed) {
  if (!job || c >= job->nchroms) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "bad chromosome index");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (out_band)
    HIPCHK(ctx, hipMemcpy(out_band, job->d_bands + job->band_off[c], job->band_len[c] * 4,
                          hipMemcpyDeviceToHost));
  if (out_missed) {
    unsigned long long m = 0;
    HIPCHK(ctx, hipMemcpy(&m, job->d_missed + c, 8, hipMemcpyDeviceToHost));
    *out_missed = m;
  }
  return MGPU_OK;
}

int mgpu_job_launch_stats(mgpu_job* job, mgpu_launch_stats* out, uint32_t max, uint32_t* n) {
  if (!job || !n) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint32_t threads[5] = {64, 0, 0, 0, 0};
  uint32_t grid[5] = {!det_uses_wave(job) ? job->grid
                      : job->wave_stride ? (uint32_t)job->wave_slots
                                         : (uint32_t)job->nitems,
                      0, 0, 0, 0};
  uint64_t cells[5] = {job->nitems, 0, 0, 0, 0}, begin[5] = {0, 0, 0, 0, 0};
  uint32_t k = 1;
  const uint32_t* order = nullptr;
  if (job->mode == MGPU_MODE_FAST) {
    k = fast_job_tiers(&job->fast, threads, grid, cells, 5, begin);
    order = fast_job_order(&job->fast);
  }
  *n = k;
  if (job->launch_ms.size() != k)
    return fail(ctx, MGPU_EINVAL, "no uncaptured run with the current launch shape yet");
  std::vector<uint64_t> cs(job->nitems * 4);
  HIPCHK(ctx, hipMemcpy(cs.data(), job->d_cell_stats, cs.size() * 8, hipMemcpyDeviceToHost));
  for (uint32_t l = 0; l < k && l < max && out; ++l) {
    mgpu_launch_stats o{};
    o.threads = threads[l];
    o.grid = grid[l];
    o.cells = cells[l];
    for (uint64_t i = begin[l]; i < begin[l] + cells[l]; ++i) {
      const uint64_t it = order ? order[i] : i;
      o.epochs += cs[4 * it];
      o.contacts += cs[4 * it + 2];
      o.barriers += job->chroms[job->item_chrom[it]].nbar;
    }
    o.ms = job->launch_ms[l];
    out[l] = o;
  }
  return MGPU_OK;
}

int mgpu_job_read_occupancy(mgpu_job* job, uint64_t c, uint64_t* out, uint64_t* nbins) {
  if (!job || c >= job->nchroms) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "bad chromosome index");
  if (!job->d_occ)
    return fail(job->ctx, MGPU_EINVAL, "the job was created without track_1d_lef_position");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (nbins) *nbins = job->occ_len[c];
  if (out)
    HIPCHK(ctx, hipMemcpy(out, job->d_occ + job->occ_off[c], job->occ_len[c] * 8,
                          hipMemcpyDeviceToHost));
  return MGPU_OK;
}

int mgpu_job_set_state_log(mgpu_job* job, uint64_t capacity) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  mgpu_ctx* ctx = job->ctx;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  if (capacity > (1ull << 32)) return fail(ctx, MGPU_ERANGE, "state log capacity above 2^32 records");
  drop_graph(job);  // a captured run keeps the log pointers it was recorded with
  if (job->d_log) {
    HIPCHK(ctx, hipStreamSynchronize(ctx->stream));
    HIPCHK(ctx, hipFree(job->d_log));
  }
  job->d_log = nullptr;
  job->log = StateLog{};
  if (capacity) {
    const uint64_t bytes = 256 + capacity * sizeof(StateRecord);
    hipError_t e = hipMalloc(&job->d_log, bytes);
    if (e != hipSuccess)
      return fail(ctx, MGPU_ENOMEM, "state log of %llu records: %s", (unsigned long long)capacity,
                  hipGetErrorString(e));
    job->log.count = static_cast<unsigned long long*>(job->d_log);
    job->log.rec = reinterpret_cast<StateRecord*>(static_cast<uint8_t*>(job->d_log) + 256);
    job->log.cap = capacity;
    HIPCHK(ctx, hipMemset(job->log.count, 0, 8));
  }
  if (job->mode == MGPU_MODE_FAST) fast_job_set_log(&job->fast, job->log);
  return MGPU_OK;
}

int mgpu_job_read_state_log(mgpu_job* job, mgpu_state_record* out, uint64_t cap,
                            uint64_t* n_records) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  mgpu_ctx* ctx = job->ctx;
  if (!job->log.rec) return fail(ctx, MGPU_EINVAL, "the job has no state log");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  unsigned long long n = 0;
  HIPCHK(ctx, hipMemcpy(&n, job->log.count, 8, hipMemcpyDeviceToHost));
  if (n_records) *n_records = n;
  if (n > job->log.cap || n > cap)
    return fail(ctx, MGPU_ERANGE, "the state log holds %llu records: capacity %llu, output room %llu",
                (unsigned long long)n, (unsigned long long)job->log.cap, (unsigned long long)cap);
  if (!out || n == 0) return MGPU_OK;
  std::vector<StateRecord> rec(n);
  HIPCHK(ctx, hipMemcpy(rec.data(), job->log.rec, n * sizeof(StateRecord), hipMemcpyDeviceToHost));
  std::vector<WorkItem> items(job->nitems);
  HIPCHK(ctx, hipMemcpy(items.data(), job->d_items, items.size() * sizeof(WorkItem),
                        hipMemcpyDeviceToHost));
  std::sort(rec.begin(), rec.end(), [](const StateRecord& a, const StateRecord& b) {
    return a.item != b.item ? a.item < b.item : a.epoch < b.epoch;
  });
  for (uint64_t i = 0; i < n; ++i) {
    const StateRecord& r = rec[i];
    if (r.item >= job->nitems) return fail(ctx, MGPU_EDEVICE, "state log record with a bad item");
    const WorkItem& w = items[r.item];
    mgpu_state_record o{};
    o.chrom = w.chrom;
    o.cell_id = w.cell_id;
    o.epoch = r.epoch;
    o.burnin = (int32_t)r.burnin;
    o.num_active_lefs = r.nact;
    o.num_barriers = job->chroms[w.chrom].nbar;
    o.active_barriers = r.active_barriers;
    o.stalls_rev = r.stalls_rev;
    o.stalls_fwd = r.stalls_fwd;
    o.stalls_both = r.stalls_both;
    o.lef_bar = r.lef_bar;
    o.primary = r.primary;
    o.secondary = r.secondary;
    o.avg_loop_size = r.avg_loop_size;
    out[i] = o;
  }
  return MGPU_OK;
}

int mgpu_job_band_device_ptr(mgpu_job* job, uint64_t c, void** dptr, uint64_t* nelems) {
  if (!job || c >= job->nchroms) return fail(nullptr, MGPU_EINVAL, "bad chromosome index");
  if (dptr) *dptr = job->d_bands + job->band_off[c];
  if (nelems) *nelems = job->band_len[c];
  return MGPU_OK;
}

void mgpu_job_destroy(mgpu_job* job) {
  if (!job) return;
  (void)hipSetDevice(job->ctx->device);
  drop_graph(job);
  fast_job_destroy(&job->fast);
  if (job->d_pool) (void)hipFree(job->d_pool);
  if (job->d_log) (void)hipFree(job->d_log);
  delete job;
}

int mgpu_simulate_chrom(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                        const mgpu_barriers* barriers, uint64_t nlefs, uint64_t cell_begin,
                        uint64_t ncells, const uint64_t* target_contacts, uint64_t target_epochs,
                        int mode, uint32_t* out_band, uint64_t* out_missed, mgpu_stats* stats) {
  if (!ctx || !out_band) return fail(ctx, MGPU_EINVAL, "null argument");
  mgpu_job* job = nullptr;
  int rc = job_create(ctx, p, false, 1, chrom, barriers, &nlefs, &ncells, &target_contacts,
                      &target_epochs, mode, &job, &cell_begin);
  if (rc) return rc;
  rc = mgpu_job_run(job, stats);
  if (!rc) {
    std::vector<uint32_t> tmp(job->band_len[0]);
    rc = mgpu_job_read_band(job, 0, tmp.data(), out_missed);
    if (!rc) {
      for (uint64_t i = 0; i < tmp.size(); ++i) out_band[i] += tmp[i];
    }
  }
  mgpu_job_destroy(job);
  return rc;
}

int mgpu_simulate_chrom_noalloc(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                                const mgpu_barriers* barriers, uint64_t nlefs,
                                uint64_t cell_begin, uint64_t ncells,
                                const uint64_t* target_contacts, uint64_t target_epochs,
                                uint32_t* d_band, uint64_t* out_missed, mgpu_stats* stats) {
  const auto t0 = std::chrono::steady_clock::now();
  if (!ctx || !d_band || !out_missed || !barriers || !target_contacts)
    return fail(ctx, MGPU_EINVAL, "null argument");
  if (!ctx->noalloc_err.empty()) return fail(ctx, MGPU_EDEVICE, "%s", ctx->noalloc_err.c_str());
  if (int rc = check_params(p)) return fail(ctx, rc, "%s", g_err.c_str());
  if (int rc = check_chrom(chrom)) return fail(ctx, rc, "%s", g_err.c_str());
  if (p->track_1d_lef_position || p->log_model_internal_state)
    return fail(ctx, MGPU_EINVAL,
                "the allocation-free entry point keeps no 1D occupancy or state log");
  if (chrom->end > kFastMaxEnd)
    return fail(ctx, MGPU_ERANGE, "chromosome end %llu: fast mode takes ends up to %llu bp",
                (unsigned long long)chrom->end, (unsigned long long)kFastMaxEnd);
  if (nlefs == 0 || nlefs > 65535) return fail(ctx, MGPU_EINVAL, "invalid LEF count");
  if (ncells == 0 || ncells > kNoAllocMaxCells)
    return fail(ctx, MGPU_ERANGE, "%llu cells: the allocation-free entry point takes 1 to %llu",
                (unsigned long long)ncells, (unsigned long long)kNoAllocMaxCells);
  const uint64_t nb = barriers->n;
  if (nb > kNoAllocMaxBarriers)
    return fail(ctx, MGPU_ERANGE, "%llu barriers: the allocation-free entry point takes up to %llu",
                (unsigned long long)nb, (unsigned long long)kNoAllocMaxBarriers);
  if (nb && (!barriers->pos || !barriers->dir || !barriers->stp_active || !barriers->stp_inactive))
    return fail(ctx, MGPU_EINVAL, "null barrier arrays");
  for (uint64_t i = 0; i < nb; ++i) {
    if ((i && barriers->pos[i] < barriers->pos[i - 1]) || barriers->pos[i] < chrom->start ||
        barriers->pos[i] >= chrom->end || (barriers->dir[i] != 1 && barriers->dir[i] != 2))
      return fail(ctx, MGPU_EINVAL, "barrier %llu unsorted, outside the range or with an invalid "
                  "direction", (unsigned long long)i);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t sim = chrom->end - chrom->start;
  const uint64_t ncols = ceil_div(sim, p->bin_size);
  DetChrom ch{};
  ch.start = chrom->start;
  ch.end = chrom->end;
  ch.nlefs = nlefs;
  ch.nbar = nb;
  ch.band = d_band;
  ch.nrows = std::min(ceil_div(p->diagonal_width, p->bin_size), ncols);
  ch.cpe = contacts_per_epoch(p, nlefs);
  ch.target_epochs = target_epochs;
  ch.bind_rate = (double)nlefs / (double)p->burnin_target_epochs_for_lef_activation;
  std::vector<WorkItem> items(ncells);
  for (uint64_t k = 0; k < ncells; ++k) {
    items[k].chrom = 0;
    items[k].cell_id = cell_begin + k;
    items[k].seed = mgpu_cell_seed(chrom, cell_begin + k, p->seed);
    items[k].target = target_contacts[k];
  }
  uint64_t missed = 0, dropped[2] = {0, 0}, sums[3] = {0, 0, 0};
  if (fast_simulate_noalloc(ctx->stream, ctx->device, (uint32_t)ctx->num_cus, *p, ch,
                            barriers->pos, barriers->dir, barriers->stp_active,
                            barriers->stp_inactive, p->prob_release, p->prob_release_burnin,
                            items, ctx->ev0, ctx->ev1, &missed, dropped, sums) != 0)
    return fail(ctx, MGPU_EDEVICE, "%s", fast_last_error());
  *out_missed = missed;
  if (stats) {
    float ms = 0.f;
    HIPCHK(ctx, hipEventElapsedTime(&ms, ctx->ev0, ctx->ev1));
    *stats = mgpu_stats{};
    stats->cells = ncells;
    stats->epochs = sums[0];
    stats->burnin_epochs = sums[1];
    stats->contacts = sums[2];
    stats->missed = missed;
    stats->dropped = dropped[0];
    stats->dropped_occupancy = dropped[1];
    stats->kernel_ms = ms;
    stats->total_ms =
        std::chrono::duration<double, std::milli>(std::chrono::steady_clock::now() - t0).count();
  }
  return MGPU_OK;
}

// ---- debug / known-answer entry points --------------------------------------------------------

int mgpu_debug_step(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom, int op,
                    uint64_t n, uint64_t* rev, uint64_t* fwd, uint64_t* epoch,
                    uint64_t* rev_ranks, uint64_t* fwd_ranks, uint64_t* rev_moves,
                    uint64_t* fwd_moves, uint64_t* rev_coll, uint64_t* fwd_coll,
                    const mgpu_barriers* barriers, const uint8_t* barrier_state,
                    uint64_t* rng_state) {
  if (!ctx || !p || !chrom || !barriers || !rng_state) return fail(ctx, MGPU_EINVAL, "null argument");
  if (n == 0 || n > (1u << 24)) return fail(ctx, MGPU_EINVAL, "n out of range");
  if (op < 0 || op > 11) return fail(ctx, MGPU_EINVAL, "unknown op %d", op);
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t nb = barriers->n;
  const uint64_t bytes = n * 8 * 7 + n * 4 * 3 + std::max<uint64_t>(nb, 1) * (8 * 3 + 2) + 64;
  uint8_t* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, bytes + 4096));
  uint64_t* d_rev = reinterpret_cast<uint64_t*>(d);
  uint64_t* d_fwd = d_rev + n;
  uint64_t* d_ep = d_fwd + n;
  uint64_t* d_rm = d_ep + n;
  uint64_t* d_fm = d_rm + n;
  uint64_t* d_rc = d_fm + n;
  uint64_t* d_fc = d_rc + n;
  uint64_t* d_bpos = d_fc + n;
  double* d_bsa = reinterpret_cast<double*>(d_bpos + std::max<uint64_t>(nb, 1));
  double* d_bsi = d_bsa + std::max<uint64_t>(nb, 1);
  uint64_t* d_rng = reinterpret_cast<uint64_t*>(d_bsi + std::max<uint64_t>(nb, 1));
  uint32_t* d_rr = reinterpret_cast<uint32_t*>(d_rng + 4);
  uint32_t* d_fr = d_rr + n;
  uint32_t* d_tmp = d_fr + n;
  uint8_t* d_bdir = reinterpret_cast<uint8_t*>(d_tmp + n);
  uint8_t* d_bst = d_bdir + std::max<uint64_t>(nb, 1);
  std::vector<uint32_t> rr(n), fr(n);
  for (uint64_t i = 0; i < n; ++i) {
    rr[i] = (uint32_t)rev_ranks[i];
    fr[i] = (uint32_t)fwd_ranks[i];
  }
  auto up = [&](void* dst, const void* src, size_t sz) {
    return sz ? hipMemcpy(dst, src, sz, hipMemcpyHostToDevice) : hipSuccess;
  };
  hipError_t e = hipSuccess;
  if (e == hipSuccess) e = up(d_rev, rev, n * 8);
  if (e == hipSuccess) e = up(d_fwd, fwd, n * 8);
  if (e == hipSuccess) e = up(d_ep, epoch, n * 8);
  if (e == hipSuccess) e = up(d_rm, rev_moves, n * 8);
  if (e == hipSuccess) e = up(d_fm, fwd_moves, n * 8);
  if (e == hipSuccess) e = up(d_rc, rev_coll, n * 8);
  if (e == hipSuccess) e = up(d_fc, fwd_coll, n * 8);
  if (e == hipSuccess) e = up(d_rr, rr.data(), n * 4);
  if (e == hipSuccess) e = up(d_fr, fr.data(), n * 4);
  if (e == hipSuccess) e = up(d_rng, rng_state, 32);
  if (e == hipSuccess) e = up(d_bpos, barriers->pos, nb * 8);
  if (e == hipSuccess) e = up(d_bsa, barriers->stp_active, nb * 8);
  if (e == hipSuccess) e = up(d_bsi, barriers->stp_inactive, nb * 8);
  if (e == hipSuccess) e = up(d_bdir, barriers->dir, nb);
  if (e == hipSuccess && barrier_state) e = up(d_bst, barrier_state, nb);
  DetChrom ch{};
  ch.start = chrom->start;
  ch.end = chrom->end;
  ch.nlefs = n;
  ch.nbar = nb;
  ch.bar_pos = d_bpos;
  ch.bar_dir = d_bdir;
  ch.bar_stp_active = d_bsa;
  ch.bar_stp_inactive = d_bsi;
  if (e == hipSuccess)
    e = launch_det_step(ctx->stream, det_params(p), ch, op, n, d_rev, d_fwd, d_ep, d_rr, d_fr,
                        d_tmp, d_rm, d_fm, d_rc, d_fc, d_bst, d_rng);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  auto down = [&](void* dst, const void* src, size_t sz) {
    return hipMemcpy(dst, src, sz, hipMemcpyDeviceToHost);
  };
  if (e == hipSuccess) e = down(rev, d_rev, n * 8);
  if (e == hipSuccess) e = down(fwd, d_fwd, n * 8);
  if (e == hipSuccess) e = down(epoch, d_ep, n * 8);
  if (e == hipSuccess) e = down(rev_moves, d_rm, n * 8);
  if (e == hipSuccess) e = down(fwd_moves, d_fm, n * 8);
  if (e == hipSuccess) e = down(rev_coll, d_rc, n * 8);
  if (e == hipSuccess) e = down(fwd_coll, d_fc, n * 8);
  if (e == hipSuccess) e = down(rr.data(), d_rr, n * 4);
  if (e == hipSuccess) e = down(fr.data(), d_fr, n * 4);
  if (e == hipSuccess) e = down(rng_state, d_rng, 32);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "debug step failed: %s", hipGetErrorString(e));
  for (uint64_t i = 0; i < n; ++i) {
    rev_ranks[i] = rr[i];
    fwd_ranks[i] = fr[i];
  }
  return MGPU_OK;
}

int mgpu_debug_draws(mgpu_ctx* ctx, uint64_t seed, int kind, double a, double b, double c,
                     uint64_t n, uint64_t* out) {
  if (!ctx || !out) return fail(ctx, MGPU_EINVAL, "null argument");
  if (kind < 0 || kind > 8) return fail(ctx, MGPU_EINVAL, "unknown sampler kind %d", kind);
  if (n == 0) return MGPU_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint64_t* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, n * 8));
  hipError_t e = launch_det_draws(ctx->stream, seed, kind, a, b, c, n, d);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d, n * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "debug draws failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}

int mgpu_debug_cell_trace(mgpu_ctx* ctx, const mgpu_params* p, const mgpu_chrom* chrom,
                          const mgpu_barriers* barriers, uint64_t nlefs, uint64_t cell_id,
                          uint64_t target_contacts, uint64_t target_epochs, uint64_t max_trace,
                          uint64_t* trace, uint64_t* out_epochs) {
  if (!ctx || !trace || !out_epochs) return fail(ctx, MGPU_EINVAL, "null argument");
  mgpu_job* job = nullptr;
  const uint64_t one = 1;
  const uint64_t* tc = &target_contacts;
  int rc = mgpu_job_create(ctx, p, 1, chrom, barriers, &nlefs, &one, &tc, &target_epochs,
                           MGPU_MODE_DETERMINISTIC, &job);
  if (rc) return rc;
  const uint64_t bytes =
      std::max(det_cell_bytes_host(nlefs, barriers->n, p->burnin_history_length),
               det_wave_cell_bytes_host(nlefs, barriers->n, p->burnin_history_length));
  uint8_t* d = nullptr;
  hipError_t e = hipMalloc(&d, bytes + max_trace * 8 + 64);
  if (e == hipSuccess) e = hipMemsetAsync(job->d_bands, 0, job->band_total * 4, ctx->stream);
  if (e == hipSuccess) e = hipMemsetAsync(job->d_missed, 0, 8, ctx->stream);
  uint64_t* d_trace = reinterpret_cast<uint64_t*>(d + bytes);
  uint64_t* d_out = d_trace + max_trace;
  if (e == hipSuccess && det_wave_fits(job))
    e = launch_det_wave_trace(ctx->stream, det_params(p), job->chroms[0],
                              mgpu_cell_seed(chrom, cell_id, p->seed), target_contacts,
                              job->wave_bytes, job->wave_stride ? d : nullptr, d_trace, max_trace,
                              d_out);
  else if (e == hipSuccess)
    e = launch_det_trace(ctx->stream, det_params(p), job->chroms[0],
                         mgpu_cell_seed(chrom, cell_id, p->seed), target_contacts, d, d_trace,
                         max_trace, d_out);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess && max_trace) e = hipMemcpy(trace, d_trace, max_trace * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(out_epochs, d_out, 32, hipMemcpyDeviceToHost);
  if (d) (void)hipFree(d);
  mgpu_job_destroy(job);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "trace failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}

int mgpu_context_stream(mgpu_ctx* ctx, void** stream, int* device) {
  if (!ctx) return fail(ctx, MGPU_EINVAL, "null context");
  if (stream) *stream = ctx->stream;
  if (device) *device = ctx->device;
  return MGPU_OK;
}

int mgpu_debug_phase_times(mgpu_ctx* ctx, int enable, uint64_t* out, uint64_t n) {
  if (!ctx) return fail(ctx, MGPU_EINVAL, "null context");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  ctx->phase_timers = enable ? 1 : 0;
  if (out && n) {
    std::vector<uint64_t> v(kFastPhaseCount + kDetWavePhases, 0);
    HIPCHK(ctx, fast_phase_times(v.data(), kFastPhaseCount));
    HIPCHK(ctx, det_wave_phase_times(v.data() + kFastPhaseCount, kDetWavePhases));
    for (uint64_t k = 0; k < n; ++k) out[k] = k < v.size() ? v[k] : 0;
  }
  return MGPU_OK;
}

int mgpu_job_cell_stats(mgpu_job* job, uint64_t* out, uint64_t max, uint64_t* n) {
  if (!job || !n) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  mgpu_ctx* ctx = job->ctx;
  *n = job->nitems;
  if (!out || max == 0) return MGPU_OK;
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t k = std::min<uint64_t>(max, job->nitems * 4);
  HIPCHK(ctx, hipMemcpy(out, job->d_cell_stats, k * 8, hipMemcpyDeviceToHost));
  return MGPU_OK;
}

int mgpu_job_kernel(mgpu_job* job, char* name, uint64_t cap) {
  if (!job || !name || cap == 0) return fail(job ? job->ctx : nullptr, MGPU_EINVAL, "null argument");
  const char* k = job->mode == MGPU_MODE_FAST ? "fast_cells_kernel"
                  : det_uses_wave(job) && job->wave_stride ? "det_wave_slab_kernel"
                  : det_uses_wave(job)         ? "det_wave_kernel"
                                               : "det_cells_kernel";
  snprintf(name, cap, "%s", k);
  return MGPU_OK;
}

int mgpu_debug_det_slab_slots(mgpu_ctx* ctx, uint64_t slots) {
  if (!ctx) return fail(ctx, MGPU_EINVAL, "null context");
  ctx->det_slab_slots = slots;
  return MGPU_OK;
}

int mgpu_debug_det_kernel(mgpu_ctx* ctx, int serial) {
  if (!ctx) return fail(ctx, MGPU_EINVAL, "null context");
  ctx->det_serial = serial ? 1 : 0;
  return MGPU_OK;
}

int mgpu_debug_fast_serial_fixups(mgpu_job* job, int serial) {
  if (!job) return fail(nullptr, MGPU_EINVAL, "null job");
  if (job->mode != MGPU_MODE_FAST) return fail(job->ctx, MGPU_EINVAL, "not a fast-mode job");
  drop_graph(job);
  fast_job_set_serial_fixups(&job->fast, serial);
  return MGPU_OK;
}

int mgpu_debug_philox(mgpu_ctx* ctx, const uint32_t* in, uint64_t n, uint32_t* out) {
  if (!ctx || !in || !out) return fail(ctx, MGPU_EINVAL, "null argument");
  if (n == 0) return MGPU_OK;
  if (n > (1u << 24)) return fail(ctx, MGPU_EINVAL, "n out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint32_t* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, n * 40));
  hipError_t e = hipMemcpy(d, in, n * 24, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_philox_kat(ctx->stream, d, d + 6 * n, (int)n);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d + 6 * n, n * 16, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "philox failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}

int mgpu_debug_band_increments(mgpu_ctx* ctx, int mode, uint64_t nrows, uint64_t ncols,
                               uint64_t n, const uint64_t* rows, const uint64_t* cols,
                               uint32_t* band, uint64_t* missed) {
  if (!ctx || !band || !missed || (n && (!rows || !cols)))
    return fail(ctx, MGPU_EINVAL, "null argument");
  if (mode != MGPU_MODE_DETERMINISTIC && mode != MGPU_MODE_FAST)
    return fail(ctx, MGPU_EINVAL, "unknown mode %d", mode);
  if (nrows == 0 || ncols == 0 || nrows * ncols > (1ull << 28) || n > (1ull << 24))
    return fail(ctx, MGPU_EINVAL, "band or increment count out of range");
  for (uint64_t i = 0; i < n; ++i) {
    // the bounds ContactMatrixDense checks before an increment: the column within the matrix
    // (larger coordinate), the offset below nrows or counted as missed
    const uint64_t j = std::max(rows[i], cols[i]);
    if (j >= ncols || (mode == MGPU_MODE_FAST && j > 0xFFFFFFFFull))
      return fail(ctx, MGPU_ERANGE, "increment %llu at (%llu, %llu) outside the matrix",
                  (unsigned long long)i, (unsigned long long)rows[i], (unsigned long long)cols[i]);
  }
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t belems = nrows * ncols + 1;
  uint8_t* d = nullptr;
  const uint64_t off_band = 0, off_missed = ((belems * 4 + 255) & ~255ull);
  const uint64_t off_coords = off_missed + 256;
  HIPCHK(ctx, hipMalloc(&d, off_coords + 16 * std::max<uint64_t>(n, 1)));
  uint32_t* d_band = reinterpret_cast<uint32_t*>(d + off_band);
  unsigned long long* d_missed = reinterpret_cast<unsigned long long*>(d + off_missed);
  hipError_t e = hipMemcpy(d_band, band, belems * 4, hipMemcpyHostToDevice);
  unsigned long long m0 = *missed;
  if (e == hipSuccess) e = hipMemcpy(d_missed, &m0, 8, hipMemcpyHostToDevice);
  if (e == hipSuccess && n) {
    if (mode == MGPU_MODE_DETERMINISTIC) {
      uint64_t* d_rows = reinterpret_cast<uint64_t*>(d + off_coords);
      uint64_t* d_cols = d_rows + n;
      e = hipMemcpy(d_rows, rows, n * 8, hipMemcpyHostToDevice);
      if (e == hipSuccess) e = hipMemcpy(d_cols, cols, n * 8, hipMemcpyHostToDevice);
      DetChrom ch{};
      ch.band = d_band;
      ch.missed = d_missed;
      ch.nrows = nrows;
      if (e == hipSuccess) e = launch_det_band(ctx->stream, ch, d_rows, d_cols, n);
    } else {
      std::vector<uint32_t> rc(2 * n);
      for (uint64_t i = 0; i < n; ++i) {
        rc[i] = (uint32_t)rows[i];
        rc[n + i] = (uint32_t)cols[i];
      }
      uint32_t* d_rc = reinterpret_cast<uint32_t*>(d + off_coords);
      e = hipMemcpy(d_rc, rc.data(), n * 8, hipMemcpyHostToDevice);
      if (e == hipSuccess)
        e = launch_fast_band(ctx->stream, d_band, d_missed, (uint32_t)nrows, d_rc, d_rc + n, n);
    }
  }
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(band, d_band, belems * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(&m0, d_missed, 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(ctx, MGPU_EDEVICE, "band increments failed: %s", hipGetErrorString(e));
  *missed = m0;
  return MGPU_OK;
}

int mgpu_debug_collisions(mgpu_ctx* ctx, uint64_t n, const uint64_t* idx, const uint8_t* ev,
                          uint64_t* words, uint64_t* index, uint32_t* event, uint32_t* flags) {
  if (!ctx || (n && (!idx || !ev || !words || !index || !event || !flags)))
    return fail(ctx, MGPU_EINVAL, "null argument");
  if (n == 0) return MGPU_OK;
  if (n > (1u << 24)) return fail(ctx, MGPU_EINVAL, "n out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  uint8_t* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, n * 41 + 256));
  uint64_t* d_idx = reinterpret_cast<uint64_t*>(d);
  uint64_t* d_words = d_idx + n;
  uint64_t* d_index = d_words + n;
  uint32_t* d_event = reinterpret_cast<uint32_t*>(d_index + n);
  uint32_t* d_flags = d_event + n;
  uint8_t* d_ev = reinterpret_cast<uint8_t*>(d_flags + n);
  hipError_t e = hipMemcpy(d_idx, idx, n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(d_ev, ev, n, hipMemcpyHostToDevice);
  if (e == hipSuccess)
    e = launch_det_collisions(ctx->stream, d_idx, d_ev, n, d_words, d_index, d_event, d_flags);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(words, d_words, n * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(index, d_index, n * 8, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(event, d_event, n * 4, hipMemcpyDeviceToHost);
  if (e == hipSuccess) e = hipMemcpy(flags, d_flags, n * 4, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess)
    return fail(ctx, MGPU_EDEVICE, "collision words failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}

int mgpu_debug_fast_math(mgpu_ctx* ctx, int fn, const double* x, uint64_t n, double* out) {
  if (!ctx || !x || !out) return fail(ctx, MGPU_EINVAL, "null argument");
  if (fn < 0 || fn > 2) return fail(ctx, MGPU_EINVAL, "unknown function %d", fn);
  if (n == 0) return MGPU_OK;
  if (n > (1u << 24)) return fail(ctx, MGPU_EINVAL, "n out of range");
  HIPCHK(ctx, hipSetDevice(ctx->device));
  const uint64_t nout = fn == 2 ? 2 * n : n;
  double* d = nullptr;
  HIPCHK(ctx, hipMalloc(&d, (n + nout) * 8));
  hipError_t e = hipMemcpy(d, x, n * 8, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = launch_fast_math(ctx->stream, fn, d, d + n, (int)n);
  if (e == hipSuccess) e = hipStreamSynchronize(ctx->stream);
  if (e == hipSuccess) e = hipMemcpy(out, d + n, nout * 8, hipMemcpyDeviceToHost);
  (void)hipFree(d);
  if (e != hipSuccess) return fail(ctx, MGPU_EDEVICE, "fast math failed: %s", hipGetErrorString(e));
  return MGPU_OK;
}

}  // extern "C"

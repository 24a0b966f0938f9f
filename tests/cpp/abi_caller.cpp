// A C++ caller of the C-ABI, as Simulation::run_simulate would be after the integration change:
// Config -> mgpu_params through fill_params_from_config, then mgpu_init -> mgpu_prepare_barriers ->
// mgpu_plan_chrom -> mgpu_simulate_chrom for one chromosome. The band it gets back is checked
// against the oracle (tests may link it; the library never does).
//
//   abi_caller params   host-only: the parameter fill against mgpu_config_transform and the
//                       defaults' closed forms (no GPU needed)
//   abi_caller run      the above, then deterministic mode bit for bit against the oracle and fast
//                       mode's contact conservation, on the GPU
#include <cmath>
#include <cstdio>
#include <cstring>
#include <random>
#include <string>
#include <vector>

#include "modle_gpu.h"
#include "modle_gpu_config.hpp"
#include "refcpu.h"

namespace {

// The simulate fields of modle::Config (simulation_config.hpp:24-135), with the values
// Cli::transform_args leaves for `modle simulate -r 5kb -w 1Mbp --extrusion-barrier-occupancy 0.825
// --ncells 4 --seed 1`: 0.8 bin of
// extrusion per epoch and side, the 5 % standard deviation in bp, release probability
// (v_rev + v_fwd) / processivity, 5 processivities of activation at burn-in speed, barrier
// occupancy 0.825 with not-occupied self-transition 0.7 (so occupied 1 - 0.3 * 0.175 / 0.825).
struct Config {
  struct ContactSamplingStrategy {
    unsigned v;
    static constexpr unsigned noisify = 1u, tad = 2u, loop = 4u;
    friend unsigned operator&(ContactSamplingStrategy a, unsigned b) { return a.v & b; }
  };
  unsigned long long target_simulation_epochs = 2000;
  double target_contact_density = 1.0;
  unsigned long long bin_size = 5000, diagonal_width = 1000000;
  ContactSamplingStrategy contact_sampling_strategy{7};
  double tad_to_loop_contact_ratio = 5.0;
  double genextreme_mu = 0, genextreme_sigma = 5000, genextreme_xi = 0.001;
  unsigned long long fwd_extrusion_speed = 4000, rev_extrusion_speed = 4000;
  double fwd_extrusion_speed_std = 200, rev_extrusion_speed_std = 200;
  double number_of_lefs_per_mbp = 20;
  double prob_of_lef_release = 8000.0 / 300000.0, prob_of_lef_release_burnin = 8000.0 / 300000.0;
  unsigned long long contact_sampling_interval = 50000;
  bool track_1d_lef_position = false;
  bool override_extrusion_barrier_occupancy = true;  // the option was given
  double barrier_occupied_stp = 1.0 - 0.3 * (1.0 - 0.825) / 0.825;
  double barrier_not_occupied_stp = 0.7;
  double hard_stall_lef_stability_multiplier = 5.0, soft_stall_lef_stability_multiplier = 1.0;
  double probability_of_extrusion_unit_bypass = 0.1;
  double lef_bar_major_collision_pblock = 1.0, lef_bar_minor_collision_pblock = 0.0;
  unsigned long long seed = 1;
  bool skip_burnin = false;
  unsigned long long burnin_history_length = 100, burnin_smoothing_window_size = 5;
  unsigned long long min_burnin_epochs = 0, max_burnin_epochs = ~0ull;
  unsigned long long burnin_target_epochs_for_lef_activation = 5 * 300000 / 8000;
  unsigned long long fwd_extrusion_speed_burnin = 4000, rev_extrusion_speed_burnin = 4000;
  bool log_model_internal_state = false;
};

int failures = 0;

void check(bool ok, const char* what) {
  if (!ok) {
    std::printf("FAIL %s\n", what);
    ++failures;
  }
}

bool close(double a, double b) { return std::fabs(a - b) <= 1e-12 * std::fmax(1.0, std::fabs(b)); }

// The fill against the library's own transform of the same options.
void check_params(mgpu_params& p) {
  Config cfg;
  fill_params_from_config(cfg, p);
  mgpu_config c;
  check(mgpu_config_defaults(&c) == MGPU_OK, "mgpu_config_defaults");
  c.diagonal_width = 1000000;
  c.seed = 1;
  c.extrusion_barrier_occupancy = 0.825;
  mgpu_params q;
  check(mgpu_config_transform(&c, MGPU_SET_BARRIER_OCCUPANCY, &q) == MGPU_OK,
        "mgpu_config_transform");
  check(p.bin_size == q.bin_size && p.diagonal_width == q.diagonal_width, "band shape");
  check(p.rev_speed == q.rev_speed && p.fwd_speed == q.fwd_speed, "speeds");
  check(p.rev_speed_burnin == q.rev_speed_burnin && p.fwd_speed_burnin == q.fwd_speed_burnin,
        "burn-in speeds");
  check(close(p.rev_speed_std, q.rev_speed_std) && close(p.fwd_speed_std, q.fwd_speed_std), "std");
  check(close(p.prob_release, q.prob_release), "release");
  check(close(p.prob_release_burnin, q.prob_release_burnin), "release (burn-in)");
  check(close(p.p_bypass, q.p_bypass), "bypass");
  check(p.pblock_major == q.pblock_major && p.pblock_minor == q.pblock_minor, "pblock");
  check(p.hard_stall_mult == q.hard_stall_mult && p.soft_stall_mult == q.soft_stall_mult, "stall");
  check(p.tad_to_loop_ratio == q.tad_to_loop_ratio, "tad:loop");
  check(p.gev_mu == q.gev_mu && p.gev_sigma == q.gev_sigma && p.gev_xi == q.gev_xi, "gev");
  check(p.noisify == q.noisify && p.skip_burnin == q.skip_burnin, "flags");
  check(p.contact_sampling_interval == q.contact_sampling_interval, "sampling interval");
  check(p.burnin_history_length == q.burnin_history_length &&
            p.burnin_smoothing_window == q.burnin_smoothing_window,
        "burn-in history");
  check(p.min_burnin_epochs == q.min_burnin_epochs && p.max_burnin_epochs == q.max_burnin_epochs,
        "burn-in bounds");
  check(p.burnin_target_epochs_for_lef_activation == q.burnin_target_epochs_for_lef_activation,
        "activation epochs");
  check(p.target_contact_density == q.target_contact_density, "density");
  check(p.seed == q.seed && p.lefs_per_mbp == q.lefs_per_mbp, "seed, LEF density");
  check(close(p.barrier_occupied_stp, q.barrier_occupied_stp) &&
            close(p.barrier_not_occupied_stp, q.barrier_not_occupied_stp),
        "barrier stp");
  check(p.override_barrier_occupancy == q.override_barrier_occupancy, "override");
}

rc_config oracle_config(const mgpu_params& p) {
  rc_config r{};
  r.bin_size = p.bin_size;
  r.diagonal_width = p.diagonal_width;
  r.rev_speed = p.rev_speed;
  r.fwd_speed = p.fwd_speed;
  r.rev_speed_burnin = p.rev_speed_burnin;
  r.fwd_speed_burnin = p.fwd_speed_burnin;
  r.rev_speed_std = p.rev_speed_std;
  r.fwd_speed_std = p.fwd_speed_std;
  r.prob_release = p.prob_release;
  r.prob_release_burnin = p.prob_release_burnin;
  r.p_bypass = p.p_bypass;
  r.pblock_major = p.pblock_major;
  r.pblock_minor = p.pblock_minor;
  r.hard_stall_mult = p.hard_stall_mult;
  r.soft_stall_mult = p.soft_stall_mult;
  r.tad_to_loop_ratio = p.tad_to_loop_ratio;
  r.gev_mu = p.gev_mu;
  r.gev_sigma = p.gev_sigma;
  r.gev_xi = p.gev_xi;
  r.noisify = p.noisify;
  r.skip_burnin = p.skip_burnin;
  r.contact_sampling_interval = p.contact_sampling_interval;
  r.burnin_history_length = p.burnin_history_length;
  r.burnin_smoothing_window = p.burnin_smoothing_window;
  r.min_burnin_epochs = p.min_burnin_epochs;
  r.max_burnin_epochs = p.max_burnin_epochs;
  r.burnin_target_epochs_for_lef_activation = p.burnin_target_epochs_for_lef_activation;
  r.target_contact_density = p.target_contact_density;
  r.target_simulation_epochs = p.target_simulation_epochs;
  r.seed = p.seed;
  return r;
}

int run(const mgpu_params& p) {
  // one 2 Mbp chromosome with 40 barriers (BED records of 1 bp, seeded), 4 cells
  const std::string name = "chrT";
  const uint64_t size = 2000000, ncells = 4;
  std::mt19937_64 rng(7);
  std::vector<mgpu_bed_record> recs;
  for (int i = 0; i < 40; ++i) {
    const uint64_t s = rng() % (size - 1);
    recs.push_back({s, s + 1, 0.8, (i & 1) ? '+' : '-'});
  }
  std::vector<uint64_t> pos(recs.size());
  std::vector<uint8_t> dir(recs.size());
  std::vector<double> sa(recs.size()), si(recs.size());
  uint64_t nbar = 0;
  check(mgpu_prepare_barriers(&p, recs.data(), recs.size(), 0, size, pos.data(), dir.data(),
                              sa.data(), si.data(), &nbar) == MGPU_OK,
        "mgpu_prepare_barriers");
  const mgpu_barriers mb{nbar, pos.data(), dir.data(), sa.data(), si.data()};
  const mgpu_chrom mc{name.data(), name.size(), size, 0, size};
  uint64_t nlefs = 0, nrows = 0, ncols = 0, target_epochs = 0;
  std::vector<uint64_t> targets(ncells);
  check(mgpu_plan_chrom(&p, &mc, ncells, &nlefs, &nrows, &ncols, targets.data(),
                        &target_epochs) == MGPU_OK,
        "mgpu_plan_chrom");

  mgpu_ctx* ctx = nullptr;
  if (mgpu_init(0, &ctx) != MGPU_OK) {
    std::printf("FAIL mgpu_init: %s\n", mgpu_last_error(nullptr));
    return 1;
  }
  std::vector<uint32_t> band(nrows * ncols + 1, 0);
  uint64_t missed = 0;
  mgpu_stats st{};
  if (mgpu_simulate_chrom(ctx, &p, &mc, &mb, nlefs, 0, ncells, targets.data(), target_epochs,
                          MGPU_MODE_DETERMINISTIC, band.data(), &missed, &st) != MGPU_OK) {
    std::printf("FAIL mgpu_simulate_chrom: %s\n", mgpu_last_error(ctx));
    mgpu_destroy(ctx);
    return 1;
  }

  // the oracle on the same inputs
  const rc_config rc = oracle_config(p);
  const rc_chrom rch{name.data(), name.size(), size, 0, size};
  const rc_barriers rb{nbar, pos.data(), dir.data(), sa.data(), si.data()};
  std::vector<uint32_t> oband(nrows * ncols + 1, 0);
  rc_band ob{nrows, ncols, oband.data(), 0};
  std::vector<rc_cell_stats> ost(ncells);
  check(rc_simulate_cells(&rc, &rch, &rb, nlefs, 0, ncells, targets.data(), target_epochs, &ob, 4,
                          ost.data()) == 0,
        "oracle");
  check(band == oband, "deterministic band bit for bit against the oracle");
  check(missed == ob.missed, "missed updates against the oracle");
  uint64_t tot = 0, sum = 0;
  for (uint64_t t : targets) tot += t;
  for (uint32_t v : band) sum += v;
  check(sum + missed == tot, "deterministic contacts conserved");

  // fast mode: every planned contact in the band, out of it or counted as dropped
  std::fill(band.begin(), band.end(), 0u);
  missed = 0;
  check(mgpu_simulate_chrom(ctx, &p, &mc, &mb, nlefs, 0, ncells, targets.data(), target_epochs,
                            MGPU_MODE_FAST, band.data(), &missed, &st) == MGPU_OK,
        "fast mode");
  sum = 0;
  for (uint32_t v : band) sum += v;
  check(sum + missed + st.dropped == tot && st.dropped == 0, "fast contacts conserved");
  std::printf("simulated %llu cells of %s: %llu contacts, %llu missed, kernel %.1f ms (fast)\n",
              (unsigned long long)ncells, name.c_str(), (unsigned long long)sum,
              (unsigned long long)missed, st.kernel_ms);
  mgpu_destroy(ctx);
  return 0;
}

}  // namespace

int main(int argc, char** argv) {
  const bool gpu = argc > 1 && std::strcmp(argv[1], "run") == 0;
  mgpu_params p;
  check_params(p);
  if (gpu && failures == 0 && run(p) != 0) ++failures;
  std::printf("%s\n", failures ? "FAILED" : "OK");
  return failures ? 1 : 0;
}

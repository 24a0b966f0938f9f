// modle_gpu_config.hpp: fills the C-ABI's mgpu_params from a modle::Config that
// Cli::transform_args (src/modle/cli.cpp:1192-1215) has already transformed, for C++ callers such
// as Simulation::run_simulate. Header-only; the field names are modle::Config's
// (src/common/include/modle/common/simulation_config.hpp:24-135), so any type with those members
// works (the reference's Config, or a mirror of it).
#ifndef MODLE_GPU_CONFIG_HPP
#define MODLE_GPU_CONFIG_HPP

#include "modle_gpu.h"

template <class ConfigT>
void fill_params_from_config(const ConfigT& c, mgpu_params& p) {
  using CS = typename ConfigT::ContactSamplingStrategy;
  p = mgpu_params{};
  p.bin_size = c.bin_size;
  p.diagonal_width = c.diagonal_width;
  p.rev_speed = c.rev_extrusion_speed;
  p.fwd_speed = c.fwd_extrusion_speed;
  p.rev_speed_burnin = c.rev_extrusion_speed_burnin;
  p.fwd_speed_burnin = c.fwd_extrusion_speed_burnin;
  // cli_update_extr_speed has turned fractional standard deviations into base pairs
  p.rev_speed_std = c.rev_extrusion_speed_std;
  p.fwd_speed_std = c.fwd_extrusion_speed_std;
  p.prob_release = c.prob_of_lef_release;
  p.prob_release_burnin = c.prob_of_lef_release_burnin;
  // already normalised by cli_normalize_probabilities when normalize_probabilities is set
  p.p_bypass = c.probability_of_extrusion_unit_bypass;
  p.pblock_major = c.lef_bar_major_collision_pblock;
  p.pblock_minor = c.lef_bar_minor_collision_pblock;
  p.hard_stall_mult = c.hard_stall_lef_stability_multiplier;
  p.soft_stall_mult = c.soft_stall_lef_stability_multiplier;
  // 0 (loop only) or infinity (TAD only) after cli_update_tad_to_loop_contact_ratio
  p.tad_to_loop_ratio = c.tad_to_loop_contact_ratio;
  p.gev_mu = c.genextreme_mu;
  p.gev_sigma = c.genextreme_sigma;
  p.gev_xi = c.genextreme_xi;
  p.noisify = bool(c.contact_sampling_strategy & CS::noisify) ? 1 : 0;
  p.skip_burnin = c.skip_burnin ? 1 : 0;
  p.contact_sampling_interval = c.contact_sampling_interval;
  p.burnin_history_length = c.burnin_history_length;
  p.burnin_smoothing_window = c.burnin_smoothing_window_size;
  p.min_burnin_epochs = c.min_burnin_epochs;
  p.max_burnin_epochs = c.max_burnin_epochs;
  p.burnin_target_epochs_for_lef_activation = c.burnin_target_epochs_for_lef_activation;
  // -1 under the simulation-epochs stopping criterion (transform_args sets it)
  p.target_contact_density = c.target_contact_density;
  p.target_simulation_epochs = c.target_simulation_epochs;
  p.seed = c.seed;
  p.lefs_per_mbp = c.number_of_lefs_per_mbp;
  // the normalised stationary-state probabilities: barriers without a score take these
  p.barrier_occupied_stp = c.barrier_occupied_stp;
  p.barrier_not_occupied_stp = c.barrier_not_occupied_stp;
  p.override_barrier_occupancy = c.override_extrusion_barrier_occupancy ? 1 : 0;
  p.track_1d_lef_position = c.track_1d_lef_position ? 1 : 0;
  p.log_model_internal_state = c.log_model_internal_state ? 1 : 0;
}

#endif

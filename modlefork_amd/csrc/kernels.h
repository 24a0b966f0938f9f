// Launch-side declarations shared by the kernel translation units and the C-ABI layer.
#pragma once
#ifndef MGPU_EMU
#include <hip/hip_runtime.h>
#endif
#include <stdint.h>

namespace mgpu {

// One (chromosome, cell) task of run_simulate (scheduler_simulate.cpp:141-153).
struct WorkItem {
  uint32_t chrom;
  uint32_t reserved;
  uint64_t cell_id;
  uint64_t seed;    // Chromosome::hash(seed, cell_id)
  uint64_t target;  // contacts this cell must register
};

// One line of the model internal-state log (Simulation::dump_stats, simulation.cpp:1166-1224),
// as a cell writes it: item is the work-item index, counts are 32-bit (LEFs per cell < 2^31).
struct StateRecord {
  uint32_t item, epoch;
  uint32_t burnin, nact;
  uint32_t active_barriers, stalls_rev;
  uint32_t stalls_fwd, stalls_both;
  uint32_t lef_bar, primary;
  uint32_t secondary, reserved;
  double avg_loop_size;
};

// Device log of StateRecords: cells take slots from `count`; slots at or past `cap` are counted
// but not written (the host reports the shortfall). rec == nullptr: logging is off.
struct StateLog {
  StateRecord* rec;
  unsigned long long* count;
  uint64_t cap;
};

// Post-transform parameters in the form the deterministic kernels read them.
struct DetParams {
  uint64_t bin_size;
  uint64_t rev_speed, fwd_speed, rev_speed_burnin, fwd_speed_burnin;
  double rev_std, fwd_std;
  double prob_release, prob_release_burnin;
  double p_bypass, pblock_major, pblock_minor;
  double inv_soft, inv_hard;  // 1/soft and 1/hard stall multipliers
  double tad_to_loop_ratio;
  double gev_mu, gev_sigma, gev_xi;
  int32_t noisify, skip_burnin;
  uint64_t hist_len, window;
  uint64_t min_burnin_epochs, max_burnin_epochs;
  double target_contact_density;
  StateLog log;
  // wave-parallel kernel only: per-phase 100 MHz wall-clock ticks are added here when set
  unsigned long long* phase;
};

// Largest chromosome end (bp) fast mode accepts: its per-epoch move arithmetic is 32-bit.
constexpr uint64_t kFastMaxEnd = 1ull << 30;

// Largest device-memory slab the wave-parallel deterministic kernel takes for cells too large for
// LDS; a job with more cells than fit runs as several launches over the same slab.
constexpr uint64_t kWaveSlabBytes = 8ull << 30;

// Device view of one chromosome: barrier table (sorted), band and its counters.
struct DetChrom {
  uint64_t start, end;
  uint64_t nlefs, nbar;
  const uint64_t* bar_pos;
  const uint8_t* bar_dir;
  const double* bar_stp_active;
  const double* bar_stp_inactive;
  uint32_t* band;
  unsigned long long* missed;
  unsigned long long* dropped;  // fast mode's abandoned samples: [0] contacts, [1] occupancy
  uint64_t nrows;
  uint64_t cpe;            // contacts per epoch
  uint64_t target_epochs;  // epochs criterion
  double bind_rate;        // nlefs / burnin_target_epochs_for_lef_activation
  const double* qn;        // pow(1 - p_loop, t) for t in [0, cpe], or null
  unsigned long long* occupancy;  // 1D LEF occupancy, ceil(size / bin) bins, or null
};

#ifndef MGPU_EMU
hipError_t launch_det_cells(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                            const WorkItem* d_items, uint64_t nitems, unsigned long long* d_next,
                            uint8_t* d_scratch, uint64_t stride, bool use_lds, uint32_t grid,
                            uint64_t* d_cell_stats);
// Wave-parallel deterministic mode: one 64-lane block per work item, block b taking item
// d_order[b] (nitems entries), lds_bytes = det_wave_bytes_host(max LEFs, max barriers, burn-in
// history) of dynamic LDS per block.
hipError_t launch_det_wave(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                           const WorkItem* d_items, const uint32_t* d_order, uint64_t nitems,
                           uint64_t lds_bytes, uint64_t* d_cell_stats, unsigned long long* d_err);
// The same for cells whose arrays live in device memory: d_slab holds nitems slots of `stride`
// bytes (det_wave_cell_bytes_host), LDS holds det_wave_fixed_bytes_host() per block.
hipError_t launch_det_wave_slab(hipStream_t stream, const DetParams& P, const DetChrom* d_chroms,
                                const WorkItem* d_items, const uint32_t* d_order, uint64_t nitems,
                                uint8_t* d_slab, uint64_t stride, uint64_t* d_cell_stats,
                                unsigned long long* d_err);
// One cell with its per-epoch hashes; d_cell: the cell's arrays in device memory, or null when
// lds_bytes holds everything.
hipError_t launch_det_wave_trace(hipStream_t stream, const DetParams& P, const DetChrom& ch,
                                 uint64_t seed, uint64_t target, uint64_t lds_bytes,
                                 uint8_t* d_cell, uint64_t* d_trace, uint64_t max_trace,
                                 uint64_t* d_out);
uint64_t det_wave_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist);
uint64_t det_wave_fixed_bytes_host();
uint64_t det_wave_cell_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist);
// The wave-parallel kernel's phase-tick accumulators (kDetWavePhases of them, a device array),
// and a read that clears them.
constexpr int kDetWavePhases = 12;
unsigned long long* det_wave_phase_ptr();
hipError_t det_wave_phase_times(uint64_t* out, int n);
hipError_t launch_det_trace(hipStream_t stream, const DetParams& P, const DetChrom& ch,
                            uint64_t seed, uint64_t target, uint8_t* d_scratch, uint64_t* d_trace,
                            uint64_t max_trace, uint64_t* d_out);
hipError_t launch_det_step(hipStream_t stream, const DetParams& P, const DetChrom& ch, int op,
                           uint64_t n, uint64_t* rev, uint64_t* fwd, uint64_t* ep, uint32_t* rr,
                           uint32_t* fr, uint32_t* tmp, uint64_t* rm, uint64_t* fm, uint64_t* rc,
                           uint64_t* fc, uint8_t* bst, uint64_t* rng);
hipError_t launch_det_draws(hipStream_t stream, uint64_t seed, int kind, double a, double b,
                            double c, uint64_t n, uint64_t* d_out);
uint64_t det_cell_bytes_host(uint64_t nlefs, uint64_t nbar, uint64_t hist);
// Band increments of deterministic mode for (rows[i], cols[i]), in order (ch.band, ch.missed,
// ch.nrows are read).
hipError_t launch_det_band(hipStream_t stream, const DetChrom& ch, const uint64_t* d_rows,
                           const uint64_t* d_cols, uint64_t n);
// One band as the block sums read it: nrows x ncols pixels, data[col * nrows + (col - row)].
struct BlockBand {
  const uint32_t* band;
  uint64_t nrows, ncols;
};
// unsafe_get_block of n queries (band chrom[i], centre (row[i], col[i])), block_size odd.
hipError_t launch_block_sums(hipStream_t stream, const BlockBand* d_bands, const uint64_t* d_chrom,
                             const uint64_t* d_row, const uint64_t* d_col, uint64_t n,
                             uint32_t block_size, uint32_t* d_out);
// Collision words of deterministic mode for n (index, event) pairs.
hipError_t launch_det_collisions(hipStream_t stream, const uint64_t* d_idx, const uint8_t* d_ev,
                                 uint64_t n, uint64_t* d_words, uint64_t* d_index,
                                 uint32_t* d_event, uint32_t* d_flags);

#endif  // MGPU_EMU

}  // namespace mgpu

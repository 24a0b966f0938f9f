// Fast-mode job interface (counter-based RNG, one workgroup per cell).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/modle_gpu.h"
#include "kernels.h"

namespace mgpu {

constexpr int kFastPhaseCount = 16;  // epoch phases with timers

struct FastJob {
  uint64_t nitems = 0;
  uint64_t nchroms = 0;
  void* impl = nullptr;
};

// Builds the 32-bit barrier tables (from the caller's host arrays, host_bars[c] for chromosome
// c) and the longest-first work order for a job whose bands already live on the device.
// release[2c] and release[2c + 1] are chromosome c's release probabilities after and during
// burn-in.
int fast_job_create(const mgpu_params& p, const std::vector<DetChrom>& chroms,
                    const mgpu_barriers* host_bars, const std::vector<double>& release,
                    const std::vector<WorkItem>& items, FastJob* job);
// timers != 0: accumulate per-phase wall-clock ticks (read with fast_phase_times). With nev at
// least twice the number of launches, ev[2k] and ev[2k + 1] are recorded before and after launch
// k. A run of several size tiers launches each on a stream of its own, forked from and joined
// back into `stream` (so the launches of one run may overlap).
hipError_t fast_job_run(hipStream_t stream, FastJob* job, const WorkItem* d_items,
                        uint64_t* d_cell_stats, int timers, hipEvent_t* ev = nullptr,
                        uint32_t nev = 0);
// Times every workgroup size on `sample` cells spread over the work order (0: all cells) and
// keeps the fastest: per size tier when the job launches by tiers, else for its single launch.
// ms (5 entries, may be null) receives the summed times of 64 to 1024 lanes. The bands are left
// dirty (runs zero them first).
int fast_job_autotune(hipStream_t stream, FastJob* job, const WorkItem* d_items,
                      uint64_t* d_cell_stats, uint64_t sample, float* ms);
// Workgroup size per cell (64 to 1024, a power of two) for one launch over every cell; 0 (the
// default) launches once per size tier instead. Results do not depend on it.
int fast_job_set_threads(FastJob* job, uint32_t threads);
uint32_t fast_job_threads(const FastJob* job);
// The launches the next run makes (one per size tier, or the single launch): workgroup size,
// grid, cells and first position in the work order of each, at most max of them; returns their
// number.
uint32_t fast_job_tiers(const FastJob* job, uint32_t* threads, uint32_t* grid, uint64_t* cells,
                        uint32_t max, uint64_t* begin = nullptr);
// The work order (work-item indices, launch by launch).
const uint32_t* fast_job_order(const FastJob* job);
// Workgroup size of one size tier (index as listed by fast_job_tiers in tiered mode).
int fast_job_set_tier_threads(FastJob* job, uint32_t tier, uint32_t threads);
// global != 0: keep every cell's working set in a per-workgroup slab of device memory even when
// it fits in LDS (the layout large chromosomes need), in one launch; 0: back to LDS where it fits.
int fast_job_set_storage(FastJob* job, int global);
int fast_job_global_storage(const FastJob* job);
// Largest grid of the launches the next run makes.
uint32_t fast_job_grid(const FastJob* job);
// Copies the first n per-phase tick totals (100 MHz wall clock, summed over cells) and clears them.
hipError_t fast_phase_times(uint64_t* out, int n);
void fast_job_destroy(FastJob* job);
const char* fast_last_error();
uint64_t fast_cell_bytes_host(uint32_t n, uint32_t nbar, uint32_t hist);
// Clears `bytes` (a multiple of 4) of device memory with a kernel (capture-safe).
hipError_t launch_zero_words(hipStream_t stream, void* p, uint64_t bytes);
// Sets the dynamic-LDS limit of every fast-mode kernel to the most one workgroup may take
// (host-side, outside captures; job creation does it too).
hipError_t fast_job_prepare(FastJob* job);
// Per-epoch state records of later runs go to log (log.rec == nullptr: none).
void fast_job_set_log(FastJob* job, const StateLog& log);
// serial != 0: bypass fix-ups walk their whole list on one lane, in the reference loop's order
// (a check of the parallel runs, which must give identical results).
void fast_job_set_serial_fixups(FastJob* job, int serial);
// Philox4x32-10 of n (counter[4], key[2]) tuples: in holds 6 words per tuple, out 4.
hipError_t launch_philox_kat(hipStream_t stream, const uint32_t* d_in, uint32_t* d_out, int n);
// ---- the allocation-free entry point ----
// Scratch memory of fast_simulate_noalloc: one statically allocated copy per GPU (4 MiB), enough
// for 65536 barriers and 16384 cells of one chromosome.
constexpr uint64_t kNoAllocScratchBytes = 4ull << 20;
constexpr uint64_t kNoAllocMaxBarriers = 65536;
constexpr uint64_t kNoAllocMaxCells = 16384;
// Resolves this device's copy of the scratch memory (call once per device with that device
// current; later calls return at once).
hipError_t fast_noalloc_init(int device);
// Simulates the work items (all of chromosome `chrom`, whose band, nrows, cpe, target epochs and
// bind rate are read from chrom; its barrier arrays are the host arrays pos/dir/sa/si) in fast mode
// with no device allocation: inputs, counters and statistics go to this device's scratch memory,
// contacts are added to chrom.band (the caller's device memory). Requires cells whose working set
// fits in LDS. ev0 / ev1 bracket the launch; *missed, dropped[0..1] and sums (epochs, burn-in
// epochs, contacts over the cells) are returned. 0 on success, -1 with fast_last_error() set.
int fast_simulate_noalloc(hipStream_t stream, int device, uint32_t num_cus, const mgpu_params& p,
                          const DetChrom& chrom, const uint64_t* pos, const uint8_t* dir,
                          const double* sa, const double* si, double release,
                          double release_burnin, const std::vector<WorkItem>& items,
                          hipEvent_t ev0, hipEvent_t ev1, uint64_t* missed, uint64_t* dropped,
                          uint64_t* sums);
// Fast mode's band increments (32-bit coordinates) for (rows[i], cols[i]), in order.
hipError_t launch_fast_band(hipStream_t stream, uint32_t* d_band, unsigned long long* d_missed,
                            uint32_t nrows, const uint32_t* d_rows, const uint32_t* d_cols,
                            uint64_t n);
// fast mode's double-precision ln (fn 0), expm1 (fn 1) or sin and cos of 2 pi x (fn 2; d_out
// holds 2n values, sines first) on n inputs.
hipError_t launch_fast_math(hipStream_t stream, int fn, const double* d_in, double* d_out, int n);

}  // namespace mgpu

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

// Builds the 32-bit barrier tables and the longest-first work order for a job whose chromosome
// descriptors (bands, barrier arrays) already live on the device.
int fast_job_create(hipStream_t stream, const mgpu_params& p, const std::vector<DetChrom>& chroms,
                    const std::vector<WorkItem>& items, const DetChrom* d_chroms, FastJob* job);
// timers != 0: accumulate per-phase wall-clock ticks (read with fast_phase_times)
hipError_t fast_job_run(hipStream_t stream, FastJob* job, const WorkItem* d_items,
                        uint64_t* d_cell_stats, int timers);
// Workgroup size per cell (256, 512 or 1024; 0 = default); results do not depend on it.
int fast_job_set_threads(FastJob* job, uint32_t threads);
uint32_t fast_job_threads(const FastJob* job);
uint32_t fast_job_grid(const FastJob* job);
// Copies the first n per-phase tick totals (100 MHz wall clock, summed over cells) and clears them.
hipError_t fast_phase_times(uint64_t* out, int n);
void fast_job_destroy(FastJob* job);
const char* fast_last_error();
uint64_t fast_cell_bytes_host(uint32_t n, uint32_t nbar, uint32_t hist);
// Philox4x32-10 of n (counter[4], key[2]) tuples: in holds 6 words per tuple, out 4.
hipError_t launch_philox_kat(hipStream_t stream, const uint32_t* d_in, uint32_t* d_out, int n);

}  // namespace mgpu

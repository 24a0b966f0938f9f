// Fast-mode job interface (counter-based RNG, wavefront-parallel cells).
#pragma once
#include <hip/hip_runtime.h>
#include <stdint.h>

#include <vector>

#include "../../include/modle_gpu.h"
#include "kernels.h"

namespace mgpu {

struct FastJob {
  void* d_mem = nullptr;
  uint64_t nitems = 0;
  uint64_t nchroms = 0;
  void* impl = nullptr;
};

int fast_job_create(hipStream_t stream, const mgpu_params& p, const std::vector<DetChrom>& chroms,
                    const std::vector<WorkItem>& items, const DetChrom* d_chroms, FastJob* job);
hipError_t fast_job_set_items(FastJob* job, const std::vector<WorkItem>& items);
hipError_t fast_job_run(hipStream_t stream, FastJob* job, uint64_t* d_cell_stats);
void fast_job_destroy(FastJob* job);
const char* fast_last_error();

}  // namespace mgpu

// Fast-mode kernels (placeholder until the wavefront-parallel engine lands).
#include <string>

#include "fast_kernels.h"

namespace mgpu {

static thread_local std::string g_fast_err;

int fast_job_create(hipStream_t, const mgpu_params&, const std::vector<DetChrom>&,
                    const std::vector<WorkItem>&, const DetChrom*, FastJob*) {
  g_fast_err = "fast mode is not available in this build";
  return -1;
}
hipError_t fast_job_set_items(FastJob*, const std::vector<WorkItem>&) { return hipErrorNotSupported; }
hipError_t fast_job_run(hipStream_t, FastJob*, uint64_t*) { return hipErrorNotSupported; }
void fast_job_destroy(FastJob*) {}
const char* fast_last_error() { return g_fast_err.c_str(); }

}  // namespace mgpu

// Instantiates the batched level kernel (FP32, the stencils of kBatchIds).
#define AN5D_BATCH_KERNELS
#include "batch.hpp"

namespace an5d {

hipError_t batch_launch_level(const BatchRec* recs, const BatchUnit* units, int nunits, int* counter,
                              int grid, hipStream_t st) {
  hipLaunchKernelGGL(batch_level_kernel, dim3(grid), dim3(64 * kBatchNW), 0, st, recs, units, nunits,
                     counter);
  return hipGetLastError();
}

hipError_t batch_launch_reset(int* counters, int n, hipStream_t st) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(256), 0, st, counters, n);
  return hipGetLastError();
}

int batch_kind_of(int stencil) {
  for (int k = 0; k < kBatchKinds; k++)
    if (kBatchIds[k] == stencil) return k;
  return -1;
}
int batch_max_bt() { return kBatchMaxBT; }
int batch_ck() { return kBatchCK; }

}  // namespace an5d

// Instantiates the batched level kernels (FP32, the stencils of kBatchIds).
#define AN5D_BATCH_KERNELS
#include "batch.hpp"

namespace an5d {

static const void* level_kernel(int kind) {
  if (kind == kLevelAdj) return reinterpret_cast<const void*>(&batch_level_kernel<kLevelAdj>);
  if (kind == kLevelFwdMisc) return reinterpret_cast<const void*>(&batch_level_kernel<kLevelFwdMisc>);
  return reinterpret_cast<const void*>(&batch_level_kernel<kLevelFwd>);
}

hipError_t batch_launch_level(int kind, const BatchRec* recs, const BatchUnit* units, int nunits,
                              int* counter, int grid, hipStream_t st) {
  if (kind == kLevelAdj)
    hipLaunchKernelGGL(batch_level_kernel<kLevelAdj>, dim3(grid), dim3(64 * kBatchNW), 0, st, recs, units,
                       nunits, counter);
  else if (kind == kLevelFwdMisc)
    hipLaunchKernelGGL(batch_level_kernel<kLevelFwdMisc>, dim3(grid), dim3(64 * kBatchNW), 0, st, recs,
                       units, nunits, counter);
  else
    hipLaunchKernelGGL(batch_level_kernel<kLevelFwd>, dim3(grid), dim3(64 * kBatchNW), 0, st, recs, units,
                       nunits, counter);
  return hipGetLastError();
}

hipError_t batch_launch_reset(int* counters, int n, hipStream_t st) {
  hipLaunchKernelGGL(batch_reset_kernel, dim3(1), dim3(256), 0, st, counters, n);
  return hipGetLastError();
}

// resident workgroups per CU of a level kernel (0 if the query fails)
int batch_occupancy(int kind) {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&n, level_kernel(kind), 64 * kBatchNW, 0) != hipSuccess)
    return 0;
  return n;
}

int batch_kind_of(int stencil) {
  for (int k = 0; k < kBatchKinds; k++)
    if (kBatchIds[k] == stencil) return k;
  return -1;
}
int batch_max_bt() { return kBatchMaxBT; }
int batch_adj_max_bt(int kind) { return batch_adj_btmax(kind); }
int batch_ck() { return kBatchCK; }
int batch_adj_ck() { return kBatchAdjCK; }

}  // namespace an5d

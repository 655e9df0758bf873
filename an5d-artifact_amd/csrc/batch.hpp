// Batched sweep: many independent 2D grids advanced together.
//
// The instances' launch schedules are laid side by side as "levels": level
// l holds launch l of every instance that has one (instances differ in size,
// stencil, T and therefore in their number of levels).  One launch per level
// covers all of its tiles; its waves pull (record, tile) work units from a
// device counter, so small and large grids, cheap and expensive stencils
// share the chip without a per-grid launch each.  Stream order between level
// launches is the only dependency (a tile of level l+1 reads the halo its
// neighbours wrote at level l), so no wave ever waits on another inside a
// launch.  The whole level sequence is captured once into a HIP graph.
#pragma once

#include "sweep.hpp"

namespace an5d {

struct BatchRec {
  const void* src;
  void* dst;
  const void* coef;
  int has_div;
  int kind;  // index into the batch kernel's stencil list
  int bt;
  int D0, D2;
  long long s0;
  int ntk, nseg, sl, SK;
};

struct BatchUnit {
  int rec, tile;
};

#ifdef AN5D_BATCH_KERNELS  // kernels: defined in one translation unit only

// Stencils the batched kernel is built for (2D, FP32, CK = 4).
constexpr int kBatchIds[] = {8 /* j2d5pt */, 9 /* j2d9pt */, 1 /* star2d2r */, 10 /* j2d9pt-gol */};
constexpr int kBatchKinds = 4;
constexpr int kBatchMaxBT = 4;
constexpr int kBatchCK = 4;
constexpr int kBatchNW = 4;

template <int ID, int BT>
__device__ __forceinline__ void batch_tile(const BatchRec& r, int tile) {
  SweepArgs<float> a;
  a.src = static_cast<const float*>(r.src);
  a.dst = static_cast<float*>(r.dst);
  a.coef = static_cast<const float*>(r.coef);
  a.gscale = 1.0;
  a.has_div = r.has_div;
  a.D0 = r.D0;
  a.D1 = 1;
  a.D2 = r.D2;
  a.s0 = r.s0;
  a.s1 = 0;
  a.ntk = r.ntk;
  a.ntj = 1;
  a.nseg = r.nseg;
  a.sl = r.sl;
  a.SK = r.SK;
  a.SJ = 1;
  a.mid = nullptr;
  a.mid_stride = 0;
  a.mid_mask = 0;
  a.nt_store = 0;
  constexpr int UNR = default_unroll<Def<ID>::rad>();
  if constexpr (Def<ID>::divisor != 0) {
    if (r.has_div == 2) sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, 2, UNR>(a, tile);
    else sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, -1, UNR>(a, tile);
  } else {
    if (r.has_div == 0) sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, 0, UNR>(a, tile);
    else sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, -1, UNR>(a, tile);
  }
}

template <int K>
__device__ __forceinline__ void batch_dispatch_bt(const BatchRec& r, int tile) {
  constexpr int ID = kBatchIds[K];
  switch (r.bt) {
    case 1: batch_tile<ID, 1>(r, tile); break;
    case 2: batch_tile<ID, 2>(r, tile); break;
    case 3: batch_tile<ID, 3>(r, tile); break;
    default: batch_tile<ID, 4>(r, tile); break;
  }
}

__global__ __launch_bounds__(64 * kBatchNW) void batch_level_kernel(const BatchRec* __restrict__ recs,
                                                                    const BatchUnit* __restrict__ units,
                                                                    int nunits, int* __restrict__ counter) {
  const int lane = threadIdx.x & 63;
  // every wave exits once the counter passes nunits
  for (;;) {
    int u = 0;
    if (lane == 0) u = atomicAdd(counter, 1);
    u = __builtin_amdgcn_readlane(u, 0);
    if (u >= nunits) break;
    const BatchUnit wu = units[u];
    const BatchRec r = recs[wu.rec];
    switch (r.kind) {
      case 0: batch_dispatch_bt<0>(r, wu.tile); break;
      case 1: batch_dispatch_bt<1>(r, wu.tile); break;
      case 2: batch_dispatch_bt<2>(r, wu.tile); break;
      default: batch_dispatch_bt<3>(r, wu.tile); break;
    }
  }
}

__global__ void batch_reset_kernel(int* __restrict__ counters, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counters[i] = 0;
}

#endif  // AN5D_BATCH_KERNELS

}  // namespace an5d

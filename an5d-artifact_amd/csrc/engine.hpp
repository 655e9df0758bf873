// Host-side view of the compiled kernel variants.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include <atomic>
#include <vector>

namespace an5d {

constexpr int kMaxBT = 16;
constexpr int kMaxMidBT = 8;  // deepest forward block with intermediate-stage outputs
// grids (both buffers) above this size stream their stores past the 256 MB
// last-level cache
constexpr size_t kStreamBytes = 512ull << 20;

constexpr int kMaxDevices = 64;

// hipFuncSetAttribute acts on the current device, so a kernel that needs more
// than the default 64 KB of LDS opts in once per device.  `done` is that
// kernel's own per-device flag array (static storage, zero-initialised).
inline hipError_t lds_opt_in(const void* kernel, size_t lds, std::atomic<int>* done) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return e;
  if (dev < 0 || dev >= kMaxDevices) return hipErrorInvalidDevice;
  if (done[dev].load(std::memory_order_acquire)) return hipSuccess;
  e = hipFuncSetAttribute(kernel, hipFuncAttributeMaxDynamicSharedMemorySize, (int)lds);
  if (e == hipSuccess) done[dev].store(1, std::memory_order_release);
  return e;
}

// kRows: one step per launch, rows streamed with per-row partial sums (2D
// boxes); launched with a one-step sweep's layout
enum KernelKind { kSweep = 0, kDirect = 1, kPlane = 2, kRows = 3 };
enum DType { kF32 = 0, kF64 = 1 };

// dtype-erased SweepArgs<T>
struct LaunchArgs {
  const void* src;
  void* dst;
  const void* coef;
  double gscale;
  int has_div;
  int D0, D1, D2;
  long long s0, s1;
  int ntk, ntj, nseg, sl;
  int SK, SJ;
  void* mid;             // intermediate-stage outputs (see SweepArgs::mid_mask)
  long long mid_stride;  // elements between consecutive intermediate outputs
  unsigned mid_mask;     // bit S: store stage S; 0: none
  int nt_store;          // 1: non-temporal stores for shallow blocks
};

using LaunchFn = hipError_t (*)(const LaunchArgs&, dim3 grid, dim3 block, size_t lds,
                                hipStream_t);

struct Variant {
  int stencil, dtype, kind;
  int ck, cj, nw, btmax;
  int regcap;              // register cap per lane of the deeper blocks (0: none)
  size_t lds[kMaxBT + 1];  // dynamic LDS bytes per workgroup, by bt
  LaunchFn fn[kMaxBT + 1];       // fn[bt], bt in [1, btmax], exact mode
  LaunchFn fn_fast[kMaxBT + 1];  // fast (FMA) mode, same shapes
  // exact mode in the shared-product form (radius-1 stars whose face taps
  // carry one coefficient); null where the stencil has no such form
  LaunchFn fn_share[kMaxBT + 1];
  // private (scratch) bytes per thread of fn_share[bt] (-1 unknown): a kernel
  // that spills is slower than the general form and is not used
  int (*share_scratch[kMaxBT + 1])();
  // resident workgroups per CU of fn[bt] (mode 0), fn_fast[bt] (1), fn_share[bt] (2)
  int (*occ[kMaxBT + 1])(int mode);
  LaunchFn fn_mid[kMaxBT + 1];  // exact mode storing intermediate stages (2D, bt <= kMaxMidBT)
};

// Every stencil x dtype translation unit appends its variants here.
std::vector<Variant>& registry();

// The adjoint workspace pads every grid with kAdjPad cells on both ends of
// each row (>= the widest tile overrun) and kAdjRowPad zero rows above and
// below it (>= the deepest plane overrun of a segment), so the adjoint
// kernel's loads need no bounds checks.
constexpr int kAdjPad = 320;
constexpr int kAdjRowPad = 64;
// blocks of 256 threads of the adjoint seed kernel: one loss slot per wave
constexpr int kSeedBlocks = 1024;

// dtype-erased AdjArgs<T>
struct AdjLaunch {
  const void* U;
  const void* Lin;
  void* Lout;
  const void* coef;
  double gscale;
  int has_div;
  double* gslot;  // NG doubles per wave tile
  int D0, D2;
  long long s0;
  int ntk, nseg, sl, SK;
};
using AdjFn = hipError_t (*)(const AdjLaunch&, dim3 grid, dim3 block, hipStream_t);
using SeedFn = hipError_t (*)(const void* uT, const void* target, void* lam, double* lslot, int D0,
                              int D2, long long s0, long long st, int R, hipStream_t);
using HaloFn = hipError_t (*)(const void* src, void* dst, int D0, int D2, long long s0, int R,
                              hipStream_t);

struct AdjVariant {
  int stencil, dtype;
  int ck, nw, btmax, reach;
  SeedFn seed;
  HaloFn halo;
  AdjFn fn[kMaxBT + 1];
  int (*occupancy)(int bt);  // resident workgroups per CU of fn[bt] (0: unknown)
};

std::vector<AdjVariant>& adj_registry();

}  // namespace an5d

// Host-side view of the compiled kernel variants.
#pragma once

#include <hip/hip_runtime.h>
#include <stddef.h>

#include <vector>

namespace an5d {

constexpr int kMaxBT = 16;

enum KernelKind { kSweep = 0, kDirect = 1 };
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
};

using LaunchFn = hipError_t (*)(const LaunchArgs&, dim3 grid, dim3 block, size_t lds,
                                hipStream_t);

struct Variant {
  int stencil, dtype, kind;
  int ck, cj, nw, btmax;
  size_t lds;              // dynamic LDS bytes per workgroup
  LaunchFn fn[kMaxBT + 1];  // fn[bt], bt in [1, btmax]
};

// Every stencil x dtype translation unit appends its variants here.
std::vector<Variant>& registry();

}  // namespace an5d

/*
 * an5d_kernel_stencil.h -- drop-in definition of the reference's per-benchmark
 * host entry point
 *
 *     double kernel_stencil(SB_TYPE *A1, int compsize, int timestep, bool scop);
 *
 * declared at common.h:42 and defined once per benchmark binary by each
 * generated host launcher (compiled/float/j2d5pt-256-10-128_host.cu:10).
 * Expand AN5D_DEFINE_KERNEL_STENCIL("<stencil>") in the one translation unit
 * that should define it (SB_TYPE must be float or double), link liban5d.so,
 * and the reference's common.h harness (or this repository's harness) runs
 * on the MI355X unchanged.
 *
 * Behaviour follows the generated host: scop != 0 runs the GPU path (kernel
 * time only), scop == 0 the CPU loop; nothing happens when
 * compsize < 1 or timestep < 1 (host.cu:17).  A failure prints
 * "HIP error: ..." (host.cu:18-30 prints "CUDA error: ...") or
 * "[AN5D ERROR] ..." to stderr and aborts, like the reference's asserts.
 */
#ifndef AN5D_KERNEL_STENCIL_H
#define AN5D_KERNEL_STENCIL_H

#include <stdio.h>
#include <stdlib.h>

#include "an5d.h"

#ifdef __cplusplus
#define AN5D_KS_LINKAGE extern "C"
#else
#include <stdbool.h>
#define AN5D_KS_LINKAGE
#endif

#define AN5D_DEFINE_KERNEL_STENCIL(NAME)                                              \
  AN5D_KS_LINKAGE double kernel_stencil(SB_TYPE* A1, int compsize, int timestep,      \
                                        bool scop)                                    \
  {                                                                                   \
    static int an5d_id = -1;                                                          \
    if (an5d_id < 0) an5d_id = an5d_stencil_id(NAME);                                 \
    if (an5d_id < 0) {                                                                \
      fprintf(stderr, "[AN5D ERROR] %s\n", an5d_last_error());                        \
      fflush(stderr);                                                                 \
      abort();                                                                        \
    }                                                                                 \
    double an5d_ms = 0.0;                                                             \
    int an5d_rc = an5d_kernel_stencil(an5d_id,                                        \
                                      sizeof(SB_TYPE) == 8 ? AN5D_FLOAT64 : AN5D_FLOAT32, \
                                      A1, compsize, timestep, scop ? 1 : 0, &an5d_ms); \
    if (an5d_rc != AN5D_OK) {                                                         \
      fprintf(stderr, an5d_rc == AN5D_E_HIP ? "HIP error: %s\n" : "[AN5D ERROR] %s\n", \
              an5d_last_error());                                                     \
      fflush(stderr);                                                                 \
      abort();                                                                        \
    }                                                                                 \
    return an5d_ms;                                                                   \
  }

#endif /* AN5D_KERNEL_STENCIL_H */

/*
 * Test infrastructure: the reference's own benchmark harness (common.h, read
 * from the reference tree at build time, unmodified) linked against
 * liban5d.so through the drop-in definition of kernel_stencil
 * (include/an5d_kernel_stencil.h).  This is the two-line host file a
 * maintainer would add in place of a generated <stencil>-<config>_host.cu.
 *
 * Built by `make -C oracle dropin` with -DBENCH_DIM/-DBENCH_FPP/-DBENCH_RAD
 * (the values of the stencil's .c file), -DSB_TYPE and -DDROPIN_STENCIL.
 */
#include "common.h"
#include "an5d_kernel_stencil.h"

AN5D_DEFINE_KERNEL_STENCIL(DROPIN_STENCIL)

/*
 * an5d.h -- C ABI of the MI355X temporal-blocking stencil engine (liban5d.so).
 *
 * Plain C types only: pointers, sizes, ints and doubles.  Every entry point
 * returns an int status (AN5D_OK == 0); on failure an5d_last_error() holds a
 * message for the calling thread.  The library never aborts and never
 * exits; the benchmark harness maps failures onto the reference's
 * stderr-and-abort behaviour.
 *
 * Data layout (the reference's, common.h:44-56 and j2d5pt.c:10-11): a
 * problem with interior extents n[] and radius r owns two buffers of
 * D = n + 2r cells per dimension, stored back to back ("buffer-major"),
 * each row-major with the last dimension fastest.  The halo (r cells on
 * every side) holds fixed (Dirichlet) values; it is read, never written.
 * After T steps buffer T%2 holds u_T.
 */
#ifndef AN5D_H
#define AN5D_H

#include <stddef.h>

#ifdef __cplusplus
extern "C" {
#endif

#define AN5D_OK 0
#define AN5D_E_INVALID (-1)      /* bad argument */
#define AN5D_E_UNSUPPORTED (-2)  /* no compiled kernel fits the request */
#define AN5D_E_HIP (-3)          /* HIP runtime error */
#define AN5D_E_ALLOC (-4)        /* host or device allocation failed */

#define AN5D_FLOAT32 0
#define AN5D_FLOAT64 1

/* A stencil problem.  `stencil` is an id from an5d_stencil_id().  n[] are
 * interior extents, outermost first (2D uses n[0], n[1]; the reference only
 * has squares, n[d] == compsize).  coefs == NULL selects the source's float
 * literals (FP64 uses (double)(float)literal, as the reference's
 * `5.1f * A[..]` does in double code, j2d5pt.c:18); otherwise it holds one
 * value per tap in source term order.  For gradient2d coefs is
 * {eps, scale}: out = C + scale / sqrt(eps + sum (C-N)^2) (reference:
 * eps = 0.0001f, scale = 1).  divisor < 0 keeps the source's divisor
 * (118 or 159 or none), 0 removes it, > 0 replaces it. */
typedef struct an5d_problem {
  int stencil;
  int dtype;
  long long n[3];
  int timesteps;
  const double* coefs;
  double divisor;
} an5d_problem;

/* Tiling.  bt: timesteps fused per launch (0: tuned/default).  variant:
 * index into an5d_variant_info() for this stencil x dtype (-1: tuned or
 * default).  sl: planes per streaming segment along dim 0 (0: default).
 * mode: AN5D_EXACT (0) reproduces the C loop bit for bit (source term order,
 * no contraction, correctly rounded division); AN5D_FAST (1) folds the
 * divisor into the coefficients and chains FMAs (results within the
 * tolerance checked by the tests, not bitwise).  In exact mode a radius-1
 * star whose face taps all carry the same coefficient value forms each
 * product c * u once per cell instead of once per use (the same rounded
 * values in the same order: bitwise the same results);
 * AN5D_EXACT_GENERAL (2) is exact mode without that form. */
#define AN5D_EXACT 0
#define AN5D_FAST 1
#define AN5D_EXACT_GENERAL 2
typedef struct an5d_config {
  int bt;
  int variant;
  int sl;
  int mode;
} an5d_config;

/* Element strides of a device-resident problem: cell (i,j,k) of buffer b is
 * at base[b*buf_stride + i*s0 + j*s1 + k] (2D: i*s0 + k).  The reference
 * layout is s1 = D2, s0 = D1*D2 (2D: s0 = D1), buf_stride = D0*s0. */
typedef struct an5d_layout {
  long long s0, s1, buf_stride;
} an5d_layout;

typedef struct an5d_stencil_info {
  char name[32];
  int dim, rad, fpp, ntaps, divisor, linear;
} an5d_stencil_info;

typedef struct an5d_variant_info {
  int kind;  /* 0: temporally blocked sweep, 1: one step per launch from global
               memory, 2: one step per launch on whole tiles, nothing
               recomputed (3D stars of radius 2-4: the plane staged in LDS;
               3D radius-1 boxes: plane-streamed partial sums), 3: one step
               per launch, row-streamed partial sums (2D boxes of radius
               2-4) */
  int ck, cj, nw, btmax;
  long long lds_bytes;
  int regcap; /* registers per lane the compiler was held to in the deeper
                 temporal blocks (bt > btmax/2), 0: no cap beyond the
                 workgroup size's; the --maxrregcount analogue
                 (make.sh:29-31) */
} an5d_variant_info;

/* ---- registry ------------------------------------------------------------
 * Replaces the compile-time BENCH_DIM/BENCH_FPP/BENCH_RAD macros of each
 * <stencil>.c (j2d5pt.c:1-3) with a runtime table of all 21 stencils. */
int an5d_stencil_count(void);
int an5d_stencil_id(const char* name); /* "j2d5pt", "j2d9pt-gol", ... ; <0 if unknown */
int an5d_stencil_get_info(int stencil, an5d_stencil_info* out);
/* taps: 4 ints per tap (sign, d_outer, d_mid, d_fast) and the literal coefficients */
int an5d_stencil_taps(int stencil, int* offsets, double* coefs, int max_taps);
int an5d_variant_count(int stencil, int dtype);
int an5d_variant_get_info(int stencil, int dtype, int variant, an5d_variant_info* out);

/* ---- launch schedule ----------------------------------------------------
 * The step counts of the launches that advance T steps at temporal block bt
 * so that u_T lands in buffer T%2: floor(T/bt) blocks of bt, with the tail
 * re-split when the launch count's parity differs from T's.  Same rule as
 * the generated host launcher (compiled/float/j2d5pt-256-10-128_host.cu:72-554).
 * Returns the number of launches written to steps[] (or <0). */
int an5d_schedule(int timesteps, int bt, int* steps, int max_launches);

/* ---- forward sweep --------------------------------------------------------
 * Drop-in core of `double kernel_stencil(SB_TYPE *A1, int compsize, int
 * timestep, bool scop)` (common.h:42; generated host.cu:10).  A1 is the
 * caller's host array (2 buffers, reference layout, square grid).  scop != 0
 * runs on the GPU: device buffers (cached across calls), H2D copy, sync,
 * timed launch schedule, sync, D2H copy (host.cu:34-41,74-77,556-564);
 * scop == 0 runs the CPU path (the OpenMP loop, host.cu:567-575).  *ms gets
 * the elapsed milliseconds of the timed region (kernels only on the GPU,
 * the whole loop on the CPU, as in the reference).  Does nothing when
 * D < 2r+1 or timestep < 1 (host.cu:17). */
int an5d_kernel_stencil(int stencil, int dtype, void* A1, int compsize, int timestep,
                        int scop, double* ms);

/* Same as the GPU branch above for any problem (rectangular grids, other
 * coefficients) and tiling.  A: host array, reference layout.  Thread-safe:
 * calls from several threads on one device take turns on its staging grid
 * (each call holds it from the upload to the download). */
int an5d_forward_host(const an5d_problem* prob, void* A, const an5d_config* cfg, double* ms);

/* Device-resident form: dA holds both buffers in `layout` (NULL: reference
 * layout).  Enqueues the launch schedule on `stream` (a hipStream_t, NULL:
 * the null stream) and returns without synchronising.  c0 is the launch
 * ordinal of the first launch (0: read buffer 0).  No replacement in the
 * reference, whose kernels are only reachable through kernel_stencil. */
int an5d_forward_device(const an5d_problem* prob, void* dA, const an5d_layout* layout,
                        const an5d_config* cfg, void* stream);

/* One launch of `steps` fused steps reading buffer c0%2 and writing
 * (c0+1)%2: the kernel0_<steps>(A, dimsize, timestep, c0) launch of the
 * generated host (compiled/float/j2d5pt-256-10-128_kernel.hu:3-12). */
int an5d_launch(const an5d_problem* prob, void* dA, const an5d_layout* layout,
                const an5d_config* cfg, int steps, int c0, void* stream);

/* The CPU path: T steps of the OpenMP loop (j2d5pt.c:23-31) on a host
 * array in the reference layout; threads <= 0 uses the OpenMP default. */
int an5d_cpu_forward(const an5d_problem* prob, void* A, int threads, double* ms);

/* Fill both buffers of a host grid (cells = cells per buffer, halo
 * included) from rand_r seeded with index + seed_base, in parallel.
 * mode 0 is the reference's init_grid (common.h:44-56, seed_base =
 * compsize): 1000*(T)rand_r/(T)RAND_MAX.  mode 1: rand_r/(RAND_MAX+1.0),
 * uniform in [0, 1). */
int an5d_init_grid(int dtype, void* A, long long cells, long long seed_base, int mode);

/* ---- adjoint (2D stencils) ------------------------------------------------
 * No counterpart in the reference, which computes no gradients.
 * Loss L = 1/2 sum over interior cells of (u_T - target)^2.  Outputs:
 * dL/du_0 on the interior (the halo is a fixed boundary condition: its
 * entries are 0), dL/dcoefs (one per tap in source order; gradient2d:
 * {d/d eps, d/d scale}) and L.  Single grids (one buffer, D0 x D1 cells,
 * reference layout).  cfg->bt is the adjoint block (0: deepest compiled),
 * cfg->variant / cfg->sl the forward kernel and the adjoint segment length.
 * The forward pass runs at its own tuned block depth and stores a checkpoint
 * every bt steps; each reverse launch recomputes the
 * states of its block from the checkpoint and runs the transposed sweep for
 * bt steps.  The gradient and loss sums are reduced in a fixed order, so
 * repeated calls give identical results.  The workspace holds one padded
 * grid per bt-step block plus three (T/bt + 3 grids: 8192^2 FP32 at T=1000,
 * bt=4 is about 75 GB); a size beyond the current device's memory is
 * rejected with AN5D_E_ALLOC and a message. */
int an5d_adjoint_workspace(const an5d_problem* prob, const an5d_config* cfg, size_t* bytes);
/* Device form: enqueues on `stream`; *ms_fwd (optional) gets the time of the
 * checkpointed forward pass (synchronises on an event when non-NULL).
 * gcoef and loss are device doubles.  The workspace (size from
 * an5d_adjoint_workspace for the same problem and cfg) holds the checkpoints,
 * two adjoint buffers and per-wave partial sums; its grids carry zero padding
 * around them, so it must be zero-filled once before its first use (the
 * library never writes the padding).  Calls that share a workspace must not
 * overlap. */
int an5d_adjoint_device(const an5d_problem* prob, const void* u0, const void* target, void* grad,
                        double* gcoef, double* loss, void* workspace, size_t workspace_bytes,
                        const an5d_config* cfg, void* stream, double* ms_fwd);
/* The same without any device allocation, for callers that cannot allocate
 * (code under HIP graph capture, tight serving loops, harnesses that forbid
 * hipMalloc).  No counterpart in the reference.  The workspace and the
 * coefficient table live in a fixed 64 MiB device array that the library
 * holds on every device it runs on until the process ends; problems whose
 * an5d_adjoint_workspace() size exceeds 64 MiB - 1 KiB are rejected with
 * AN5D_E_INVALID.  Each call clears the part it uses with a fill kernel (no
 * memset) and passes the coefficients by value through a kernel argument, so
 * the enqueued work reads no host memory.  All calls on one device share that
 * array: calls that overlap on one device (other streams or threads) race.
 * Results are bit for bit those of an5d_adjoint_device for the same inputs. */
int an5d_adjoint_device_noalloc(const an5d_problem* prob, const void* u0, const void* target,
                                void* grad, double* gcoef, double* loss, const an5d_config* cfg,
                                void* stream);
/* Host form: copies in, runs, copies out; *ms_total = forward + adjoint. */
int an5d_adjoint_host(const an5d_problem* prob, const void* u0, const void* target, void* grad,
                      double* gcoef, double* loss, const an5d_config* cfg, double* ms_fwd,
                      double* ms_total);

/* ---- batched forward and batched forward + adjoint (2D, FP32) -------------
 * No counterpart in the reference (one grid per binary, one grid per launch:
 * compiled/float/j2d5pt-256-10-128_host.cu:74-77).  Stencils j2d5pt, j2d9pt,
 * star2d2r and j2d9pt-gol; any sizes and timesteps.
 *
 * Each instance's work is a chain of levels; one kernel launch covers one
 * level of every instance, its waves pulling (instance, tile) work units of
 * every kind from a device counter, and the whole level sequence is captured
 * into one HIP graph on the first run (replayed afterwards).  Chains are
 * placed on the level axis so that the estimated work per level is even.
 *
 * an5d_batch_create: forward only.  dA[i] holds both buffers of instance i
 * (reference layout); after a run buffer T_i%2 holds u_T, bit for bit the
 * single-grid result.  cfgs (optional): per-instance bt (<= 5; 0: ranked by
 * the autotuner's model at the batch kernel's shape) and sl (0: batch-wide).
 *
 * an5d_batch_adjoint_create: forward + adjoint.  Currently switched off: it
 * returns AN5D_E_UNSUPPORTED (its first GPU run stalled; cause not found).
 * Designed behaviour: for every instance, the
 * same loss, gradients and conventions as an5d_adjoint_device (single grids
 * in the reference layout; grad's halo is 0), with the forward states
 * checkpointed every bt steps in a workspace the batch allocates and owns
 * ((T/bt + 3) padded grids per instance).  When the workspaces would exceed
 * the memory budget ($AN5D_BATCH_MEM bytes, else half the free device
 * memory), the largest instances keep only every C-th checkpoint (C ~
 * sqrt(T/bt)) and recompute the others segment by segment (about sqrt(T/bt)
 * grids, one more forward pass).  cfgs (optional): per-instance
 * adjoint block bt (0: the deepest compiled: 5 for j2d5pt, 4 for j2d9pt-gol,
 * 3 for the radius-2 stars) and sl.  Loss and gradient sums are reduced in a
 * fixed order, so repeated runs give identical results.
 *
 * an5d_batch_tune: times the level sequence (`reps` replays each) at each
 * candidate batch-wide segment length (32-256) and, for a forward batch, bt
 * policy (each instance's model-prior bt, or bt = 3, 4, 5 for all instances
 * without a caller bt), and keeps the fastest; a forward batch's grids are
 * restored afterwards.  an5d_batch_run: enqueue one run on
 * `stream` (a hipStream_t); runs that overlap on one batch race. */
typedef struct an5d_batch an5d_batch;
typedef struct an5d_batch_grad {
  const void* u0;     /* initial state, one grid (halo included) */
  const void* target; /* one grid */
  void* grad;         /* out: dL/du0, one grid */
  double* gcoef;      /* out: dL/dcoefs, one per tap */
  double* loss;       /* out: L */
} an5d_batch_grad;
int an5d_batch_create(an5d_batch** out, int n, const an5d_problem* probs, void* const* dA,
                      const an5d_config* cfgs);
int an5d_batch_adjoint_create(an5d_batch** out, int n, const an5d_problem* probs,
                              const an5d_batch_grad* io, const an5d_config* cfgs);
int an5d_batch_tune(an5d_batch* batch, int reps, int* best_sl, double* best_ms);
int an5d_batch_run(an5d_batch* batch, void* stream);
int an5d_batch_info(const an5d_batch* batch, int* levels, long long* work_units);
int an5d_batch_destroy(an5d_batch* batch);

/* ---- autotuner ------------------------------------------------------------
 * Replaces the reference's Excel performance model + top-5 compile/run
 * scripts (Model_V100.xlsm, compiled/nvcc_compile_top5.sh, compiled/run.sh).
 * an5d_autotune ranks every (kernel shape, bt, sl) for `prob` with the
 * analytic model (an5d_model_estimate), times the best max_measured of them
 * on the GPU (dA: the caller's device grid, or NULL for a scratch grid) and
 * records the winner in the table that the forward entry points consult
 * when cfg->variant < 0.  The shapes include register-capped twins
 * (an5d_variant_info.regcap), so the cap is searched like the reference's
 * --maxrregcount builds.  A caller's grid is kept in host memory during the
 * search and written back before the call returns.  an5d_tune_save/load
 * persist the table as text (default file: an5d-artifact_amd/tuned.tsv, or
 * $AN5D_TUNE_FILE; its last column records the chosen kernel's cap). */
int an5d_autotune(const an5d_problem* prob, void* dA, const an5d_layout* layout, int mode,
                  int max_measured, an5d_config* best, double* best_gstencil);
int an5d_model_estimate(const an5d_problem* prob, const an5d_config* cfg, double* gstencil);
/* The configuration the forward entry points would run for `prob` and `cfg`
 * (NULL: all defaults): the variant, bt and sl after the tuned-table lookup
 * and the per-stencil defaults, with sl = the planes per segment actually
 * launched (0 for one-step kernels).  The reference fixes these at compile
 * time per binary (compiled/nvcc_compile_top5.sh); this reports the choice. */
int an5d_resolve_config(const an5d_problem* prob, const an5d_config* cfg, an5d_config* out);
int an5d_tune_load(const char* path);
int an5d_tune_save(const char* path);
int an5d_tune_clear(void);

/* Release cached device buffers. */
int an5d_release(void);

const char* an5d_last_error(void);

#ifdef __cplusplus
}
#endif

#endif /* AN5D_H */

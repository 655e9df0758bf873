// Benchmark binary: the reference's common.h harness (common.h:21-180)
// restated over liban5d, with the same CLI (-s -t -n -c -h), the same
// output lines (parsed by compiled/run.sh:29 and summarize.sh:32-42), and
// extra lines after them: GStencil/s, a relative-error verdict and, for
// float, the error against an FP64 CPU run of the same input.
//
// Built once per stencil x type:  -DAN5D_STENCIL="j2d5pt" -DSB_TYPE=float.
// The stencil's dim / radius / FPP come from the library registry instead of
// the BENCH_* macros.  Tiling can be forced with AN5D_BT, AN5D_VARIANT and
// AN5D_SL in the environment (read by the library).
#include <float.h>
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include "../../include/an5d.h"
#include "../../include/an5d_kernel_stencil.h"

#ifndef SB_TYPE
#define SB_TYPE float
#endif
#ifndef AN5D_STENCIL
#error "AN5D_STENCIL must name the stencil"
#endif

AN5D_DEFINE_KERNEL_STENCIL(AN5D_STENCIL)

static an5d_stencil_info g_info;

static void usage(char* cmd) {
  printf("%s [OPTION] ...\n", cmd);
  printf("\n");
  printf("-s\tSpecify compute size\n");
  printf("-t\tSpecify the number of timesteps\n");
  printf("-n\tSpecify the number of execution\n");
  printf("-c\tEnable comparison to CPU execution\n");
  printf("-h\tShow this usage\n");
}

// common.h:44-56
template <typename T>
static void init_grid(T* A, int compsize) {
  size_t gridarea = 1, dimsize = compsize + g_info.rad * 2;
  for (int i = 0; i < g_info.dim; i++) gridarea *= dimsize;
#pragma omp parallel for
  for (size_t i = 0; i < gridarea; i++) {
    unsigned int seed = i + compsize;
    A[0 * gridarea + i] = 1000 * (T)(rand_r(&seed)) / (T)(RAND_MAX);
    A[1 * gridarea + i] = A[0 * gridarea + i];
  }
}

// common.h:58-78, plus a relative-error / non-finite check
template <typename T>
static bool check_result(const T* A, const T* B, int compsize, int timestep) {
  int t = timestep % 2;
  size_t gridarea = 1, dimsize = compsize + g_info.rad * 2;
  for (int i = 0; i < g_info.dim; i++) gridarea *= dimsize;
  double error_max = 0.0, rel_max = 0.0;
  long double error_sum = 0.0;
  long long nonfinite_mismatch = 0;
  const double floor_ = sizeof(T) == 4 ? (double)FLT_MIN : DBL_MIN;
#pragma omp parallel for reduction(max : error_max, rel_max) reduction(+ : error_sum, nonfinite_mismatch)
  for (size_t i = 0; i < gridarea; i++) {
    const T a = A[t * gridarea + i], b = B[t * gridarea + i];
    double error = fabs((double)(a - b));
    error_max = (error > error_max) ? error : error_max;
    error_sum += error * error;
    if (isfinite(a) && isfinite(b)) {
      double rel = error / fmax(fabs((double)b), floor_);
      rel_max = rel > rel_max ? rel : rel_max;
    } else if (!(isnan(a) && isnan(b)) && a != b) {
      nonfinite_mismatch++;
    }
  }
  printf("Max Error: %e\n", error_max);
  printf("RMS Error: %Le\n", sqrtl(error_sum / gridarea));
  const double rtol = sizeof(T) == 4 ? 1e-5 : 1e-12;
  printf("Max Rel Error: %e (floor %e, non-finite mismatches %lld)\n", rel_max, floor_,
         nonfinite_mismatch);
  const bool pass = rel_max <= rtol && nonfinite_mismatch == 0;
  printf("Verdict: %s (rtol %g)\n", pass ? "PASS" : "FAIL", rtol);
  return pass;
}

// float runs: GPU result against the CPU path in FP64 on the same input
static void check_fp64(const float* A, int compsize, int timestep) {
  size_t gridarea = 1, dimsize = compsize + g_info.rad * 2;
  for (int i = 0; i < g_info.dim; i++) gridarea *= dimsize;
  double* S = (double*)malloc(2 * gridarea * sizeof(double));
  float* F = (float*)malloc(2 * gridarea * sizeof(float));
  if (!S || !F) {
    printf("Max Error vs FP64: skipped (host memory)\n");
    free(S);
    free(F);
    return;
  }
  init_grid<float>(F, compsize);
  for (size_t i = 0; i < 2 * gridarea; i++) S[i] = F[i];
  an5d_problem p;
  memset(&p, 0, sizeof(p));
  p.stencil = an5d_stencil_id(AN5D_STENCIL);
  p.dtype = AN5D_FLOAT64;
  for (int d = 0; d < g_info.dim; d++) p.n[d] = compsize;
  p.timesteps = timestep;
  p.divisor = -1.0;
  double ms = 0.0;
  if (an5d_cpu_forward(&p, S, 0, &ms) != AN5D_OK) {
    printf("Max Error vs FP64: skipped (%s)\n", an5d_last_error());
  } else {
    int t = timestep % 2;
    double emax = 0.0;
    for (size_t i = 0; i < gridarea; i++) {
      double e = fabs((double)A[t * gridarea + i] - S[t * gridarea + i]);
      emax = e > emax ? e : emax;
    }
    printf("Max Error vs FP64: %e\n", emax);
  }
  free(S);
  free(F);
}

static double calc_gflops(double total_flop, double ms) { return (total_flop / ms) / (1000 * 1000); }

static int benchmark(int compsize, int timestep, int reptnum, bool cpucomp) {
  size_t gridarea = 1, comparea = 1;
  for (int i = 0; i < g_info.dim; i++) {
    gridarea *= (compsize + g_info.rad * 2);
    comparea *= compsize;
  }
  double total_flop = (double)timestep * comparea * g_info.fpp;
  SB_TYPE* A = (SB_TYPE*)malloc(2 * gridarea * sizeof(SB_TYPE));
  SB_TYPE* B = (SB_TYPE*)malloc(2 * gridarea * sizeof(SB_TYPE));
  if (A == NULL || B == NULL) {
    fprintf(stderr, "Host memory allocation failed\n");
    exit(EXIT_FAILURE);
  }
  puts("[Device warm-up]");
  printf("%.4f ms\n", kernel_stencil(A, compsize, timestep, true));
  double elapsed = 0.0f;
  puts("\n[Device execution]");
  for (int n = 0; n < reptnum; n++) {
    init_grid<SB_TYPE>(A, compsize);
    double ms = kernel_stencil(A, compsize, timestep, true);
    printf("%7d: %10.4lf GFLOPS, %10.4f ms\n", n, calc_gflops(total_flop, ms), ms);
    elapsed += ms;
  }
  printf("Average: %10.4lf GFLOPS, %10.4f ms\n", calc_gflops(total_flop, elapsed / reptnum),
         elapsed / reptnum);
  printf("GStencil/s: %10.4lf\n", (double)timestep * comparea / (elapsed / reptnum) / 1e6);
  int rc = EXIT_SUCCESS;
  if (cpucomp) {
    puts("\n[CPU execution]");
    init_grid<SB_TYPE>(B, compsize);
    printf("%.4f ms\n", kernel_stencil(B, compsize, timestep, false));
    if (!check_result<SB_TYPE>(A, B, compsize, timestep)) rc = EXIT_FAILURE;
    if (sizeof(SB_TYPE) == 4) check_fp64((const float*)A, compsize, timestep);
  }
  free(A);
  free(B);
  return rc;
}

int main(int argc, char** argv) {
  int id = an5d_stencil_id(AN5D_STENCIL);
  if (id < 0 || an5d_stencil_get_info(id, &g_info) != AN5D_OK) {
    fprintf(stderr, "[AN5D ERROR] %s\n", an5d_last_error());
    return EXIT_FAILURE;
  }
  int compsize = (g_info.dim == 2) ? 8192 : 512;
  int timestep = 120;
  int reptnum = 2;
  bool cpucomp = false;
  int opt;
  while ((opt = getopt(argc, argv, "s:t:n:ch")) != -1) {
    switch (opt) {
      case 's':
        compsize = (int)strtol(optarg, NULL, 10);
        if (compsize < 1) {
          fprintf(stderr, "Larger compute size required: %d\n", compsize);
          exit(EXIT_FAILURE);
        }
        break;
      case 't':
        timestep = (int)strtol(optarg, NULL, 10);
        if (timestep < 1) {
          fprintf(stderr, "More than one time-step required: %d\n", timestep);
          exit(EXIT_FAILURE);
        }
        break;
      case 'n':
        reptnum = (int)strtol(optarg, NULL, 10);
        if (reptnum < 1) {
          fprintf(stderr, "More than one execution required: %d\n", reptnum);
          exit(EXIT_FAILURE);
        }
        break;
      case 'c':
        cpucomp = true;
        break;
      case 'h':
        usage(argv[0]);
        return EXIT_SUCCESS;
    }
  }
  return benchmark(compsize, timestep, reptnum, cpucomp);
}

// Benchmark binary bin/<stencil>.<float|double>: the reference's a.out for one
// stencil, over liban5d.
//
// The command line (-s -t -n -c -h), the error messages and exit codes, and
// the output lines up to "Average:" and "Max Error:" / "RMS Error:" follow the
// reference harness (common.h:21-30,85-180): compiled/run.sh:29 and
// summarize.sh read them.  After them this harness prints its own lines: the
// GStencil/s rate, a relative-error verdict, and for float runs the error
// against an FP64 run of the CPU path on the same input.
//
// Built once per stencil x type:  -DAN5D_STENCIL="j2d5pt" -DSB_TYPE=float.
// Dimension, radius and FLOPs per update come from the library's registry.
// Tiling can be forced with AN5D_BT, AN5D_VARIANT, AN5D_SL and AN5D_MODE in
// the environment (read by the library).
#include <float.h>
#include <getopt.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <vector>

#include "../../include/an5d.h"
#include "../../include/an5d_kernel_stencil.h"

#ifndef SB_TYPE
#define SB_TYPE float
#endif
#ifndef AN5D_STENCIL
#error "AN5D_STENCIL must name the stencil"
#endif

AN5D_DEFINE_KERNEL_STENCIL(AN5D_STENCIL)

namespace {

constexpr int kDtype = sizeof(SB_TYPE) == 8 ? AN5D_FLOAT64 : AN5D_FLOAT32;

struct Shape {
  an5d_stencil_info info;
  long long cells;     // per buffer, halo included
  long long interior;  // updated cells per step
};

Shape shape_of(const an5d_stencil_info& info, int n) {
  Shape s{info, 1, 1};
  for (int d = 0; d < info.dim; d++) {
    s.cells *= n + 2LL * info.rad;
    s.interior *= n;
  }
  return s;
}

void print_usage(const char* cmd) {
  printf("%s [OPTION] ...\n\n", cmd);
  printf("-s\tSpecify compute size\n");
  printf("-t\tSpecify the number of timesteps\n");
  printf("-n\tSpecify the number of execution\n");
  printf("-c\tEnable comparison to CPU execution\n");
  printf("-h\tShow this usage\n");
}

// Both buffers from the reference's generator (common.h:44-56: rand_r seeded
// with cell index + n, scaled to [0, 1000]).
void fill(SB_TYPE* A, const Shape& s, int n) {
  if (an5d_init_grid(kDtype, A, s.cells, n, 0) != AN5D_OK) {
    fprintf(stderr, "[AN5D ERROR] %s\n", an5d_last_error());
    exit(EXIT_FAILURE);
  }
}

// Buffer T%2 of the GPU run against the CPU run: the reference's two lines
// (max and RMS absolute error, common.h:58-78), then the relative error with
// the smallest normal as floor, non-finite mismatches and the verdict.
bool compare(const SB_TYPE* gpu, const SB_TYPE* cpu, const Shape& s, int T) {
  const SB_TYPE* g = gpu + (size_t)(T % 2) * s.cells;
  const SB_TYPE* c = cpu + (size_t)(T % 2) * s.cells;
  const double floor_rel = kDtype == AN5D_FLOAT64 ? DBL_MIN : (double)FLT_MIN;
  double amax = 0.0, rmax = 0.0;
  long double sq = 0.0;
  long long bad = 0;
#pragma omp parallel for reduction(max : amax, rmax) reduction(+ : sq, bad)
  for (long long i = 0; i < s.cells; i++) {
    const double d = fabs((double)(g[i] - c[i]));
    if (d > amax) amax = d;
    sq += (long double)d * d;
    if (isfinite(g[i]) && isfinite(c[i])) {
      const double r = d / fmax(fabs((double)c[i]), floor_rel);
      if (r > rmax) rmax = r;
    } else if (!(isnan(g[i]) && isnan(c[i])) && g[i] != c[i]) {
      bad++;
    }
  }
  printf("Max Error: %e\n", amax);
  printf("RMS Error: %Le\n", sqrtl(sq / s.cells));
  const double rtol = kDtype == AN5D_FLOAT64 ? 1e-12 : 1e-5;
  printf("Max Rel Error: %e (floor %e, non-finite mismatches %lld)\n", rmax, floor_rel, bad);
  const bool ok = rmax <= rtol && bad == 0;
  printf("Verdict: %s (rtol %g)\n", ok ? "PASS" : "FAIL", rtol);
  return ok;
}

// Float runs: the GPU result against the CPU path in FP64 from the same input.
void compare_fp64(const SB_TYPE* gpu, const Shape& s, int n, int T) {
  if (kDtype != AN5D_FLOAT32) return;
  std::vector<double> D;
  try {
    D.resize((size_t)2 * s.cells);
  } catch (...) {
    printf("Max Error vs FP64: skipped (host memory)\n");
    return;
  }
  std::vector<float> F((size_t)2 * s.cells);
  fill(reinterpret_cast<SB_TYPE*>(F.data()), s, n);
  for (size_t i = 0; i < D.size(); i++) D[i] = F[i];
  an5d_problem p;
  memset(&p, 0, sizeof(p));
  p.stencil = an5d_stencil_id(AN5D_STENCIL);
  p.dtype = AN5D_FLOAT64;
  for (int d = 0; d < s.info.dim; d++) p.n[d] = n;
  p.timesteps = T;
  p.divisor = -1.0;
  double ms = 0.0;
  if (an5d_cpu_forward(&p, D.data(), 0, &ms) != AN5D_OK) {
    printf("Max Error vs FP64: skipped (%s)\n", an5d_last_error());
    return;
  }
  const SB_TYPE* g = gpu + (size_t)(T % 2) * s.cells;
  const double* r = D.data() + (size_t)(T % 2) * s.cells;
  double e = 0.0;
  for (long long i = 0; i < s.cells; i++) e = fmax(e, fabs((double)g[i] - r[i]));
  printf("Max Error vs FP64: %e\n", e);
}

// The reference's run sequence (common.h:85-129): one warm-up call, `reps`
// timed calls each on a freshly initialised grid, their average, then the
// optional CPU comparison.
int run(const Shape& s, int n, int T, int reps, bool check) {
  std::vector<SB_TYPE> gpu, cpu;
  try {
    gpu.resize((size_t)2 * s.cells);
    if (check) cpu.resize((size_t)2 * s.cells);
  } catch (...) {
    fprintf(stderr, "Host memory allocation failed\n");
    exit(EXIT_FAILURE);
  }
  const double flops = (double)T * (double)s.interior * s.info.fpp;
  auto gflops = [&](double ms) { return flops / ms / 1e6; };
  puts("[Device warm-up]");
  printf("%.4f ms\n", kernel_stencil(gpu.data(), n, T, true));
  puts("\n[Device execution]");
  double total_ms = 0.0;
  for (int k = 0; k < reps; k++) {
    fill(gpu.data(), s, n);
    const double ms = kernel_stencil(gpu.data(), n, T, true);
    printf("%7d: %10.4lf GFLOPS, %10.4f ms\n", k, gflops(ms), ms);
    total_ms += ms;
  }
  const double avg = total_ms / reps;
  printf("Average: %10.4lf GFLOPS, %10.4f ms\n", gflops(avg), avg);
  printf("GStencil/s: %10.4lf\n", (double)T * s.interior / avg / 1e6);
  if (!check) return EXIT_SUCCESS;
  puts("\n[CPU execution]");
  fill(cpu.data(), s, n);
  printf("%.4f ms\n", kernel_stencil(cpu.data(), n, T, false));
  const bool ok = compare(gpu.data(), cpu.data(), s, T);
  compare_fp64(gpu.data(), s, n, T);
  return ok ? EXIT_SUCCESS : EXIT_FAILURE;
}

// -s/-t/-n values below 1 end the program with the reference's messages
// (common.h:147-166).
int positive_arg(const char* text, const char* what) {
  const int v = (int)strtol(text, nullptr, 10);
  if (v < 1) {
    fprintf(stderr, "%s: %d\n", what, v);
    exit(EXIT_FAILURE);
  }
  return v;
}

}  // namespace

int main(int argc, char** argv) {
  an5d_stencil_info info;
  const int id = an5d_stencil_id(AN5D_STENCIL);
  if (id < 0 || an5d_stencil_get_info(id, &info) != AN5D_OK) {
    fprintf(stderr, "[AN5D ERROR] %s\n", an5d_last_error());
    return EXIT_FAILURE;
  }
  int n = info.dim == 2 ? 8192 : 512, T = 120, reps = 2;
  bool check = false;
  for (int opt; (opt = getopt(argc, argv, "s:t:n:ch")) != -1;) {
    if (opt == 's') n = positive_arg(optarg, "Larger compute size required");
    else if (opt == 't') T = positive_arg(optarg, "More than one time-step required");
    else if (opt == 'n') reps = positive_arg(optarg, "More than one execution required");
    else if (opt == 'c') check = true;
    else if (opt == 'h') {
      print_usage(argv[0]);
      return EXIT_SUCCESS;
    }
  }
  return run(shape_of(info, n), n, T, reps, check);
}

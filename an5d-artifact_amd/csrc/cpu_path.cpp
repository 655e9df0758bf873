// CPU path of the drop-in: the OpenMP loop of each <stencil>.c (e.g.
// j2d5pt.c:23-31, star3d1r.c:27-38) and of the generated host's else branch
// (compiled/float/j2d5pt-256-10-128_host.cu:567-575).  T steps, one
// `omp parallel for` over the outermost interior dimension per step, the
// halo untouched.  Built with g++ -O3 -fopenmp -ffp-contract=off and no
// -march flag, so every expression rounds exactly like the C source; the
// tap loop is unrolled at compile time from the generated tables.
//
// gradient2d follows the C semantics of gradient2d.c:18-22 (the float sum is
// passed to the double sqrt), like the GPU kernels.
#include <math.h>
#include <stdlib.h>
#include <omp.h>
#include <string.h>
#include <time.h>

#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "stencil_defs.hpp"

namespace an5d {
int fail(int code, const std::string& msg);  // api.cpp: sets an5d_last_error()
}

namespace {

using an5d::Def;
using an5d::TapDef;
using an5d::static_for;


template <int ID, typename T>
void run_cpu(T* A, long long D0, long long D1, long long D2, int steps, const T* c, T div,
             bool has_div, double gscale, int threads) {
  using D = Def<ID>;
  constexpr int R = D::rad;
  constexpr bool DIM3 = D::dim == 3;
  const long long s1 = D2, s0 = D1 * D2, area = D0 * D1 * D2;
  const long long jlo = DIM3 ? R : 0, jhi = DIM3 ? D1 - R : 1;
  for (int t = 0; t < steps; t++) {
    const T* in = A + (long long)(t % 2) * area;
    T* out = A + (long long)((t + 1) % 2) * area;
#pragma omp parallel for schedule(static) num_threads(threads)
    for (long long i = R; i < D0 - R; i++)
      for (long long j = jlo; j < jhi; j++)
        for (long long k = R; k < D2 - R; k++) {
          const T* p = in + i * s0 + j * s1 + k;
          T acc;
          if constexpr (D::linear) {
            static_for<0, D::ntaps>([&](auto qc) {
              constexpr int Q = decltype(qc)::value;
              constexpr TapDef tp = D::taps[Q];
              const T v = p[tp.d0 * s0 + tp.d1 * s1 + tp.d2];
              if constexpr (Q == 0) acc = c[0] * v;
              else if constexpr (tp.sign > 0) acc = acc + c[Q] * v;
              else acc = acc - c[Q] * v;
            });
            if (has_div) acc = acc / div;
          } else {
            const T cc = p[0];
            const T dn = cc - p[-s0], ds = cc - p[s0], de = cc - p[1], dw = cc - p[-1];
            T sum = c[0] + dn * dn;
            sum = sum + ds * ds;
            sum = sum + de * de;
            sum = sum + dw * dw;
            double q = 1.0 / sqrt((double)sum);
            if (gscale != 1.0) q = gscale * q;
            acc = (T)((double)cc + q);
          }
          out[i * s0 + j * s1 + k] = acc;
        }
  }
}

template <typename T>
using RunFn = void (*)(T*, long long, long long, long long, int, const T*, T, bool, double, int);

template <typename T>
RunFn<T> cpu_fn(int id) {
  static const RunFn<T> table[] = {
#define AN5D_CPU_FN(id_, ident, dim, rad, fpp, n, div, lin) &run_cpu<id_, T>,
      AN5D_STENCILS(AN5D_CPU_FN)};
  return table[id];
}

double now_ms() {
  struct timespec tp;
  clock_gettime(CLOCK_MONOTONIC_RAW, &tp);
  return (tp.tv_sec + tp.tv_nsec * 1.0e-9) * 1.0e3;
}

template <typename T>
int cpu_forward(const an5d_problem* p, void* Av, int threads, double* ms) {
  static const int dims[] = {
#define AN5D_CPU_DIM(id_, ident, dim, rad, fpp, n, div, lin) dim,
      AN5D_STENCILS(AN5D_CPU_DIM)};
  static const int rads[] = {
#define AN5D_CPU_RAD(id_, ident, dim, rad, fpp, n, div, lin) rad,
      AN5D_STENCILS(AN5D_CPU_RAD)};
  static const int ntaps[] = {
#define AN5D_CPU_NT(id_, ident, dim, rad, fpp, n, div, lin) (lin ? n : 1),
      AN5D_STENCILS(AN5D_CPU_NT)};
  static const int divs[] = {
#define AN5D_CPU_DIV(id_, ident, dim, rad, fpp, n, div, lin) div,
      AN5D_STENCILS(AN5D_CPU_DIV)};
  static const bool linear[] = {
#define AN5D_CPU_LIN(id_, ident, dim, rad, fpp, n, div, lin) (lin != 0),
      AN5D_STENCILS(AN5D_CPU_LIN)};
  static const std::vector<std::vector<float>> literals = [] {
    std::vector<std::vector<float>> v;
#define AN5D_CPU_LIT(id_, ident, dim, rad, fpp, n, div, lin)          \
  {                                                                    \
    std::vector<float> c;                                              \
    for (int q = 0; q < n; q++) c.push_back(Def<id_>::taps[q].coef);   \
    v.push_back(c);                                                    \
  }
    AN5D_STENCILS(AN5D_CPU_LIT)
    return v;
  }();
  const int id = p->stencil;
  const int dim = dims[id], r = rads[id];
  for (int d = 0; d < dim; d++)
    if (p->n[d] < 1) return an5d::fail(AN5D_E_INVALID, "interior extents must be >= 1");
  const long long D0 = p->n[0] + 2 * r;
  const long long D1 = dim == 3 ? p->n[1] + 2 * r : 1;
  const long long D2 = (dim == 3 ? p->n[2] : p->n[1]) + 2 * r;
  const int nt = ntaps[id];
  std::vector<T> c(nt);
  T div = 1;
  bool has_div = false;
  double gscale = 1.0;
  if (linear[id]) {
    for (int q = 0; q < nt; q++) c[q] = (T)(p->coefs ? p->coefs[q] : (double)literals[id][q]);
    if (p->divisor < 0) { has_div = divs[id] != 0; div = (T)divs[id]; }
    else if (p->divisor > 0) { has_div = true; div = (T)p->divisor; }
  } else {
    c[0] = (T)(p->coefs ? p->coefs[0] : (double)0.0001f);
    gscale = p->coefs ? p->coefs[1] : 1.0;
  }
  if (threads <= 0) threads = omp_get_max_threads();
  const double t0 = now_ms();
  cpu_fn<T>(id)(static_cast<T*>(Av), D0, D1, D2, p->timesteps, c.data(), div, has_div, gscale,
                threads);
  if (ms) *ms = now_ms() - t0;
  return AN5D_OK;
}

}  // namespace

extern "C" int an5d_cpu_forward(const an5d_problem* p, void* A, int threads, double* ms) {
  if (!p || !A) return an5d::fail(AN5D_E_INVALID, "null argument");
  if (p->stencil < 0 || p->stencil >= AN5D_NUM_STENCILS) return an5d::fail(AN5D_E_INVALID, "bad stencil id");
  if (p->timesteps < 0) return an5d::fail(AN5D_E_INVALID, "timesteps must be >= 0");
  if (ms) *ms = 0.0;
  if (p->dtype == AN5D_FLOAT32) return cpu_forward<float>(p, A, threads, ms);
  if (p->dtype == AN5D_FLOAT64) return cpu_forward<double>(p, A, threads, ms);
  return an5d::fail(AN5D_E_INVALID, "bad dtype");
}

// Grid initialisation for both buffers (halo included), per cell i of a
// buffer with seed = i + seed_base (the reference seeds with i + compsize,
// common.h:52):
//   mode 0: 1000 * (T)rand_r / (T)RAND_MAX           (common.h:53)
//   mode 1: rand_r / (RAND_MAX + 1.0), uniform [0, 1)
extern "C" int an5d_init_grid(int dtype, void* A, long long cells, long long seed_base, int mode) {
  if (!A || cells < 0 || (mode != 0 && mode != 1)) return an5d::fail(AN5D_E_INVALID, "bad init_grid arguments");
  if (dtype == AN5D_FLOAT32) {
    float* F = static_cast<float*>(A);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < cells; i++) {
      unsigned int seed = (unsigned int)(i + seed_base);
      const int r = rand_r(&seed);
      F[i] = mode == 0 ? 1000 * (float)r / (float)RAND_MAX : (float)(r / ((double)RAND_MAX + 1.0));
      F[cells + i] = F[i];
    }
    return AN5D_OK;
  }
  if (dtype == AN5D_FLOAT64) {
    double* F = static_cast<double*>(A);
#pragma omp parallel for schedule(static)
    for (long long i = 0; i < cells; i++) {
      unsigned int seed = (unsigned int)(i + seed_base);
      const int r = rand_r(&seed);
      F[i] = mode == 0 ? 1000 * (double)r / (double)RAND_MAX : r / ((double)RAND_MAX + 1.0);
      F[cells + i] = F[i];
    }
    return AN5D_OK;
  }
  return an5d::fail(AN5D_E_INVALID, "bad dtype");
}

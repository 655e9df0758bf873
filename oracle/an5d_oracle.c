/*
 * CPU oracle for the temporal-blocking stencil sweep.
 *
 * TEST INFRASTRUCTURE ONLY.  Only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load this library, and only as the
 * checker.  The product (liban5d.so) never links or calls it.
 *
 * It restates, in plain C compiled with -ffp-contract=off (no FMA, source
 * evaluation order), the reference's CPU semantics:
 *
 *   - oracle_run_*      the sequential `#pragma scop` loop of each
 *                       <stencil>.c (e.g. j2d5pt.c:14-21, star3d1r.c:15-25,
 *                       gradient2d.c:14-23): for t in [0,T) the interior of
 *                       A[(t+1)%2] is computed from A[t%2]; the halo is never
 *                       written.  Linear stencils evaluate
 *                       acc = c0*A[p0]; acc = acc (+|-) c_n*A[p_n] ...; acc/div
 *                       left to right, exactly as C parses the source sum.
 *                       FP64 keeps the float literals: coefficient =
 *                       (double)(float)literal (j2d5pt.c:18 with SB_TYPE=double).
 *                       gradient2d follows C promotion: the float sum goes
 *                       through the double sqrt() and a double 1.0f/..., then
 *                       the double sum is rounded once on the store.
 *   - oracle_init_grid  common.h:44-56 (rand_r seeded with index+compsize).
 *   - oracle_check      common.h:58-78 (max |a-b| and RMS over buffer T%2).
 *   - oracle_adjoint_*  FP64 reverse-mode sweep (no reference counterpart):
 *                       lambda_t = J_t^T lambda_{t+1} and the coefficient
 *                       gradient sum_t sum_x lambda_{t+1}(x) u_t(x+d)/div.
 *
 * Pinned by tests/test_oracle.py against: the reference's own sources built
 * into oracle/_ref (bitwise, every stencil x dtype), the committed golden
 * fixtures made from them, impulse-response / constant-field closed forms and
 * finite differences for the adjoint.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "../an5d-artifact_amd/csrc/stencils_gen.h"

typedef struct { int sign, d0, d1, d2; float coef; } otap;
typedef struct { int dim, rad, fpp, ntaps, divisor, linear; const otap* taps; } ostencil;

#define OT(s, a, b, c, v) {s, a, b, c, v},
#define OP(a, b, c) {1, a, b, c, 1.0f},
#define DECL_TAPS(id, ident, dim, rad, fpp, n, div, lin) \
  static const otap taps_##ident[] = { AN5D_TAPLIST_##lin(ident) };
#define AN5D_TAPLIST_1(ident) AN5D_TAPS_##ident(OT)
#define AN5D_TAPLIST_0(ident) AN5D_POINTS_##ident(OP)
AN5D_STENCILS(DECL_TAPS)
#define DECL_DESC(id, ident, dim, rad, fpp, n, div, lin) {dim, rad, fpp, n, div, lin, taps_##ident},
static const ostencil stencils[AN5D_NUM_STENCILS] = { AN5D_STENCILS(DECL_DESC) };

int oracle_num_stencils(void) { return AN5D_NUM_STENCILS; }

/* info[0..5] = dim, rad, fpp, ntaps, divisor, linear */
int oracle_stencil_info(int id, int* info)
{
  if (id < 0 || id >= AN5D_NUM_STENCILS) return -1;
  const ostencil* s = &stencils[id];
  info[0] = s->dim; info[1] = s->rad; info[2] = s->fpp;
  info[3] = s->ntaps; info[4] = s->divisor; info[5] = s->linear;
  return 0;
}

/* taps as (sign, d0, d1, d2) ints and the float coefficient literal values */
int oracle_stencil_taps(int id, int* offs, float* coefs)
{
  if (id < 0 || id >= AN5D_NUM_STENCILS) return -1;
  const ostencil* s = &stencils[id];
  for (int t = 0; t < s->ntaps; t++) {
    offs[4 * t + 0] = s->taps[t].sign;
    offs[4 * t + 1] = s->taps[t].d0;
    offs[4 * t + 2] = s->taps[t].d1;
    offs[4 * t + 3] = s->taps[t].d2;
    coefs[t] = s->taps[t].coef;
  }
  return s->ntaps;
}

/* Grid geometry: extents D[0..2] include the halo (D = n + 2r); 2D uses
 * D[1] = 1 for the absent middle dimension.  Buffers are stored back to
 * back: buffer b starts at b*area. */
typedef struct { int64_t D0, D1, D2, area; int r; } ogeom;

static ogeom make_geom(const ostencil* s, const int64_t* n)
{
  ogeom g;
  g.r = s->rad;
  g.D0 = n[0] + 2 * s->rad;
  if (s->dim == 2) { g.D1 = 1; g.D2 = n[1] + 2 * s->rad; }
  else { g.D1 = n[1] + 2 * s->rad; g.D2 = n[2] + 2 * s->rad; }
  g.area = g.D0 * g.D1 * g.D2;
  return g;
}

/* Linear stencils: `coefs` (ntaps values) overrides the literals when not
 * NULL; `divisor` < 0 keeps the source divisor, 0 removes it, > 0 replaces
 * it.  The float path rounds every override to float first. */
#define DEFINE_RUN(T, SUF)                                                        \
static void run_linear_##SUF(const ostencil* s, T* A, ogeom g, int T_steps,      \
                             const T* c, T div, int has_div)                     \
{                                                                                \
  const int64_t s1 = g.D2, s0 = g.D1 * g.D2;                                     \
  const int64_t lo1 = (s->dim == 3) ? g.r : 0, hi1 = (s->dim == 3) ? g.D1 - g.r : 1; \
  int64_t* off = (int64_t*)malloc(sizeof(int64_t) * s->ntaps);                   \
  for (int t = 0; t < s->ntaps; t++)                                             \
    off[t] = s->taps[t].d0 * s0 + s->taps[t].d1 * s1 + s->taps[t].d2;            \
  for (int t = 0; t < T_steps; t++) {                                            \
    const T* in = A + (int64_t)(t % 2) * g.area;                                 \
    T* out = A + (int64_t)((t + 1) % 2) * g.area;                                \
    for (int64_t i = g.r; i < g.D0 - g.r; i++)                                   \
      for (int64_t j = lo1; j < hi1; j++)                                        \
        for (int64_t k = g.r; k < g.D2 - g.r; k++) {                             \
          const int64_t x = i * s0 + j * s1 + k;                                 \
          T acc = c[0] * in[x + off[0]];                                         \
          for (int q = 1; q < s->ntaps; q++) {                                   \
            T prod = c[q] * in[x + off[q]];                                      \
            acc = (s->taps[q].sign > 0) ? acc + prod : acc - prod;               \
          }                                                                      \
          if (has_div) acc = acc / div;                                          \
          out[x] = acc;                                                          \
        }                                                                        \
  }                                                                              \
  free(off);                                                                     \
}
DEFINE_RUN(float, f32)
DEFINE_RUN(double, f64)

/* gradient2d.c:18-22 under C promotion rules.  eps is the `0.0001f`
 * literal; scale multiplies the reciprocal-root term (1.0 reproduces the
 * reference exactly, since multiplying a double by 1.0 is exact). */
static void run_gradient_f32(float* A, ogeom g, int T_steps, float eps, double scale)
{
  for (int t = 0; t < T_steps; t++) {
    const float* in = A + (int64_t)(t % 2) * g.area;
    float* out = A + (int64_t)((t + 1) % 2) * g.area;
    for (int64_t i = 1; i < g.D0 - 1; i++)
      for (int64_t j = 1; j < g.D2 - 1; j++) {
        const int64_t x = i * g.D2 + j;
        const float c = in[x];
        const float dn = c - in[x - g.D2], ds = c - in[x + g.D2];
        const float de = c - in[x + 1], dw = c - in[x - 1];
        float sum = eps + dn * dn;
        sum = sum + ds * ds;
        sum = sum + de * de;
        sum = sum + dw * dw;
        double q = 1.0 / sqrt((double)sum);
        if (scale != 1.0) q = scale * q;
        out[x] = (float)((double)c + q);
      }
  }
}

static void run_gradient_f64(double* A, ogeom g, int T_steps, double eps, double scale)
{
  for (int t = 0; t < T_steps; t++) {
    const double* in = A + (int64_t)(t % 2) * g.area;
    double* out = A + (int64_t)((t + 1) % 2) * g.area;
    for (int64_t i = 1; i < g.D0 - 1; i++)
      for (int64_t j = 1; j < g.D2 - 1; j++) {
        const int64_t x = i * g.D2 + j;
        const double c = in[x];
        const double dn = c - in[x - g.D2], ds = c - in[x + g.D2];
        const double de = c - in[x + 1], dw = c - in[x - 1];
        double sum = eps + dn * dn;
        sum = sum + ds * ds;
        sum = sum + de * de;
        sum = sum + dw * dw;
        double q = 1.0 / sqrt(sum);
        if (scale != 1.0) q = scale * q;
        out[x] = c + q;
      }
  }
}

/* Run T steps of stencil `id` on A (2 buffers).  n = interior extents
 * (outermost first; 2D uses n[0], n[1]).  For gradient2d, coefs (if given)
 * is {eps, scale}.  Returns 0 or -1 on a bad id. */
int oracle_run_f32(int id, float* A, const int64_t* n, int T_steps,
                   const double* coefs, double divisor)
{
  if (id < 0 || id >= AN5D_NUM_STENCILS) return -1;
  const ostencil* s = &stencils[id];
  ogeom g = make_geom(s, n);
  if (T_steps < 1) return 0;
  if (!s->linear) {
    float eps = coefs ? (float)coefs[0] : 0.0001f;
    double scale = coefs ? coefs[1] : 1.0;
    run_gradient_f32(A, g, T_steps, eps, scale);
    return 0;
  }
  float* c = (float*)malloc(sizeof(float) * s->ntaps);
  for (int q = 0; q < s->ntaps; q++) c[q] = coefs ? (float)coefs[q] : s->taps[q].coef;
  int has_div = (divisor < 0) ? (s->divisor != 0) : (divisor > 0);
  float div = (divisor < 0) ? (float)s->divisor : (float)divisor;
  run_linear_f32(s, A, g, T_steps, c, div, has_div);
  free(c);
  return 0;
}

int oracle_run_f64(int id, double* A, const int64_t* n, int T_steps,
                   const double* coefs, double divisor)
{
  if (id < 0 || id >= AN5D_NUM_STENCILS) return -1;
  const ostencil* s = &stencils[id];
  ogeom g = make_geom(s, n);
  if (T_steps < 1) return 0;
  if (!s->linear) {
    double eps = coefs ? coefs[0] : (double)0.0001f;
    double scale = coefs ? coefs[1] : 1.0;
    run_gradient_f64(A, g, T_steps, eps, scale);
    return 0;
  }
  double* c = (double*)malloc(sizeof(double) * s->ntaps);
  for (int q = 0; q < s->ntaps; q++) c[q] = coefs ? coefs[q] : (double)s->taps[q].coef;
  int has_div = (divisor < 0) ? (s->divisor != 0) : (divisor > 0);
  double div = (divisor < 0) ? (double)s->divisor : divisor;
  run_linear_f64(s, A, g, T_steps, c, div, has_div);
  free(c);
  return 0;
}

/* common.h:44-56: every cell of the whole grid (halo included), seed
 * index+compsize, 1000*(T)rand_r/(T)RAND_MAX, both buffers identical. */
void oracle_init_grid_f32(float* A, int dim, int rad, int compsize)
{
  size_t area = 1, D = (size_t)compsize + 2 * (size_t)rad;
  for (int i = 0; i < dim; i++) area *= D;
  for (size_t i = 0; i < area; i++) {
    unsigned int seed = (unsigned int)(i + compsize);
    A[i] = 1000 * (float)(rand_r(&seed)) / (float)(RAND_MAX);
    A[area + i] = A[i];
  }
}

void oracle_init_grid_f64(double* A, int dim, int rad, int compsize)
{
  size_t area = 1, D = (size_t)compsize + 2 * (size_t)rad;
  for (int i = 0; i < dim; i++) area *= D;
  for (size_t i = 0; i < area; i++) {
    unsigned int seed = (unsigned int)(i + compsize);
    A[i] = 1000 * (double)(rand_r(&seed)) / (double)(RAND_MAX);
    A[area + i] = A[i];
  }
}

/* common.h:58-78 restated; results returned instead of printed:
 * out[0] = max |A-B| over buffer T%2 (NaN ignored by the `>` test, as in
 * the reference), out[1] = RMS with a long double sum. */
#define DEFINE_CHECK(T, SUF)                                                    \
void oracle_check_##SUF(const T* A, const T* B, size_t area, int T_steps, double* out) \
{                                                                               \
  int t = T_steps % 2;                                                          \
  double error_max = 0.0;                                                       \
  long double error_sum = 0.0;                                                  \
  for (size_t i = 0; i < area; i++) {                                           \
    double error = fabs((double)(A[t * area + i] - B[t * area + i]));           \
    error_max = (error > error_max) ? error : error_max;                        \
    error_sum += error * error;                                                 \
  }                                                                             \
  out[0] = error_max;                                                           \
  out[1] = (double)sqrtl(error_sum / area);                                     \
}
DEFINE_CHECK(float, f32)
DEFINE_CHECK(double, f64)

/* ------------------------------------------------------------------------
 * FP64 adjoint oracle (no reference counterpart).
 *
 * Forward: u_{t+1} = S(u_t) on the interior, halo fixed.  Loss
 * L = 0.5 * sum over interior cells of (u_T - target)^2.  The halo is a
 * boundary condition, not a variable: its gradient entries are 0.
 * With lam_T = u_T - target on the interior (0 on the halo):
 *   linear:    lam_t(y) = sum_taps s_q c_q/div * lam_{t+1}(y - d_q), y interior
 *              dL/dc_q += sum_x lam_{t+1}(x) * s_q u_t(x + d_q) / div
 *   gradient2d: J_t^T from u_t (all states stored; oracle sizes only),
 *              dL/deps += lam * (-scale/2) rs^3, dL/dscale += lam * rs.
 * gu (out) = lam_0 on the interior, 0 on the halo.
 * ------------------------------------------------------------------------ */
int oracle_adjoint_f64(int id, const double* u0, const double* target,
                       const int64_t* n, int T_steps, const double* coefs,
                       double divisor, double* gu, double* gc, double* loss)
{
  if (id < 0 || id >= AN5D_NUM_STENCILS || T_steps < 1) return -1;
  const ostencil* s = &stencils[id];
  ogeom g = make_geom(s, n);
  const int64_t s1 = g.D2, s0 = g.D1 * g.D2, area = g.area;
  const int64_t lo1 = (s->dim == 3) ? g.r : 0, hi1 = (s->dim == 3) ? g.D1 - g.r : 1;
  double* states = (double*)malloc(sizeof(double) * area * (size_t)(T_steps + 1));
  double* buf = (double*)malloc(sizeof(double) * area * 2);
  if (!states || !buf) { free(states); free(buf); return -2; }
  memcpy(states, u0, sizeof(double) * area);
  for (int t = 0; t < T_steps; t++) {
    memcpy(buf, states + (size_t)t * area, sizeof(double) * area);
    memcpy(buf + area, states + (size_t)t * area, sizeof(double) * area);
    oracle_run_f64(id, buf, n, 1, coefs, divisor);
    memcpy(states + (size_t)(t + 1) * area, buf + area, sizeof(double) * area);
  }
  const double* uT = states + (size_t)T_steps * area;
  double L = 0.0;
  double* lam = (double*)calloc((size_t)area, sizeof(double));
  double* nlam = (double*)calloc((size_t)area, sizeof(double));
  for (int64_t x = 0; x < area; x++) gu[x] = 0.0;
  for (int64_t i = 0; i < g.D0; i++)
    for (int64_t j = 0; j < g.D1; j++)
      for (int64_t k = 0; k < g.D2; k++) {
        int64_t x = i * s0 + j * s1 + k;
        int interior = i >= g.r && i < g.D0 - g.r && k >= g.r && k < g.D2 - g.r &&
                       (s->dim == 2 || (j >= g.r && j < g.D1 - g.r));
        if (interior) {
          lam[x] = uT[x] - target[x];
          L += 0.5 * lam[x] * lam[x];
        }
      }
  int nc = s->linear ? s->ntaps : 2;
  for (int q = 0; q < nc; q++) gc[q] = 0.0;
  double* c = (double*)malloc(sizeof(double) * (s->linear ? s->ntaps : 1));
  int has_div = 0; double div = 1.0;
  if (s->linear) {
    for (int q = 0; q < s->ntaps; q++) c[q] = coefs ? coefs[q] : (double)s->taps[q].coef;
    has_div = (divisor < 0) ? (s->divisor != 0) : (divisor > 0);
    div = has_div ? ((divisor < 0) ? (double)s->divisor : divisor) : 1.0;
  }
  double eps = coefs ? coefs[0] : (double)0.0001f, scale = coefs ? coefs[1] : 1.0;
  for (int t = T_steps - 1; t >= 0; t--) {
    const double* ut = states + (size_t)t * area;
    memset(nlam, 0, sizeof(double) * area);
    for (int64_t i = g.r; i < g.D0 - g.r; i++)
      for (int64_t j = lo1; j < hi1; j++)
        for (int64_t k = g.r; k < g.D2 - g.r; k++) {
          const int64_t x = i * s0 + j * s1 + k;
          const double l = lam[x];
          if (s->linear) {
            for (int q = 0; q < s->ntaps; q++) {
              int64_t y = x + s->taps[q].d0 * s0 + s->taps[q].d1 * s1 + s->taps[q].d2;
              double w = (s->taps[q].sign > 0 ? 1.0 : -1.0) / div;
              nlam[y] += w * c[q] * l;
              gc[q] += w * l * ut[y];
            }
          } else {
            /* out = C + scale / sqrt(eps + sum_m (C - N_m)^2) */
            const int64_t nb[4] = {x - g.D2, x + g.D2, x + 1, x - 1};
            double sum = eps;
            for (int m = 0; m < 4; m++) sum += (ut[x] - ut[nb[m]]) * (ut[x] - ut[nb[m]]);
            double rs = 1.0 / sqrt(sum);
            double f = -scale * rs * rs * rs;  /* d(scale*sum^-1/2)/dsum * 2 / 2 */
            double dc = 1.0;
            for (int m = 0; m < 4; m++) {
              double dd = ut[x] - ut[nb[m]];
              dc += f * dd;
              nlam[nb[m]] += -f * dd * l;
            }
            nlam[x] += dc * l;
            gc[0] += l * (-0.5 * scale * rs * rs * rs);
            gc[1] += l * rs;
          }
        }
    /* lam on halo cells would be a sensitivity to the fixed boundary: drop it */
    for (int64_t i = 0; i < g.D0; i++)
      for (int64_t j = 0; j < g.D1; j++)
        for (int64_t k = 0; k < g.D2; k++) {
          int64_t x = i * s0 + j * s1 + k;
          int interior = i >= g.r && i < g.D0 - g.r && k >= g.r && k < g.D2 - g.r &&
                         (s->dim == 2 || (j >= g.r && j < g.D1 - g.r));
          if (!interior) nlam[x] = 0.0;
        }
    double* tmp = lam; lam = nlam; nlam = tmp;
  }
  for (int64_t x = 0; x < area; x++) gu[x] = lam[x];
  *loss = L;
  free(c); free(lam); free(nlam); free(states); free(buf);
  return 0;
}

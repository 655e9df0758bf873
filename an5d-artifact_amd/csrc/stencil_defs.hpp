// Compile-time stencil descriptors built from the generated tap tables, shared
// by the HIP kernels and the host-side CPU path.
//
// A stencil's *structure* (tap offsets, term order, signs, radius) is fixed at
// compile time so that every register index in the sweep resolves statically.
// Its *coefficients* (and the divisor) are runtime data: the reference
// literals by default (e.g. j2d5pt.c:18-20), or any other set, e.g. the
// seeded normalised sets of the benchmark presets or the values an adjoint
// differentiates against.
#pragma once

#include <stdint.h>

#include <type_traits>
#include <utility>

#include "stencils_gen.h"

#if !defined(__HIPCC__)
#define __host__
#define __device__
#define __forceinline__ inline
#endif

namespace an5d {

struct TapDef {
  int sign, d0, d1, d2;  // d0: outermost (streamed) dim, d2: fastest dim
  float coef;            // reference literal (float, as in the sources)
};

template <int ID> struct Def;

#define AN5D_TAP_ITEM(s, a, b, c, v) TapDef{s, a, b, c, v},
#define AN5D_PT_ITEM(a, b, c) TapDef{1, a, b, c, 1.0f},
#define AN5D_LIST_1(ident) AN5D_TAPS_##ident(AN5D_TAP_ITEM)
#define AN5D_LIST_0(ident) AN5D_POINTS_##ident(AN5D_PT_ITEM)
#define AN5D_DEF(id_, ident, dim_, rad_, fpp_, n_, div_, lin_)                 \
  template <> struct Def<id_> {                                                \
    static constexpr int id = id_, dim = dim_, rad = rad_, fpp = fpp_;         \
    static constexpr int ntaps = n_, divisor = div_;                           \
    static constexpr bool linear = lin_;                                       \
    static constexpr const char* name = #ident;                                \
    static constexpr TapDef taps[n_] = {AN5D_LIST_##lin_(ident)};              \
  };
AN5D_STENCILS(AN5D_DEF)
#undef AN5D_DEF

constexpr int kNumStencils = AN5D_NUM_STENCILS;
constexpr int kGradient2d = 11;  // the one nonlinear stencil (gradient2d.c:18-22)

// Does plane offset d0 need neighbours inside its plane (middle or fastest dim)?
template <class D>
constexpr bool plane_needs_ext(int d0) {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d0 == d0 && (D::taps[q].d1 != 0 || D::taps[q].d2 != 0)) return true;
  return false;
}
template <class D>
constexpr bool needs_corners() {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d1 != 0 && D::taps[q].d2 != 0) return true;
  return false;
}
template <class D>
constexpr bool needs_mid() {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d1 != 0) return true;
  return false;
}
template <class D>
constexpr bool needs_fast() {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d2 != 0) return true;
  return false;
}
template <class D>
constexpr int num_ext_planes() {
  int n = 0;
  for (int d = -D::rad; d <= D::rad; d++) n += plane_needs_ext<D>(d) ? 1 : 0;
  return n;
}
// slot of plane offset d0 among the extended planes (-1 if not extended)
template <class D>
constexpr int ext_slot(int d0) {
  int n = 0;
  for (int d = -D::rad; d <= D::rad; d++) {
    if (d == d0) return plane_needs_ext<D>(d) ? n : -1;
    n += plane_needs_ext<D>(d) ? 1 : 0;
  }
  return -1;
}

// index of the centre tap (offset 0, 0, 0), -1 if there is none
template <class D>
constexpr int center_tap() {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d0 == 0 && D::taps[q].d1 == 0 && D::taps[q].d2 == 0) return q;
  return -1;
}
// Radius-1 stars (a centre and its 2*dim face neighbours): when the runtime
// coefficients of all face taps are equal, each cell's product c * u is the
// same rounded value in all 2*dim sums that use it, so it can be formed once
// per cell (the shared-product form of the sweep).
template <class D>
constexpr bool share_capable() {
  return D::linear && D::rad == 1 && !needs_corners<D>() && center_tap<D>() >= 0 && D::ntaps == 2 * D::dim + 1;
}
// the tap whose coefficient the face taps share
template <class D>
constexpr int share_tap() {
  return center_tap<D>() == 0 ? 1 : 0;
}

// Compile-time loop over [B, E): f(std::integral_constant<int, i>) in order.
template <int B, class F, int... Is>
__host__ __device__ __forceinline__ void static_for_impl(F&& f, std::integer_sequence<int, Is...>) {
  (f(std::integral_constant<int, B + Is>{}), ...);
}
template <int B, int E, class F>
__host__ __device__ __forceinline__ void static_for(F&& f) {
  if constexpr (B < E) static_for_impl<B>(f, std::make_integer_sequence<int, E - B>{});
}

}  // namespace an5d

// Which kernel shapes are compiled for each stencil x dtype.
//
// A shape fixes CK (cells per lane along the fastest dim), CJ (rows per lane
// along the middle dim, 3D), NW (waves per workgroup) and the deepest
// temporal block BTMAX; every bt in [1, BTMAX] is instantiated so that the
// remainder launches of any schedule exist (the reference generates
// kernel0_1..kernel0_bt for the same reason).  The choice keeps the
// VGPR-resident pipeline below ~256 registers:
//   rings  BT * (2R+1) * CJ * CK   +   extended planes (CJ+2R) * (CK+2R) * NX.
#pragma once

#include "stencil_defs.hpp"

namespace an5d {

struct Shape {
  int ck, cj, nw, btmax;
};

constexpr int kMaxShapes = 5;

template <class D>
constexpr bool is_box() {
  return needs_corners<D>();
}

// Shapes with btmax == 0 are unused slots.
template <int ID, typename T>
constexpr Shape sweep_shape(int which) {
  using D = Def<ID>;
  constexpr bool f64 = sizeof(T) == 8;
  constexpr int R = D::rad;
  if constexpr (D::dim == 2) {
    // 16-byte runs per lane where the radius allows it
    constexpr int ck = f64 ? (R <= 2 ? 2 : 4) : 4;
    int bt = 0;
    if (R == 1) bt = is_box<D>() || !D::linear ? 8 : 12;
    if (R == 2) bt = is_box<D>() ? 4 : 8;
    if (R == 3) bt = is_box<D>() ? 3 : 6;
    if (R == 4) bt = is_box<D>() ? 2 : 4;
    if (f64 && bt > 8) bt = 8;
    if (which == 0) return Shape{ck, 1, 4, bt};
    if (which == 1 && !f64 && R == 1) return Shape{8, 1, 4, is_box<D>() ? 4 : 8};
    if (which == 2 && !f64 && R == 1 && D::linear && !is_box<D>()) return Shape{2, 1, 4, 8};
    return Shape{0, 0, 0, 0};
  } else {
    if constexpr (R == 1) {
      if (is_box<D>()) {
        if (which == 0) return Shape{2, 4, 8, 4};
        if (which == 1) return Shape{4, 2, 8, 3};
        if (which == 2 && !f64) return Shape{3, 4, 8, 3};  // 512 = three 192-cell tiles
        return Shape{0, 0, 0, 0};
      }
      if (which == 0) return Shape{2, 4, 8, 6};
      if (which == 1) return Shape{4, 4, 4, 4};
      if (which == 2) return Shape{2, 2, 16, 6};
      // 192-cell rows (3 per lane): 512 = 3 tiles of 186 at bt 3 instead of
      // 5 of 122 (FP32, within 128 VGPRs)
      if (which == 3 && !f64) return Shape{3, 2, 16, 3};
      // the same 192 x 32 tile in 8 waves of 4 rows: 256 VGPRs per wave at one
      // workgroup per CU (room for the shared-product form's second ring)
      if (which == 4 && !f64) return Shape{3, 4, 8, 3};
      return Shape{0, 0, 0, 0};
    }
    if constexpr (R == 2) {
      if (is_box<D>()) return Shape{0, 0, 0, 0};  // direct kernel only
      if (which == 0) return Shape{2, 4, 8, 3};
      if (which == 1 && !f64) return Shape{3, 4, 8, 2};
      return Shape{0, 0, 0, 0};
    }
    if constexpr (R >= 3) {
      if (is_box<D>()) return Shape{0, 0, 0, 0};
      // FP32: 8 waves of 3-4 rows keep the (2R+1)-plane rings in 256 VGPRs
      // (16-wave workgroups cap a wave at 128 and spill); FP64 keeps the
      // 16-wave shape
      if (!f64 && which == 0) return R == 3 ? Shape{4, 4, 8, 1} : Shape{4, 3, 8, 1};
      if (which == (f64 ? 0 : 1)) return Shape{4, 2, 16, R == 3 ? 2 : 1};
      return Shape{0, 0, 0, 0};
    }
    return Shape{0, 0, 0, 0};
  }
}

// cells per thread of the direct (one step per launch) kernel
template <int ID, typename T>
constexpr int direct_ck() {
  return Def<ID>::ntaps > 100 ? 1 : 2;
}

}  // namespace an5d

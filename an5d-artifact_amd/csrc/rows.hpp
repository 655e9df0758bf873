// One step per launch for the 2D boxes, streamed by rows with per-row partial
// sums (the reference's associative "partial-sum" code shape for box
// stencils, re-designed for gfx950).
//
// The temporally blocked sweep keeps 2R+1 input rows per stage in registers,
// each extended by R cells on both sides, which for radius >= 2 boxes (25 to
// 81 taps) exceeds the register file and spills.  Here each input row is read
// once and, extended by DPP, adds its 2R+1 row-convolutions into the partial
// sums of the 2R+1 output rows it touches; an output row is complete (and
// stored) after its last input row.  Registers: (2R+1) x CK partial sums and
// one extended row, instead of (2R+1) extended rows per stage.  The row loop
// is unrolled by 2R+1, so each output row keeps one register slot from its
// first input row to its store (no moves between slots).
//
// Exactness: the sources list a box's taps row-major (d0, then d2, both
// ascending: box2d4r.c:14-108), and the input rows arrive in ascending d0, so
// every output accumulates the same rounded products in the same order as the
// C expression (first term assigned, the others added; -ffp-contract=off):
// the results are the C loop's bit for bit.  FMA mode chains fused
// multiply-adds with the folded coefficients in the same order.
//
// Launch layout: wave tiles of 64*CK cells along the row that do not overlap
// (lanes 0 and 63 read the R cells beyond the tile edge themselves),
// segments of sl rows, four waves per workgroup.  The coefficients live in
// LDS and are read one box row at a time (81 FP64 values would not fit the
// registers next to the partial sums).
#pragma once

#include "engine.hpp"
#include "sweep.hpp"

namespace an5d {

// Taps ordered row-major (d0 then d2 ascending) over the full (2R+1)^2 box:
// the condition for the partial-sum form to keep the source order.
template <class D>
constexpr bool rows_order() {
  if (D::dim != 2 || !D::linear) return false;
  constexpr int R = D::rad, P = 2 * R + 1;
  if (D::ntaps != P * P) return false;
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d0 != q / P - R || D::taps[q].d2 != q % P - R || D::taps[q].d1 != 0) return false;
  return true;
}

template <int ID, typename T, int CK, int NW, bool EXACT, int DM>
__device__ __forceinline__ void rows_body(const SweepArgs<T>& a) {
#pragma clang fp contract(off)
  using D = Def<ID>;
  constexpr int R = D::rad;
  constexpr int P = 2 * R + 1;
  constexpr int NT = D::ntaps;
  constexpr unsigned SZ = sizeof(T);
  static_assert(rows_order<D>(), "row-major box taps only");
  static_assert(CK >= R, "a lane must hold at least R cells");

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = blockIdx.x * NW + w;
  // the coefficients in LDS, read one box row at a time (broadcast reads);
  // filled before any wave of the workgroup leaves
  __shared__ T cs[NT];
  for (int q = threadIdx.x; q < NT; q += 64 * NW) cs[q] = a.coef[q];
  __syncthreads();
  if (gw >= a.ntk * a.nseg) return;
  const int tk = gw % a.ntk, seg = gw / a.ntk;
  const int kv0 = R + tk * a.SK;  // the tile's first cell: SK = 64 * CK, no overlap
  const int kl = kv0 + lane * CK;
  const int i0 = R + seg * a.sl;
  const int sle = min(a.sl, a.D0 - R - i0);
  const int nit = sle + 2 * R;

  bool kst[CK];
  bool kst_all = true;
  unsigned vclamp[CK];
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    int k = kl + kk;
    kst[kk] = k < a.D2 - R;
    kst_all = kst_all && kst[kk];
    k = k >= a.D2 ? a.D2 - 1 : k;
    vclamp[kk] = (unsigned)k * SZ;
  }
  // beyond the tile edge: lane 0 reads the R cells left of kv0, lane 63 the
  // R cells right of kv0 + SK; the other lanes read their own first cell
  // (unused)
  unsigned vedge[R];
#pragma unroll
  for (int m = 0; m < R; m++) {
    int k = lane == 0 ? kv0 - R + m : (lane == 63 ? kv0 + a.SK + m : kl);
    k = k < 0 ? 0 : (k >= a.D2 ? a.D2 - 1 : k);
    vedge[m] = (unsigned)k * SZ;
  }
  const T div = a.coef[NT];
  const T rc = a.coef[NT + 1];

  // partial sums of the P output rows in flight, one slot per row m mod P
  T acc[P][CK];
#pragma unroll
  for (int s = 0; s < P; s++)
#pragma unroll
    for (int kk = 0; kk < CK; kk++) acc[s][kk] = T(0);

  auto load_row = [&](T (&x)[CK], T (&e)[R], int n) {
    int r = i0 - R + n;
    r = r < 0 ? 0 : (r >= a.D0 ? a.D0 - 1 : r);
    const auto rs = make_rsrc(a.src + (long long)r * a.s0);
#pragma unroll
    for (int kk = 0; kk < CK; kk++) x[kk] = bload<false, T>(rs, vclamp[kk], 0);
#pragma unroll
    for (int m = 0; m < R; m++) e[m] = bload<false, T>(rs, vedge[m], 0);
  };
  T nxt[CK], nxe[R];
  load_row(nxt, nxe, 0);
  // P rows per trip: input row n = n0 + U adds box row DI into the sums of
  // output n - DI, held in acc[(U - DI) mod P] (n0 is a multiple of P), so the
  // sums never move between registers
  for (int n0 = 0; n0 < nit; n0 += P) {
    static_for<0, P>([&](auto uc) {
      constexpr int U = decltype(uc)::value;
      const int n = n0 + U;
      if (n < nit) {
        T x[CK], e[R];
#pragma unroll
        for (int kk = 0; kk < CK; kk++) x[kk] = nxt[kk];
#pragma unroll
        for (int m = 0; m < R; m++) e[m] = nxe[m];
        load_row(nxt, nxe, n + 1);  // one row ahead (the last is a clamped stand-in)
        T xe[CK + 2 * R];
#pragma unroll
        for (int kk = 0; kk < CK; kk++) xe[kk + R] = x[kk];
#pragma unroll
        for (int m = 1; m <= R; m++) {
          const T l = from_left(x[CK - m]);
          const T rr = from_right(x[m - 1]);
          xe[R - m] = lane == 0 ? e[R - m] : l;
          xe[R + CK - 1 + m] = lane == 63 ? e[m - 1] : rr;
        }
        // box row DI's coefficients are in crb[DI & 1]; row DI + 1's are read
        // while row DI is summed.  Their offsets pass through an empty asm, so
        // the reads stay in the loop (81 FP64 values would not fit next to the
        // partial sums) and none is issued before the previous row's sums.
        T crb[2][P];
        auto read_coef = [&](T (&cr)[P], int di) {
          int off = di * P;
          asm volatile("" : "+s"(off));
#pragma unroll
          for (int d = 0; d < P; d++) cr[d] = cs[off + d];
        };
        read_coef(crb[0], 0);
        static_for<0, P>([&](auto dc) {
          constexpr int DI = decltype(dc)::value;
          if constexpr (DI + 1 < P) read_coef(crb[(DI + 1) & 1], DI + 1);
          const T(&cr)[P] = crb[DI & 1];
          // tap by tap over the CK cells: each cell's sum keeps the source order
          static_for<0, P>([&](auto ec) {
            constexpr int E = decltype(ec)::value;
            constexpr TapDef t = D::taps[DI * P + E];
#pragma unroll
            for (int kk = 0; kk < CK; kk++) {
              T& s = acc[(U - DI + P) % P][kk];
              const T v = xe[kk + E];
              if constexpr (EXACT) {
                const T prod = cr[E] * v;
                if constexpr (DI == 0 && E == 0) s = t.sign > 0 ? prod : -prod;
                else if constexpr (t.sign > 0) s = s + prod;
                else s = s - prod;
              } else {
                if constexpr (DI == 0 && E == 0) s = cr[E] * v;
                else s = fma_t(cr[E], v, s);
              }
            }
          });
          // the sums of this box row are formed here: without this the compiler
          // sinks them below the store branch, and the reads of every box row's
          // coefficients then sit together at the top of the loop (81 values
          // live at once: for box2d4r FP64, 284 registers and one wave per SIMD
          // instead of about 150 and three)
#pragma unroll
          for (int kk = 0; kk < CK; kk++) asm volatile("" : "+v"(acc[(U - DI + P) % P][kk]));
        });
        // output m = n - 2R (slot (U + 1) mod P) received its last row
        const int m = n - 2 * R;
        if (m >= 0) {
          T out[CK];
#pragma unroll
          for (int kk = 0; kk < CK; kk++) {
            if constexpr (DM > 0) out[kk] = div_by(acc[(U + 1) % P][kk], div, rc, DM);
            else out[kk] = acc[(U + 1) % P][kk];
          }
          T* row = a.dst + (long long)(i0 + m) * a.s0 + kl;
          if (kst_all) {
            if (a.nt_store) gstore_run<true, T, CK>(row, out);
            else gstore_run<false, T, CK>(row, out);
          } else {
#pragma unroll
            for (int kk = 0; kk < CK; kk++)
              if (kst[kk]) row[kk] = out[kk];
          }
        }
      }
    });
  }
}

template <int ID, typename T, int CK, int NW, bool EXACT>
__device__ __forceinline__ void rows_dispatch(const SweepArgs<T>& a) {
  if constexpr (!EXACT) {
    rows_body<ID, T, CK, NW, false, 0>(a);
  } else {
    if (a.has_div == 2) rows_body<ID, T, CK, NW, true, 2>(a);
    else if (a.has_div == 1) rows_body<ID, T, CK, NW, true, 1>(a);
    else rows_body<ID, T, CK, NW, true, 0>(a);
  }
}

template <int ID, typename T, int CK, int NW, bool EXACT>
__global__ __launch_bounds__(64 * NW) void rows_kernel(SweepArgs<T> a) {
  rows_dispatch<ID, T, CK, NW, EXACT>(a);
}

// Register-capped twin (at least WPE waves per SIMD): the bound lets the
// register allocator aim for more waves than the workgroup size alone asks
// for (box2d4r FP64 exact: 127 VGPRs and four waves per SIMD, against 156 and
// three).  The tuner measures both.
template <int ID, typename T, int CK, int NW, bool EXACT, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void rows_kernel_cap(
    SweepArgs<T> a) {
  rows_dispatch<ID, T, CK, NW, EXACT>(a);
}


// ---- 3D: plane-streamed partial sums for the radius-1 boxes -----------------
//
// The register sweep of a 27-point box keeps three extended planes per stage
// (237 VGPRs for j3d27pt FP64 at one step: two waves per SIMD) and
// recomputes the overlap of its tiles.  Here, one step per launch, each
// input plane is read once and adds its in-plane taps into the partial sums
// of the 2R+1 output planes it touches, as the 2D kernel above does with
// rows:
//   - a wave owns CJ rows x 64*CK cells of every plane of its segment
//     (workgroups of NW waves stacked along dim 1, the plane kernel's tiles:
//     nothing is recomputed); it reads its rows plus R rows above and below
//     straight from memory (neighbouring waves' rows, mostly from L2), so the
//     waves need no LDS exchange and no barrier;
//   - cells of the lanes' neighbours come by DPP; lanes 0 and 63 read their
//     R cells beyond the tile edge themselves;
//   - partial sums: (2R+1) x CJ x CK per lane; a plane's coefficients are
//     read from LDS one dim-0 offset at a time.
// Exactness: the taps must be grouped by plane in ascending d0 (j3d27pt.c
// and box3d1r.c list the d0 = -1 terms, then 0, then +1); within a plane the
// source order is kept, so each output adds the same rounded products in the
// same order as the C expression.
template <class D>
constexpr bool rows3_order() {
  if (D::dim != 3 || !D::linear) return false;
  for (int q = 1; q < D::ntaps; q++)
    if (D::taps[q].d0 < D::taps[q - 1].d0) return false;
  return true;
}

// first tap with d0 >= v (taps grouped by plane)
template <class D>
constexpr int first_tap_from(int v) {
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d0 >= v) return q;
  return D::ntaps;
}

template <int ID, typename T, int CK, int CJ, int NW, bool EXACT, int DM>
__device__ __forceinline__ void rows3_body(const SweepArgs<T>& a) {
#pragma clang fp contract(off)
  using D = Def<ID>;
  constexpr int R = D::rad;
  constexpr int P = 2 * R + 1;
  constexpr int NT = D::ntaps;
  constexpr unsigned SZ = sizeof(T);
  constexpr int TK = 64 * CK, TJ = NW * CJ;
  constexpr int LR = CJ + 2 * R;  // rows a wave reads per plane
  static_assert(rows3_order<D>(), "taps grouped by plane in ascending d0 only");
  static_assert(CK >= R, "a lane must hold at least R cells");

  __shared__ T cs[NT];
  for (int q = threadIdx.x; q < NT; q += 64 * NW) cs[q] = a.coef[q];
  __syncthreads();

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tk = b % a.ntk;
  b /= a.ntk;
  const int tj = b % a.ntj;
  const int seg = b / a.ntj;
  const int k0 = R + tk * TK;           // first cell of the tile
  const int j0 = R + tj * TJ + w * CJ;  // first own row of the wave
  const int i0 = R + seg * a.sl;
  const int sle = min(a.sl, a.D0 - R - i0);
  if (sle <= 0) return;  // the whole workgroup: uniform
  const int nit = sle + 2 * R;
  const int kl = k0 + lane * CK;

  bool kst[CK];
  bool kst_all = true;
  unsigned vk[CK];
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    int k = kl + kk;
    kst[kk] = k < a.D2 - R;
    kst_all = kst_all && kst[kk];
    k = k >= a.D2 ? a.D2 - 1 : k;
    vk[kk] = (unsigned)k * SZ;
  }
  // beyond the tile edge: lane 0 the R cells left of k0, lane 63 the R cells
  // right of k0 + TK; the other lanes read their own first cell (unused)
  unsigned ve[R];
#pragma unroll
  for (int m = 0; m < R; m++) {
    int k = lane == 0 ? k0 - R + m : (lane == 63 ? k0 + TK + m : kl);
    k = k < 0 ? 0 : (k >= a.D2 ? a.D2 - 1 : k);
    ve[m] = (unsigned)k * SZ;
  }
  const T div = a.coef[NT];
  const T rc = a.coef[NT + 1];

  // acc[d] holds output plane m = n - d while input plane n is added
  T acc[P][CJ][CK];
#pragma unroll
  for (int d = 0; d < P; d++)
#pragma unroll
    for (int jj = 0; jj < CJ; jj++)
#pragma unroll
      for (int kk = 0; kk < CK; kk++) acc[d][jj][kk] = T(0);

  struct Plane {
    T x[LR][CK];
    T e[LR][R];
  };
  auto load_plane = [&](Plane& pl, int n) {
    int p = i0 - R + n;
    p = p < 0 ? 0 : (p >= a.D0 ? a.D0 - 1 : p);
    const T* base = a.src + (long long)p * a.s0;
#pragma unroll
    for (int r = 0; r < LR; r++) {
      int j = j0 - R + r;
      j = j < 0 ? 0 : (j >= a.D1 ? a.D1 - 1 : j);
      const auto rs = make_rsrc(base + (long long)j * a.s1);
#pragma unroll
      for (int kk = 0; kk < CK; kk++) pl.x[r][kk] = bload<false, T>(rs, vk[kk], 0);
#pragma unroll
      for (int m = 0; m < R; m++) pl.e[r][m] = bload<false, T>(rs, ve[m], 0);
    }
  };

  Plane nxt;
  load_plane(nxt, 0);
  for (int n = 0; n < nit; n++) {
    Plane cur = nxt;
    load_plane(nxt, n + 1);  // one plane ahead (the last is a clamped stand-in)
    // extended rows: xe[r][c] is cell kl - R + c of row j0 - R + r
    T xe[LR][CK + 2 * R];
#pragma unroll
    for (int r = 0; r < LR; r++) {
#pragma unroll
      for (int kk = 0; kk < CK; kk++) xe[r][kk + R] = cur.x[r][kk];
#pragma unroll
      for (int m = 1; m <= R; m++) {
        const T l = from_left(cur.x[r][CK - m]);
        const T rr = from_right(cur.x[r][m - 1]);
        xe[r][R - m] = lane == 0 ? cur.e[r][R - m] : l;
        xe[r][R + CK - 1 + m] = lane == 63 ? cur.e[r][m - 1] : rr;
      }
    }
    static_for<0, P>([&](auto dc) {
      constexpr int DI = decltype(dc)::value;  // this plane is d0 = DI - R of output n - DI
      constexpr int QA = first_tap_from<D>(DI - R), QB = first_tap_from<D>(DI - R + 1);
      constexpr int NQ = QB - QA;
      if constexpr (NQ > 0) {
        // the block's coefficients through an offset the compiler cannot
        // hoist out of the loop
        int off = QA;
        asm volatile("" : "+s"(off));
        T cr[NQ];
#pragma unroll
        for (int q = 0; q < NQ; q++) cr[q] = cs[off + q];
        static_for<0, NQ>([&](auto qc) {
          constexpr int QI = decltype(qc)::value;
          constexpr TapDef t = D::taps[QA + QI];
#pragma unroll
          for (int jj = 0; jj < CJ; jj++)
#pragma unroll
            for (int kk = 0; kk < CK; kk++) {
              T& s = acc[DI][jj][kk];
              const T v = xe[jj + R + t.d1][kk + R + t.d2];
              if constexpr (EXACT) {
                const T prod = cr[QI] * v;
                if constexpr (QA + QI == 0) s = t.sign > 0 ? prod : -prod;
                else if constexpr (t.sign > 0) s = s + prod;
                else s = s - prod;
              } else {
                if constexpr (QA + QI == 0) s = cr[QI] * v;
                else s = fma_t(cr[QI], v, s);
              }
            }
        });
        // formed here, not sunk below the store branch
#pragma unroll
        for (int jj = 0; jj < CJ; jj++)
#pragma unroll
          for (int kk = 0; kk < CK; kk++) asm volatile("" : "+v"(acc[DI][jj][kk]));
      }
    });
    // output plane m = n - 2R received its last input plane
    const int m = n - 2 * R;
    if (m >= 0) {
#pragma unroll
      for (int jj = 0; jj < CJ; jj++) {
        if (j0 + jj >= a.D1 - R) continue;  // uniform
        T out[CK];
#pragma unroll
        for (int kk = 0; kk < CK; kk++) {
          if constexpr (DM > 0) out[kk] = div_by(acc[P - 1][jj][kk], div, rc, DM);
          else out[kk] = acc[P - 1][jj][kk];
        }
        T* row = a.dst + (long long)(i0 + m) * a.s0 + (long long)(j0 + jj) * a.s1 + kl;
        if (kst_all) {
          if (a.nt_store) gstore_run<true, T, CK>(row, out);
          else gstore_run<false, T, CK>(row, out);
        } else {
#pragma unroll
          for (int kk = 0; kk < CK; kk++)
            if (kst[kk]) row[kk] = out[kk];
        }
      }
    }
#pragma unroll
    for (int d = P - 1; d > 0; d--)
#pragma unroll
      for (int jj = 0; jj < CJ; jj++)
#pragma unroll
        for (int kk = 0; kk < CK; kk++) acc[d][jj][kk] = acc[d - 1][jj][kk];
  }
}

template <int ID, typename T, int CK, int CJ, int NW, bool EXACT>
__global__ __launch_bounds__(64 * NW) void rows3_kernel(SweepArgs<T> a) {
  if constexpr (!EXACT) {
    rows3_body<ID, T, CK, CJ, NW, false, 0>(a);
  } else {
    if (a.has_div == 2) rows3_body<ID, T, CK, CJ, NW, true, 2>(a);
    else if (a.has_div == 1) rows3_body<ID, T, CK, CJ, NW, true, 1>(a);
    else rows3_body<ID, T, CK, CJ, NW, true, 0>(a);
  }
}

}  // namespace an5d

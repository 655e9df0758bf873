// Instantiates and registers the forward kernels of one stencil x dtype.
// Built once per pair with -DAN5D_SID=<stencil id> -DAN5D_F64=<0|1>, so the
// 42 translation units compile in parallel.
#include "engine.hpp"
#include "plane.hpp"
#include "rows.hpp"
#include "sweep.hpp"
#include "variants.hpp"

#ifndef AN5D_SID
#error "AN5D_SID must be defined"
#endif

namespace an5d {
namespace {

using T = std::conditional_t<AN5D_F64, double, float>;
constexpr int SID = AN5D_SID;

template <typename U>
SweepArgs<U> to_args(const LaunchArgs& a) {
  SweepArgs<U> s;
  s.src = static_cast<const U*>(a.src);
  s.dst = static_cast<U*>(a.dst);
  s.coef = static_cast<const U*>(a.coef);
  s.gscale = a.gscale;
  s.has_div = a.has_div;
  s.D0 = a.D0; s.D1 = a.D1; s.D2 = a.D2;
  s.s0 = a.s0; s.s1 = a.s1;
  s.ntk = a.ntk; s.ntj = a.ntj; s.nseg = a.nseg; s.sl = a.sl;
  s.SK = a.SK; s.SJ = a.SJ;
  s.mid = static_cast<U*>(a.mid);
  s.mid_stride = a.mid_stride;
  s.mid_mask = a.mid_mask;
  s.nt_store = a.nt_store;
  return s;
}

constexpr int kUnroll = default_unroll<Def<SID>::rad>();
constexpr bool kShare = share_capable<Def<SID>>();

// The kernel of one instance: WPE > 0 selects the register-capped form.
template <int BT, int CK, int CJ, int NW, bool EXACT, bool MIDS, bool SHARE, int WPE>
constexpr auto sweep_fn() {
  if constexpr (WPE > 0) return &sweep_kernel_cap<SID, T, BT, CK, CJ, NW, EXACT, WPE>;
  else return &sweep_kernel<SID, T, BT, CK, CJ, NW, EXACT, MIDS, kUnroll, SHARE>;
}

template <int BT, int CK, int CJ, int NW, bool EXACT, bool MIDS = false, bool SHARE = false, int WPE = 0>
hipError_t launch_sweep(const LaunchArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  constexpr auto kern = sweep_fn<BT, CK, CJ, NW, EXACT, MIDS, SHARE, WPE>();
  if (lds > 65536) {
    // gfx950 has 160 KB of LDS per CU
    static std::atomic<int> done[kMaxDevices];
    const hipError_t e = lds_opt_in(reinterpret_cast<const void*>(kern), lds, done);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL(kern, grid, block, lds, st, to_args<T>(a));
  return hipGetLastError();
}

// Private bytes per thread of the shared-product kernel (its register
// spills), asked once; -1 if the query fails.
template <int BT, int CK, int CJ, int NW>
int share_scratch_once() {
  static const int n = [] {
    hipFuncAttributes attr;
    const void* k = reinterpret_cast<const void*>(&sweep_kernel<SID, T, BT, CK, CJ, NW, true, false, kUnroll, true>);
    if (hipFuncGetAttributes(&attr, k) != hipSuccess) return -1;
    return (int)attr.localSizeBytes;
  }();
  return n;
}

// Resident workgroups per CU (0 if the query fails); computed once, on the
// first device that asks (every device is a gfx950 with the same resources).
template <int BT, int CK, int CJ, int NW, bool EXACT, bool SHARE = false, int WPE = 0>
int occupancy_once() {
  static const int n = [] {
    const void* k = reinterpret_cast<const void*>(sweep_fn<BT, CK, CJ, NW, EXACT, false, SHARE, WPE>());
    const size_t lds = sweep_lds_bytes<SID, T, CK, CJ, NW>(BT);
    static std::atomic<int> done[kMaxDevices];
    if (lds > 65536 && lds_opt_in(k, lds, done) != hipSuccess) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64 * NW, lds) != hipSuccess) return 0;
    return blocks;
  }();
  return n;
}
template <int BT, int CK, int CJ, int NW, int WPE = 0>
int occupancy_sweep(int mode) {
  if constexpr (kShare && WPE == 0)
    if (mode == 2) return occupancy_once<BT, CK, CJ, NW, true, true>();
  return mode == 1 ? occupancy_once<BT, CK, CJ, NW, false, false, WPE>()
                   : occupancy_once<BT, CK, CJ, NW, true, false, WPE>();
}

template <int CK>
hipError_t launch_direct(const LaunchArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((direct_kernel<SID, T, CK>), grid, block, lds, st, to_args<T>(a));
  return hipGetLastError();
}

template <int CK, int CJ, int NW, bool EXACT>
hipError_t launch_plane(const LaunchArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  // the kernel keeps in-plane offsets in 32 bits
  if ((long long)a.D1 * a.s1 >= (1LL << 31)) return hipErrorInvalidValue;
  if (lds > 65536) {
    static std::atomic<int> done[kMaxDevices];
    const hipError_t e =
        lds_opt_in(reinterpret_cast<const void*>(&plane_kernel<SID, T, CK, CJ, NW, EXACT>), lds, done);
    if (e != hipSuccess) return e;
  }
  hipLaunchKernelGGL((plane_kernel<SID, T, CK, CJ, NW, EXACT>), grid, block, lds, st, to_args<T>(a));
  return hipGetLastError();
}

template <int CK, int CJ, int NW, bool EXACT>
int occupancy_plane_once() {
  static const int n = [] {
    const void* k = reinterpret_cast<const void*>(&plane_kernel<SID, T, CK, CJ, NW, EXACT>);
    constexpr size_t lds = plane_lds_bytes<T, CK, CJ, NW, Def<SID>::rad>();
    static std::atomic<int> done[kMaxDevices];
    if (lds > 65536 && lds_opt_in(k, lds, done) != hipSuccess) return 0;
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64 * NW, lds) != hipSuccess) return 0;
    return blocks;
  }();
  return n;
}
template <int CK, int CJ, int NW>
int occupancy_plane(int mode) {
  return mode ? occupancy_plane_once<CK, CJ, NW, false>() : occupancy_plane_once<CK, CJ, NW, true>();
}

template <int CK, bool EXACT, int WPE>
constexpr auto rows_fn() {
  if constexpr (WPE > 0) return &rows_kernel_cap<SID, T, CK, 4, EXACT, WPE>;
  else return &rows_kernel<SID, T, CK, 4, EXACT>;
}

template <int CK, bool EXACT, int WPE>
hipError_t launch_rows(const LaunchArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  constexpr auto kern = rows_fn<CK, EXACT, WPE>();
  hipLaunchKernelGGL(kern, grid, block, 0, st, to_args<T>(a));
  return hipGetLastError();
}

template <int CK, bool EXACT, int WPE>
int occupancy_rows_once() {
  static const int n = [] {
    int blocks = 0;
    const void* k = reinterpret_cast<const void*>(rows_fn<CK, EXACT, WPE>());
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64 * 4, 0) != hipSuccess) return 0;
    return blocks;
  }();
  return n;
}
template <int CK, int WPE>
int occupancy_rows(int mode) {
  return mode == 1 ? occupancy_rows_once<CK, false, WPE>() : occupancy_rows_once<CK, true, WPE>();
}

// Row-streamed partial-sum kernels for the 2D boxes of radius >= 2 (whose
// register rings spill), one per run width CK, optionally register-capped.
template <int CK, int WPE = 0>
void add_rows(std::vector<Variant>& out) {
  if constexpr (rows_order<Def<SID>>() && Def<SID>::rad >= 2 && CK >= Def<SID>::rad) {
    Variant v{};
    v.stencil = SID;
    v.dtype = AN5D_F64 ? kF64 : kF32;
    v.kind = kRows;
    v.ck = CK; v.cj = 1; v.nw = 4; v.btmax = 1;
    v.regcap = WPE > 0 ? 512 / WPE : 0;
    v.fn[1] = &launch_rows<CK, true, WPE>;
    v.fn_fast[1] = &launch_rows<CK, false, WPE>;
    v.occ[1] = &occupancy_rows<CK, WPE>;
    out.push_back(v);
  }
}

template <int CK, int CJ, int NW, bool EXACT>
hipError_t launch_rows3(const LaunchArgs& a, dim3 grid, dim3 block, size_t lds, hipStream_t st) {
  hipLaunchKernelGGL((rows3_kernel<SID, T, CK, CJ, NW, EXACT>), grid, block, 0, st, to_args<T>(a));
  return hipGetLastError();
}

template <int CK, int CJ, int NW, bool EXACT>
int occupancy_rows3_once() {
  static const int n = [] {
    int blocks = 0;
    const void* k = reinterpret_cast<const void*>(&rows3_kernel<SID, T, CK, CJ, NW, EXACT>);
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, k, 64 * NW, 0) != hipSuccess) return 0;
    return blocks;
  }();
  return n;
}
template <int CK, int CJ, int NW>
int occupancy_rows3(int mode) {
  return mode == 1 ? occupancy_rows3_once<CK, CJ, NW, false>() : occupancy_rows3_once<CK, CJ, NW, true>();
}

// Plane-streamed partial sums for the radius-1 3D boxes (taps grouped by
// plane): one step per launch on the plane kernel's whole tiles (nothing
// recomputed), so they are planned like it (kind kPlane).
template <int CK, int CJ, int NW>
void add_rows3(std::vector<Variant>& out) {
  if constexpr (rows3_order<Def<SID>>() && is_box<Def<SID>>() && Def<SID>::rad == 1) {
    Variant v{};
    v.stencil = SID;
    v.dtype = AN5D_F64 ? kF64 : kF32;
    v.kind = kPlane;
    v.ck = CK; v.cj = CJ; v.nw = NW; v.btmax = 1;
    v.fn[1] = &launch_rows3<CK, CJ, NW, true>;
    v.fn_fast[1] = &launch_rows3<CK, CJ, NW, false>;
    v.occ[1] = &occupancy_rows3<CK, CJ, NW>;
    out.push_back(v);
  }
}

// The shape's btmax lowered until the exchange buffers fit gfx950's 160 KB of LDS.
template <int W>
constexpr Shape fitted_shape() {
  Shape sh = sweep_shape<SID, T>(W);
  while (sh.btmax > 0 && sweep_lds_bytes<SID, T, sweep_shape<SID, T>(W).ck, sweep_shape<SID, T>(W).cj,
                                         sweep_shape<SID, T>(W).nw>(sh.btmax) > 160 * 1024)
    sh.btmax--;
  return sh;
}

template <int W>
void add_sweep(std::vector<Variant>& out) {
  constexpr Shape sh = fitted_shape<W>();
  if constexpr (sh.btmax > 0) {
    Variant v{};
    v.stencil = SID;
    v.dtype = AN5D_F64 ? kF64 : kF32;
    v.kind = kSweep;
    v.ck = sh.ck; v.cj = sh.cj; v.nw = sh.nw; v.btmax = sh.btmax;
    static_for<1, sh.btmax + 1>([&](auto bc) {
      constexpr int BT = decltype(bc)::value;
      v.lds[BT] = sweep_lds_bytes<SID, T, sh.ck, sh.cj, sh.nw>(BT);
      v.fn[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, true>;
      v.fn_fast[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, false>;
      if constexpr (kShare) {
        v.fn_share[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, true, false, true>;
        v.share_scratch[BT] = &share_scratch_once<BT, sh.ck, sh.cj, sh.nw>;
      }
      v.occ[BT] = &occupancy_sweep<BT, sh.ck, sh.cj, sh.nw>;
      // intermediate-stage outputs: the adjoint's checkpointing forward pass (2D, exact)
      if constexpr (Def<SID>::dim == 2 && BT >= 2 && BT <= kMaxMidBT)
        v.fn_mid[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, true, true>;
    });
    out.push_back(v);
  }
}

// Register-capped twin of shape W (the cap as a search dimension): workgroups
// of 4 waves are capped at 2 waves per SIMD (256 registers) where the rings
// are wide (FP64, radius >= 2) and at 4 (128) otherwise, workgroups of 8 at 4
// (128); 16-wave workgroups are at 128 already and have no twin.  The deeper
// half of the temporal blocks (where the registers run out) takes the capped
// kernels; the shallower ones, and the intermediate-stage launches, share the
// uncapped shape's kernels.
template <int W>
void add_capped(std::vector<Variant>& out) {
  constexpr Shape sh = fitted_shape<W>();
  if constexpr (sh.btmax > 0 && sh.nw <= 8) {
    constexpr int wpe = sh.nw == 4 && (AN5D_F64 || Def<SID>::rad >= 2) ? 2 : 4;
    Variant v{};
    v.stencil = SID;
    v.dtype = AN5D_F64 ? kF64 : kF32;
    v.kind = kSweep;
    v.ck = sh.ck; v.cj = sh.cj; v.nw = sh.nw; v.btmax = sh.btmax;
    v.regcap = 512 / wpe;
    static_for<1, sh.btmax + 1>([&](auto bc) {
      constexpr int BT = decltype(bc)::value;
      constexpr int C = 2 * BT > sh.btmax ? wpe : 0;
      v.lds[BT] = sweep_lds_bytes<SID, T, sh.ck, sh.cj, sh.nw>(BT);
      v.fn[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, true, false, false, C>;
      v.fn_fast[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, false, false, false, C>;
      v.occ[BT] = &occupancy_sweep<BT, sh.ck, sh.cj, sh.nw, C>;
      if constexpr (Def<SID>::dim == 2 && BT >= 2 && BT <= kMaxMidBT)
        v.fn_mid[BT] = &launch_sweep<BT, sh.ck, sh.cj, sh.nw, true, true>;
    });
    out.push_back(v);
  }
}

}  // namespace

#define AN5D_CAT2(a, b, c) a##b##_##c
#define AN5D_CAT(a, b, c) AN5D_CAT2(a, b, c)
void AN5D_CAT(register_variants_, AN5D_SID, AN5D_F64)(std::vector<Variant>& out) {
  add_sweep<0>(out);
  add_sweep<1>(out);
  add_sweep<2>(out);
  add_sweep<3>(out);
  add_sweep<4>(out);
  Variant d{};
  d.stencil = SID;
  d.dtype = AN5D_F64 ? kF64 : kF32;
  d.kind = kDirect;
  d.ck = direct_ck<SID, T>(); d.cj = 1; d.nw = 4; d.btmax = 1;
  for (auto& l : d.lds) l = 0;
  d.fn[1] = &launch_direct<direct_ck<SID, T>()>;
  d.fn_fast[1] = d.fn[1];
  out.push_back(d);
  // radius 2-4 3D stars: one step per launch with the plane staged in LDS
  if constexpr (plane_taps_ok<Def<SID>>() && Def<SID>::rad >= 2) {
    // 16-row x 128-cell tiles, at most 128 VGPRs; the R+2 LDS tile slots
    // take 42-157 KB
    constexpr int CK = 2, CJ = AN5D_F64 ? 1 : 2, NW = AN5D_F64 ? 16 : 8;
    static_assert(plane_lds_bytes<T, CK, CJ, NW, Def<SID>::rad>() <= 160 * 1024, "plane tiles exceed LDS");
    Variant pv{};
    pv.stencil = SID;
    pv.dtype = AN5D_F64 ? kF64 : kF32;
    pv.kind = kPlane;
    pv.ck = CK; pv.cj = CJ; pv.nw = NW; pv.btmax = 1;
    pv.lds[1] = plane_lds_bytes<T, CK, CJ, NW, Def<SID>::rad>();
    pv.fn[1] = &launch_plane<CK, CJ, NW, true>;
    pv.fn_fast[1] = &launch_plane<CK, CJ, NW, false>;
    pv.occ[1] = &occupancy_plane<CK, CJ, NW>;
    out.push_back(pv);
  }
  // register-capped twins of the sweep shapes, after the other variants so
  // the indices of those stay put
  add_capped<0>(out);
  add_capped<1>(out);
  add_capped<2>(out);
  add_capped<3>(out);
  add_capped<4>(out);
  // row-streamed partial sums (2D boxes, radius >= 2)
  add_rows<4>(out);
  if constexpr (!AN5D_F64) add_rows<8>(out);
  else add_rows<4, 2>(out);
  // plane-streamed partial sums (3D radius-1 boxes)
  add_rows3<2, 4, 4>(out);
  if constexpr (!AN5D_F64) add_rows3<1, 8, 4>(out);
  else add_rows3<2, 2, 8>(out);
}

}  // namespace an5d

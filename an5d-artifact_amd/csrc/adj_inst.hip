// Instantiates and registers the adjoint kernels of one 2D stencil x dtype
// (-DAN5D_SID=<stencil id> -DAN5D_F64=<0|1>).
#include "adjoint.hpp"
#include "engine.hpp"

namespace an5d {
namespace {

using T = std::conditional_t<AN5D_F64, double, float>;
constexpr int SID = AN5D_SID;
using D = Def<SID>;

// cells per lane and deepest block: keeps the kernels within 256 VGPRs (two
// waves per SIMD)
constexpr int adj_ck() {
  if (!D::linear) return 2;
  if (D::rad <= 2 && !AN5D_F64) return 4;
  return D::rad <= 2 ? 2 : 4;  // a lane holds at least R cells
}
constexpr int adj_btmax() {
  if (!D::linear) return 2;
  if (D::rad == 1) return D::ntaps > 5 ? 4 : 5;
  if (D::rad == 2) return AN5D_F64 ? 2 : 3;  // deeper blocks exceed 256 VGPRs
  return 1;
}

template <int BT>
hipError_t launch_adj(const AdjLaunch& l, dim3 grid, dim3 block, hipStream_t st) {
  AdjArgs<T> a;
  a.U = static_cast<const T*>(l.U);
  a.Lin = static_cast<const T*>(l.Lin);
  a.Lout = static_cast<T*>(l.Lout);
  a.coef = static_cast<const T*>(l.coef);
  a.gscale = l.gscale;
  a.has_div = l.has_div;
  a.gslot = l.gslot;
  a.D0 = l.D0; a.D2 = l.D2; a.s0 = l.s0;
  a.ntk = l.ntk; a.nseg = l.nseg; a.sl = l.sl; a.SK = l.SK;
  hipLaunchKernelGGL((adjoint_kernel<SID, T, BT, adj_ck(), 4>), grid, block, 0, st, a);
  return hipGetLastError();
}

template <int BT>
int occupancy_of() {
  int n = 0;
  if (hipOccupancyMaxActiveBlocksPerMultiprocessor(
          &n, reinterpret_cast<const void*>(&adjoint_kernel<SID, T, BT, adj_ck(), 4>), 64 * 4, 0) !=
      hipSuccess)
    return 0;
  return n;
}
int occupancy(int bt) {
  int n = 0;
  static_for<1, adj_btmax() + 1>([&](auto bc) {
    constexpr int BT = decltype(bc)::value;
    if (bt == BT) n = occupancy_of<BT>();
  });
  return n;
}
hipError_t launch_seed(const void* uT, const void* target, void* lam, double* lslot, int D0, int D2,
                       long long s0, long long stgt, int R, hipStream_t st) {
  hipLaunchKernelGGL((adjoint_seed_kernel<T>), dim3(kSeedBlocks), dim3(256), 0, st, static_cast<const T*>(uT),
                     static_cast<const T*>(target), static_cast<T*>(lam), lslot, D0, D2, s0, stgt, R);
  return hipGetLastError();
}

hipError_t launch_halo(const void* src, void* dst, int D0, int D2, long long s0, int R, hipStream_t st) {
  const long long cells = 2LL * R * D2 + 2LL * R * (D0 - 2 * R);
  long long blocks = (cells + 255) / 256;
  if (blocks > 1024) blocks = 1024;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL((halo_copy_kernel<T>), dim3((unsigned)blocks), dim3(256), 0, st, static_cast<const T*>(src),
                     static_cast<T*>(dst), D0, D2, s0, R);
  return hipGetLastError();
}

}  // namespace

#define AN5D_CAT2(a, b, c) a##b##_##c
#define AN5D_CAT(a, b, c) AN5D_CAT2(a, b, c)
void AN5D_CAT(register_adjoint_, AN5D_SID, AN5D_F64)(std::vector<AdjVariant>& out) {
  AdjVariant v{};
  v.stencil = SID;
  v.dtype = AN5D_F64 ? kF64 : kF32;
  v.ck = adj_ck();
  v.nw = 4;
  v.btmax = adj_btmax();
  v.reach = adj_reach<D>();
  v.seed = &launch_seed;
  v.halo = &launch_halo;
  v.occupancy = &occupancy;
  static_for<1, adj_btmax() + 1>([&](auto bc) {
    constexpr int BT = decltype(bc)::value;
    v.fn[BT] = &launch_adj<BT>;
  });
  out.push_back(v);
}

}  // namespace an5d

// Temporally blocked stencil sweep for gfx950 (CDNA4), hand-written HIP.
//
// One launch advances the grid by BT fused timesteps: it reads buffer c0%2
// and writes the interior of buffer (c0+1)%2, as the reference's kernel0_<k>
// does (compiled/float/j2d5pt-256-10-128_kernel.cu:82-83), and leaves every
// halo cell alone (Dirichlet: a halo value is the same at every step).
//
// Work decomposition (no relation to the reference's thread layout):
//   * dim 0 (outermost) is streamed: a wave walks `sl` output planes of a
//     segment plus a 2*BT*R warm-up, holding for every stage s < BT a ring
//     of the last 2R+1 planes of u_s in VGPRs.  Per input plane it runs all
//     BT stages back to back; stage s lags the input by s*R planes.
//   * the fastest dim is owned lane-contiguously: each of the 64 lanes holds
//     CK adjacent cells, so a wave spans 64*CK cells, loads them with one
//     coalesced vector load per row and gets the R cells left/right of its
//     run from the neighbouring lane with DPP wave_shr:1 / wave_shl:1 (no
//     LDS, no barrier).  Lanes at the wave's two ends receive garbage, which
//     spreads R cells per stage: a wave therefore produces 64*CK - 2*BT*R
//     valid cells (overlapped tiling).
//   * 3D only: the middle dim is split over the NW waves of a workgroup, CJ
//     rows per lane.  Rows owned by another wave come through LDS: for stars
//     one exchange per input plane covers every stage (radius 1 with bt >= 3:
//     double-buffered, one barrier), boxes exchange per stage.  In 2D the NW
//     waves of a workgroup are independent tiles.
//   * input planes are queued PF iterations ahead in registers.
//
// Exactness: each cell of each stage is produced by the same expression, in
// the same term order, from the same operands as the C loop (no contraction:
// the translation unit is built with -ffp-contract=off), so the result is
// bitwise independent of BT, CK, CJ, NW and sl and equal to the CPU loop.
#pragma once

#include <hip/hip_runtime.h>

#include "stencil_defs.hpp"

namespace an5d {

template <typename T>
struct SweepArgs {
  const T* __restrict__ src;   // buffer c0 % 2
  T* __restrict__ dst;         // buffer (c0 + 1) % 2
  const T* __restrict__ coef;  // ntaps coefficients, the divisor, RN(1/divisor)
  double gscale;               // gradient2d: factor on the reciprocal-root term
  int has_div;                 // 0: no division, 1: IEEE x/div, 2: corrected x*RN(1/div)
  int D0, D1, D2;              // full extents incl. halo (2D: D1 == 1)
  long long s0, s1;            // element strides of dim 0 and dim 1 (dim 2 is unit)
  int ntk, ntj, nseg, sl;      // tiles along dim 2 / dim 1, segments along dim 0, segment length
  int SK, SJ;                  // output cells per tile along dim 2 / dim 1
  // Intermediate outputs (checkpoints for the adjoint): stage S < BT with bit
  // S of mid_mask set also stores its owned cells into mid + m * mid_stride,
  // m = the number of lower set bits.  mid_mask == 0: none.
  T* __restrict__ mid;
  long long mid_stride;
  unsigned mid_mask;
  int nt_store;  // 1: non-temporal stores (grid larger than the last-level cache)
};

// ---- cross-lane moves ------------------------------------------------------
// lane L receives lane L-1's value (wave_shr:1) / lane L+1's value
// (wave_shl:1).  bound_ctrl: the lane without a source reads 0, so no "old"
// operand has to be materialised first.
__device__ __forceinline__ int dpp_shr1(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x138, 0xF, 0xF, true);
}
__device__ __forceinline__ int dpp_shl1(int v) {
  return __builtin_amdgcn_mov_dpp(v, 0x130, 0xF, 0xF, true);
}
__device__ __forceinline__ float from_left(float v) {
  return __builtin_bit_cast(float, dpp_shr1(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ float from_right(float v) {
  return __builtin_bit_cast(float, dpp_shl1(__builtin_bit_cast(int, v)));
}
__device__ __forceinline__ double from_left(double v) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = dpp_shr1((int)(b & 0xffffffffLL)), hi = dpp_shr1((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double from_right(double v) {
  long long b = __builtin_bit_cast(long long, v);
  int lo = dpp_shl1((int)(b & 0xffffffffLL)), hi = dpp_shl1((int)(b >> 32));
  return __builtin_bit_cast(double, ((long long)hi << 32) | (unsigned int)lo);
}

// ---- vector memory access with 4-byte alignment only -----------------------
template <typename T, int N>
struct Vec { T v[N]; };

template <typename T, int N>
__device__ __forceinline__ void load_run(T (&dst)[N], const T* p) {
  Vec<T, N> t;
  __builtin_memcpy(&t, p, sizeof(t));
#pragma unroll
  for (int i = 0; i < N; i++) dst[i] = t.v[i];
}
template <typename T, int N>
__device__ __forceinline__ void store_run(T* p, const T (&src)[N]) {
  Vec<T, N> t;
#pragma unroll
  for (int i = 0; i < N; i++) t.v[i] = src[i];
  __builtin_memcpy(p, &t, sizeof(t));
}

// Buffer-resource loads: a wave-uniform base in the descriptor, a
// wave-uniform byte offset (soffset) and a per-lane 32-bit byte offset
// (voffset), so the streamed plane loads need no per-lane 64-bit address
// arithmetic.  The range field is left at its maximum (no clamping: every
// access is in bounds by construction); 0x00020000 is the gfx9-family word 3
// (32-bit data format, which untyped accesses ignore).  aux bit 1 =
// non-temporal.
__device__ __forceinline__ __amdgpu_buffer_rsrc_t make_rsrc(const void* base) {
  return __builtin_amdgcn_make_buffer_rsrc(const_cast<void*>(base), 0, 0xffffffff, 0x00020000);
}
template <bool NT, typename T>
__device__ __forceinline__ T bload(__amdgpu_buffer_rsrc_t r, unsigned vo, int so) {
  constexpr int aux = NT ? 2 : 0;
  if constexpr (sizeof(T) == 4) return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b32(r, vo, so, aux));
  else return __builtin_bit_cast(T, __builtin_amdgcn_raw_buffer_load_b64(r, vo, so, aux));
}

// Global accesses of 4-byte-aligned runs; NT: non-temporal (streaming) load
// or store.  The non-temporal hint pays where a launch is bandwidth-bound
// (one or two fused steps) and costs where it is compute-bound (the halo
// re-reads of overlapping tiles then miss the cache).
template <typename T, int N>
struct NtVec;
template <> struct NtVec<float, 1> { typedef float type __attribute__((aligned(4))); };
template <> struct NtVec<float, 2> { typedef float type __attribute__((ext_vector_type(2), aligned(4))); };
template <> struct NtVec<float, 3> { typedef float type __attribute__((ext_vector_type(3), aligned(4))); };
template <> struct NtVec<float, 4> { typedef float type __attribute__((ext_vector_type(4), aligned(4))); };
template <> struct NtVec<float, 8> { typedef float type __attribute__((ext_vector_type(8), aligned(4))); };
template <> struct NtVec<double, 1> { typedef double type __attribute__((aligned(4))); };
template <> struct NtVec<double, 2> { typedef double type __attribute__((ext_vector_type(2), aligned(4))); };
template <> struct NtVec<double, 3> { typedef double type __attribute__((ext_vector_type(3), aligned(4))); };
template <> struct NtVec<double, 4> { typedef double type __attribute__((ext_vector_type(4), aligned(4))); };
template <> struct NtVec<double, 8> { typedef double type __attribute__((ext_vector_type(8), aligned(4))); };

template <bool NT, typename T, int N>
__device__ __forceinline__ void gload_run(T (&dst)[N], const T* p) {
  if constexpr (NT) {
    using V = typename NtVec<T, N>::type;
    const V v = __builtin_nontemporal_load(reinterpret_cast<const V*>(p));
    if constexpr (N == 1) {
      dst[0] = v;
    } else {
#pragma unroll
      for (int i = 0; i < N; i++) dst[i] = v[i];
    }
  } else {
    load_run<T, N>(dst, p);
  }
}
template <bool NT, typename T, int N>
__device__ __forceinline__ void gstore_run(T* p, const T (&src)[N]) {
  if constexpr (NT) {
    using V = typename NtVec<T, N>::type;
    V v;
    if constexpr (N == 1) {
      v = src[0];
    } else {
#pragma unroll
      for (int i = 0; i < N; i++) v[i] = src[i];
    }
    __builtin_nontemporal_store(v, reinterpret_cast<V*>(p));
  } else {
    store_run<T, N>(p, src);
  }
}

// ---- correctly rounded division by a constant --------------------------------
// has_div == 2: q0 = x*rc, e' = fma(q0, d, -x) (exact residual, negated),
// q = fma(-e', rc, q0) with rc = RN(1/d) (Markstein's correction step; the
// negated residual makes x = -0 give -0).  When q is +-0 or normal it equals
// IEEE x/d: checked for all 2^32 floats with d = 118 and d = 159, and for
// doubles on every multiple of the subnormal quantum below 2^28 quanta plus
// 1.3e9 random inputs over all exponents.  Other q (subnormal, inf, NaN)
// take the IEEE division.  has_div == 1: IEEE.
__device__ __forceinline__ float fma_t(float a, float b, float c) { return __builtin_fmaf(a, b, c); }
__device__ __forceinline__ double fma_t(double a, double b, double c) { return __builtin_fma(a, b, c); }
// classes: -normal | -0 | +0 | +normal
__device__ __forceinline__ bool zero_or_normal(float q) { return __builtin_amdgcn_classf(q, 0x168); }
__device__ __forceinline__ bool zero_or_normal(double q) { return __builtin_amdgcn_class(q, 0x168); }
template <typename T>
__device__ __forceinline__ T div_fast(T x, T d, T rc) {
  const T q0 = x * rc;
  const T ep = fma_t(q0, d, -x);
  return fma_t(-ep, rc, q0);
}
template <typename T>
__device__ __forceinline__ T div_by(T x, T d, T rc, int mode) {
  if (mode == 0) return x;
  if (mode == 2) {
    T q = div_fast(x, d, rc);
    // wave-uniform test keeps the IEEE fallback a skipped branch
    const bool ok = zero_or_normal(q);
    if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
      if (!ok) q = x / d;
    }
    return q;
  }
  return x / d;
}

// Fast mode (EXACT == false): tolerance-checked, not bitwise.  Linear
// stencils take pre-folded coefficients s_q*c_q/div and chain FMAs.  (The
// nonlinear gradient2d keeps the exact evaluation in both modes.)
template <class D, typename T, class Get>
__device__ __forceinline__ T eval_fast(const Get& get, const T* c) {
  static_assert(D::linear, "fast evaluation is for linear stencils");
  T acc;
  static_for<0, D::ntaps>([&](auto qc) {
    constexpr int Q = decltype(qc)::value;
    constexpr TapDef t = D::taps[Q];
    const T v = get.template at<t.d0, t.d1, t.d2>();
    if constexpr (Q == 0) acc = c[0] * v;
    else acc = fma_t(c[Q], v, acc);
  });
  return acc;
}

template <class D, typename T, int DM = -1, class Get>
__device__ __forceinline__ T eval_point(const Get& get, const T* c, T div, int has_div,
                                        double gscale) {
#pragma clang fp contract(off)
  if constexpr (D::linear) {
    T acc;
    static_for<0, D::ntaps>([&](auto qc) {
      constexpr int Q = decltype(qc)::value;
      constexpr TapDef t = D::taps[Q];
      const T v = get.template at<t.d0, t.d1, t.d2>();
      if constexpr (Q == 0) acc = c[0] * v;
      else if constexpr (t.sign > 0) acc = acc + c[Q] * v;
      else acc = acc - c[Q] * v;
    });
    if constexpr (DM < 0) {
      if (has_div) acc = div_by(acc, div, c[D::ntaps + 1], has_div);
    } else if constexpr (DM > 0) {
      acc = div_by(acc, div, c[D::ntaps + 1], DM);
    }
    return acc;
  } else {
    // gradient2d.c:18-22: C + 1.0f / sqrt(eps + (C-N)^2 + (C-S)^2 + (C-E)^2 + (C-W)^2)
    // with C's promotion: the float sum goes through the double sqrt.
    const T cc = get.template at<0, 0, 0>();
    const T dn = cc - get.template at<-1, 0, 0>();
    const T ds = cc - get.template at<1, 0, 0>();
    const T de = cc - get.template at<0, 0, 1>();
    const T dw = cc - get.template at<0, 0, -1>();
    T sum = c[0] + dn * dn;
    sum = sum + ds * ds;
    sum = sum + de * de;
    sum = sum + dw * dw;
    double q = 1.0 / __builtin_sqrt((double)sum);
    if (gscale != 1.0) q = gscale * q;
    return (T)((double)cc + q);
  }
}

// Shared-product form (radius-1 stars whose face taps all carry the same
// coefficient c_f): `get` returns the products P = c_f * u of the face
// neighbours (the in-plane ones formed once per cell of the centre plane and
// moved by DPP / the LDS exchange, the two off-plane ones formed here); the
// centre term is c_0 * u as in the C loop.  Every term is the same rounded
// product as in eval_point, summed in the same order: the result is bit for
// bit the same.
template <class D, typename T, int DM = -1, class Get>
__device__ __forceinline__ T eval_share(const Get& get, T ucen, const T* c, T div, int has_div) {
#pragma clang fp contract(off)
  static_assert(share_capable<D>(), "shared products need a radius-1 star");
  T acc;
  static_for<0, D::ntaps>([&](auto qc) {
    constexpr int Q = decltype(qc)::value;
    constexpr TapDef t = D::taps[Q];
    T v;
    if constexpr (Q == center_tap<D>()) v = c[Q] * ucen;
    else v = get.template at<t.d0, t.d1, t.d2>();
    if constexpr (Q == 0) acc = v;
    else if constexpr (t.sign > 0) acc = acc + v;
    else acc = acc - v;
  });
  if constexpr (DM < 0) {
    if (has_div) acc = div_by(acc, div, c[D::ntaps + 1], has_div);
  } else if constexpr (DM > 0) {
    acc = div_by(acc, div, c[D::ntaps + 1], DM);
  }
  return acc;
}

template <typename T, int CJ, int CK>
struct Plane { T v[CJ][CK]; };

// Stage input of output cell (JJ, KK) of a lane: taps on extended planes read
// the extension x (rows -XJ.., cells -XK..), the rest read the ring directly.
template <class D, typename T, int NXA, int CJ, int CK, int XJ, int XK, int JJ, int KK>
struct RegGet {
  const T (&x)[NXA][CJ + 2 * XJ][CK + 2 * XK];
  const Plane<T, CJ, CK> (&rin)[2 * D::rad + 1];
  template <int A0, int A1, int A2>
  __device__ __forceinline__ T at() const {
    constexpr int SL = ext_slot<D>(A0);
    if constexpr (SL >= 0) return x[SL][JJ + A1 + XJ][KK + A2 + XK];
    else return rin[D::rad + A0].v[JJ][KK];
  }
};

// Shared-product form: in-plane face products from the extension x of the
// centre plane's P, off-plane ones formed from the ring.
template <class D, typename T, int CJ, int CK, int XJ, int XK, int JJ, int KK>
struct ShareGet {
  const T (&x)[1][CJ + 2 * XJ][CK + 2 * XK];
  const Plane<T, CJ, CK> (&rin)[2 * D::rad + 1];
  T cf;
  template <int A0, int A1, int A2>
  __device__ __forceinline__ T at() const {
    if constexpr (A0 == 0) return x[0][JJ + A1 + XJ][KK + A2 + XK];
    else return cf * rin[D::rad + A0].v[JJ][KK];
  }
};

// Stage input straight from global memory (direct kernel).
template <typename T>
struct MemGet {
  const T* p;
  long long s0, s1;
  template <int A0, int A1, int A2>
  __device__ __forceinline__ T at() const { return p[A0 * s0 + A1 * s1 + A2]; }
};

template <int R>
constexpr int default_unroll() { return R <= 2 ? 2 * R + 1 : 1; }

// Workgroups are dispatched round-robin over gfx950's 8 XCDs, each with its
// own L2.  Workgroup b of nb takes logical tile (b % 8) * (nb / 8) + b / 8, so
// every XCD walks a contiguous range of tiles and neighbouring tiles, which
// re-read each other's halos, share an L2.  A bijection on [0, nb): the last
// nb % 8 workgroups keep their own index.
__device__ __forceinline__ int xcd_tile(int b, int nb) {
  constexpr int kXcds = 8;
  const int q = nb / kXcds;
  return b < q * kXcds ? (b % kXcds) * q + b / kXcds : b;
}

// Input planes in flight ahead of their use.  A depth that divides the unroll
// factor rotates by renaming; an unroll of 1 keeps a two-deep queue (one move
// per cell and plane).  At most ~40 VGPRs of queued cells.
template <int UNR, int CELLS32>
constexpr int prefetch_depth() {
  if (UNR == 1) return 2 * CELLS32 <= 40 ? 2 : 1;
  int best = 1;
  for (int d = 1; d <= UNR; d++)
    if (UNR % d == 0 && d * CELLS32 <= 40) best = d;
  return best;
}

// Rows a wave publishes per extended plane for the middle-dim exchange: its
// first and last XJ rows when it holds at least 2*XJ rows (neighbours need
// nothing else), otherwise all of them.
template <int CJ, int XJ>
constexpr int edge_rows() { return CJ >= 2 * XJ ? 2 * XJ : CJ; }
// storage index of own row RR among the published rows (-1: not published)
template <int CJ, int XJ>
constexpr int edge_index(int rr) {
  if (CJ < 2 * XJ) return rr;
  if (rr < XJ) return rr;
  if (rr >= CJ - XJ) return rr - (CJ - 2 * XJ);
  return -1;
}

// Stars (only the centre plane has in-plane taps, no corners): all stages'
// middle-dim exchanges of an iteration share one LDS round.
template <class D>
constexpr bool star_exchange() {
  return D::dim == 3 && needs_mid<D>() && num_ext_planes<D>() == 1 && ext_slot<D>(0) == 0 &&
         !needs_corners<D>();
}

// Radius-1 stars with bt >= 3 double-buffer that exchange: the edge rows a
// stage produces in one iteration go to the other buffer right away, and one
// barrier per iteration separates them from their readers.  (Shallower
// blocks are bandwidth-bound and keep the single buffer, which leaves room
// for more resident workgroups.)
template <class D>
constexpr bool star_exchange_db(int bt) {
  return star_exchange<D>() && D::rad == 1 && bt >= 3;
}

// LDS bytes of the middle-dim exchange: stars one buffer (two for radius 1)
// with a slot per stage, the others two buffers of one slot per extended plane.
template <int ID, typename T, int CK, int CJ, int NW>
constexpr size_t sweep_lds_bytes(int bt) {
  using D = Def<ID>;
  if constexpr (star_exchange<D>())
    return (star_exchange_db<D>(bt) ? 2 : 1) * (size_t)bt * NW * edge_rows<CJ, D::rad>() * 64 * CK *
           sizeof(T);
  else if constexpr (D::dim == 3 && needs_mid<D>())
    return 2ull * num_ext_planes<D>() * NW * edge_rows<CJ, D::rad>() * 64 * CK * sizeof(T);
  else
    return 0;
}

// tile >= 0 (2D only): process that wave tile instead of the one implied by
// blockIdx (used by the batched work-queue kernel).
// MIDS: the kernel also stores intermediate stages (SweepArgs::mid_mask);
// compiled only where used, since the extra store paths enlarge every stage.
// SHARE: the shared-product form (eval_share): each stage forms the products
// P = c_f * u of its centre plane once, and the in-plane neighbours (DPP and
// the middle-dim exchange, whose edge rows are published as P) move P
// instead of u.
template <int ID, typename T, int BT, int CK, int CJ, int NW, bool EXACT, int DM, int UNR,
          bool MIDS = false, bool SHARE = false>
__device__ __forceinline__ void sweep_body(const SweepArgs<T>& a, int tile = -1) {
  using D = Def<ID>;
  constexpr int R = D::rad;
  constexpr int P = 2 * R + 1;
  constexpr bool DIM3 = D::dim == 3;
  constexpr int NT = D::linear ? D::ntaps : 1;
  constexpr int NX = num_ext_planes<D>();
  constexpr bool MID = DIM3 && needs_mid<D>();
  constexpr bool FAST = needs_fast<D>();
  constexpr bool CORNER = needs_corners<D>();
  constexpr int XJ = MID ? R : 0;
  constexpr int XK = FAST ? R : 0;
  constexpr int H = BT * R;
  // Non-temporal grid loads where a launch is bandwidth-bound; non-temporal
  // stores too when the grid does not fit the last-level cache (otherwise the
  // next launch reads the output back from it).
  constexpr bool STREAM = BT <= 2;
  static_assert(!FAST || CK >= R, "a lane must hold at least R cells");
  static_assert(DIM3 || CJ == 1, "2D sweeps hold one row per lane");
  static_assert(!MID || NW * CJ > 2 * H, "tile too thin for BT");
  static_assert(!SHARE || (share_capable<D>() && EXACT), "shared products: exact radius-1 stars only");

  const int lane = threadIdx.x & 63;
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);

  int tk, tj = 0, seg;
  if constexpr (DIM3) {
    int b = xcd_tile(blockIdx.x, gridDim.x);
    tk = b % a.ntk;
    b /= a.ntk;
    tj = b % a.ntj;
    seg = b / a.ntj;
  } else {
    const int gw = tile >= 0 ? tile : blockIdx.x * NW + w;
    if (gw >= a.ntk * a.nseg) return;
    tk = gw % a.ntk;
    seg = gw / a.ntk;
  }

  const int kv0 = R + tk * a.SK;
  const int kv1 = min(kv0 + a.SK, a.D2 - R);
  const int kl = kv0 - H + lane * CK;
  int jw = 0, jv0 = 0, jv1 = 1;
  if constexpr (DIM3) {
    jv0 = R + tj * a.SJ;
    jv1 = min(jv0 + a.SJ, a.D1 - R);
    jw = jv0 - H + w * CJ;
  }
  const int i0 = R + seg * a.sl;
  const int sle = min(a.sl, a.D0 - R - i0);
  const int nit = sle + 2 * H;

  bool kin[CK], kst[CK], jin[CJ], jst[CJ];
  bool kst_all = true;
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    const int k = kl + kk;
    kin[kk] = k >= R && k < a.D2 - R;
    kst[kk] = k >= kv0 && k < kv1;
    kst_all = kst_all && kst[kk];
  }
#pragma unroll
  for (int jj = 0; jj < CJ; jj++) {
    const int j = jw + jj;
    jin[jj] = !DIM3 || (j >= R && j < a.D1 - R);
    jst[jj] = !DIM3 || (j >= jv0 && j < jv1);
  }
  // every cell of this wave interior along dims 1-2: no Dirichlet select
  bool cells_in = true;
#pragma unroll
  for (int kk = 0; kk < CK; kk++) cells_in = cells_in && kin[kk];
#pragma unroll
  for (int jj = 0; jj < CJ; jj++) cells_in = cells_in && jin[jj];
  const bool wave_in = __builtin_amdgcn_ballot_w64(!cells_in) == 0;

  T c[NT + 2];
#pragma unroll
  for (int q = 0; q < NT + 2; q++) c[q] = a.coef[q];
  const T div = c[NT];
  const int has_div = a.has_div;
  const double gscale = a.gscale;

  extern __shared__ __align__(16) unsigned char an5d_lds[];
  int xbuf = 0;

  Plane<T, CJ, CK> ring[BT][P];
  // shared-product form: the face coefficient all face taps carry
  const T cf = c[share_capable<D>() ? share_tap<D>() : 0];
  auto scaled = [&](const Plane<T, CJ, CK>& in) {
    Plane<T, CJ, CK> o;
#pragma unroll
    for (int jj = 0; jj < CJ; jj++)
#pragma unroll
      for (int kk = 0; kk < CK; kk++) o.v[jj][kk] = cf * in.v[jj][kk];
    return o;
  };

  // Every load reads real grid cells: plane, row and column indices are
  // clamped into the grid (the stand-in values only reach cells outside it,
  // which are never stored).  Loads are unconditional, so a queued plane never
  // merges with an older value, which would wait for the load at once.
  // Loads go through buffer descriptors: the plane's base (moved down by H
  // cells) in the descriptor, the row's byte offset in soffset and the cell's
  // (column k + H, clamped into the row) in voffset; all 32-bit (the host
  // keeps a plane below 2 GiB) and none computed in the loop.  Stores keep
  // the 64-bit address form, which lets a lane write its run with one
  // multi-word store (measured faster than word-wise buffer stores).
  constexpr unsigned SZ = sizeof(T);
  unsigned vclamp[CK];
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    int k = kl + kk;
    k = k < 0 ? 0 : (k >= a.D2 ? a.D2 - 1 : k);
    vclamp[kk] = (unsigned)(k + H) * SZ;
  }
  auto load_plane = [&](Plane<T, CJ, CK>& pl, int p) {
    p = p < 0 ? 0 : (p >= a.D0 ? a.D0 - 1 : p);
    const auto rs = make_rsrc(a.src + (long long)p * a.s0 - H);
#pragma unroll
    for (int jj = 0; jj < CJ; jj++) {
      int j = jw + jj;
      if (DIM3) j = j < 0 ? 0 : (j >= a.D1 ? a.D1 - 1 : j);
      const int so = DIM3 ? j * a.s1 * (int)SZ : 0;
#pragma unroll
      for (int kk = 0; kk < CK; kk++) pl.v[jj][kk] = bload<STREAM, T>(rs, vclamp[kk], so);
    }
  };

  auto store_plane = [&](const Plane<T, CJ, CK>& pl, int q, T* dstbuf) {
    T* base = dstbuf + (long long)q * a.s0;
#pragma unroll
    for (int jj = 0; jj < CJ; jj++) {
      if (!jst[jj]) continue;
      T* row = base + (DIM3 ? (long long)(jw + jj) * a.s1 : 0ll) + kl;
      if (kst_all) {
        if (STREAM && a.nt_store) gstore_run<true, T, CK>(row, pl.v[jj]);
        else gstore_run<false, T, CK>(row, pl.v[jj]);
      } else {
#pragma unroll
        for (int kk = 0; kk < CK; kk++)
          if (kst[kk]) row[kk] = pl.v[jj][kk];
      }
    }
  };

  // Stage S output plane from ring[S-1]; q = global index of that plane.
  constexpr bool STARX = star_exchange<D>();
  constexpr bool STARX_DB = star_exchange_db<D>(BT);
  constexpr int ERS = edge_rows<CJ, XJ>();
  constexpr size_t XBUF = (size_t)BT * NW * ERS * 64 * CK;  // elements per exchange buffer
  // single-buffered exchange: the halo rows of each stage's centre plane, read
  // once per iteration; double-buffered: read by each stage when it needs them
  T xh[STARX && !STARX_DB ? BT : 1][STARX && !STARX_DB ? 2 * XJ : 1][CK];
  int xrd = 0;  // double-buffered exchange: the buffer this iteration reads

  auto compute = [&](const Plane<T, CJ, CK>(&rin)[P], Plane<T, CJ, CK>& out, int q, auto sc) {
    constexpr int S = decltype(sc)::value;
    // extended copies of the planes that need in-plane neighbours (in the
    // shared-product form: of the centre plane's products P = c_f * u)
    T x[NX > 0 ? NX : 1][CJ + 2 * XJ][CK + 2 * XK];
    Plane<T, CJ, CK> pcen;
    if constexpr (SHARE) pcen = scaled(rin[R]);
    static_for<-R, R + 1>([&](auto dc) {
      constexpr int DD = decltype(dc)::value;
      constexpr int SL = ext_slot<D>(DD);
      if constexpr (SL >= 0) {
        const auto& src = SHARE ? pcen : rin[R + DD];
#pragma unroll
        for (int jj = 0; jj < CJ; jj++) {
#pragma unroll
          for (int kk = 0; kk < CK; kk++) x[SL][jj + XJ][kk + XK] = src.v[jj][kk];
          if constexpr (XK > 0) {
#pragma unroll
            for (int m = 1; m <= XK; m++) {
              x[SL][jj + XJ][XK - m] = from_left(src.v[jj][CK - m]);
              x[SL][jj + XJ][XK + CK - 1 + m] = from_right(src.v[jj][m - 1]);
            }
          }
        }
      }
    });
    if constexpr (STARX_DB) {
      // the neighbours' edge rows of this stage's centre plane, written during
      // the previous iteration into buffer xrd
      const T* lds = reinterpret_cast<const T*>(an5d_lds) + (size_t)xrd * XBUF;
      static_for<0, 2 * XJ>([&](auto hc) {
        constexpr int HI = decltype(hc)::value;
        constexpr int JR = HI < XJ ? HI - XJ : CJ + (HI - XJ);
        constexpr int WOFF = JR >= 0 ? JR / CJ : -((-JR + CJ - 1) / CJ);
        constexpr int RR = JR - WOFF * CJ;
        constexpr int E = edge_index<CJ, XJ>(RR);
        static_assert(E >= 0, "halo row must be published");
        int ws = w + WOFF;
        ws = ws < 0 ? 0 : (ws >= NW ? NW - 1 : ws);
        T run[CK];
        load_run<T, CK>(run, lds + ((((size_t)(S - 1) * NW + ws) * ERS + E) * 64 + lane) * CK);
#pragma unroll
        for (int kk = 0; kk < CK; kk++) x[0][JR + XJ][kk + XK] = run[kk];
      });
    } else if constexpr (STARX) {
#pragma unroll
      for (int h = 0; h < 2 * XJ; h++) {
        const int jr = h < XJ ? h : CJ + XJ + (h - XJ);  // x row index of the halo row
#pragma unroll
        for (int kk = 0; kk < CK; kk++) x[0][jr][kk + XK] = xh[(S - 1) * !STARX_DB][h * !STARX_DB][kk];
      }
    } else if constexpr (MID) {
      // publish own edge rows, then fetch the XJ rows above and below
      constexpr int ER = edge_rows<CJ, XJ>();
      T* lds = reinterpret_cast<T*>(an5d_lds) + (size_t)xbuf * NX * NW * ER * 64 * CK;
      static_for<0, NX>([&](auto sc) {
        constexpr int SL = decltype(sc)::value;
        static_for<0, CJ>([&](auto jc) {
          constexpr int JJ = decltype(jc)::value;
          constexpr int E = edge_index<CJ, XJ>(JJ);
          if constexpr (E >= 0) {
            T run[CK];
#pragma unroll
            for (int kk = 0; kk < CK; kk++) run[kk] = x[SL][JJ + XJ][kk + XK];
            store_run<T, CK>(lds + ((((size_t)SL * NW + w) * ER + E) * 64 + lane) * CK, run);
          }
        });
      });
      __syncthreads();
      static_for<0, NX>([&](auto sc) {
        constexpr int SL = decltype(sc)::value;
        static_for<0, 2 * XJ>([&](auto hc) {
          constexpr int HI = decltype(hc)::value;
          constexpr int JR = HI < XJ ? HI - XJ : CJ + (HI - XJ);  // row relative to own first row
          constexpr int WOFF = JR >= 0 ? JR / CJ : -((-JR + CJ - 1) / CJ);
          constexpr int RR = JR - WOFF * CJ;
          constexpr int E = edge_index<CJ, XJ>(RR);
          static_assert(E >= 0, "halo row must be published");
          int ws = w + WOFF;
          ws = ws < 0 ? 0 : (ws >= NW ? NW - 1 : ws);
          T run[CK];
          load_run<T, CK>(run, lds + ((((size_t)SL * NW + ws) * ER + E) * 64 + lane) * CK);
#pragma unroll
          for (int kk = 0; kk < CK; kk++) x[SL][JR + XJ][kk + XK] = run[kk];
          if constexpr (CORNER && XK > 0) {
#pragma unroll
            for (int m = 1; m <= XK; m++) {
              x[SL][JR + XJ][XK - m] = from_left(run[CK - m]);
              x[SL][JR + XJ][XK + CK - 1 + m] = from_right(run[m - 1]);
            }
          }
        });
      });
      xbuf ^= 1;
    }
    const bool pin = q >= R && q < a.D0 - R;
    T val[CJ][CK];
    // DM == 2: the tap sums of the whole plane first, then the divisions with
    // one wave-uniform guard for all of them
    constexpr bool BATCHDIV = EXACT && D::linear && DM == 2;
    static_for<0, CJ>([&](auto jc) {
      constexpr int JJ = decltype(jc)::value;
      static_for<0, CK>([&](auto kc) {
        constexpr int KK = decltype(kc)::value;
        if constexpr (SHARE) {
          ShareGet<D, T, CJ, CK, XJ, XK, JJ, KK> get{x, rin, cf};
          const T u0 = rin[R].v[JJ][KK];
          if constexpr (BATCHDIV) val[JJ][KK] = eval_share<D, T, 0>(get, u0, c, div, has_div);
          else val[JJ][KK] = eval_share<D, T, DM>(get, u0, c, div, has_div);
        } else {
          RegGet<D, T, (NX > 0 ? NX : 1), CJ, CK, XJ, XK, JJ, KK> get{x, rin};
          if constexpr (BATCHDIV) val[JJ][KK] = eval_point<D, T, 0>(get, c, div, has_div, gscale);
          else if constexpr (EXACT || !D::linear) val[JJ][KK] = eval_point<D, T, DM>(get, c, div, has_div, gscale);
          else val[JJ][KK] = eval_fast<D, T>(get, c);
        }
      });
    });
    if constexpr (BATCHDIV) {
      const T rc = c[NT + 1];
      T sum[CJ][CK];
      bool ok = true;
#pragma unroll
      for (int jj = 0; jj < CJ; jj++)
#pragma unroll
        for (int kk = 0; kk < CK; kk++) {
          sum[jj][kk] = val[jj][kk];
          val[jj][kk] = div_fast(sum[jj][kk], div, rc);
          ok = ok && zero_or_normal(val[jj][kk]);
        }
      if (__builtin_amdgcn_ballot_w64(!ok) != 0) {
#pragma unroll
        for (int jj = 0; jj < CJ; jj++)
#pragma unroll
          for (int kk = 0; kk < CK; kk++)
            if (!zero_or_normal(val[jj][kk])) val[jj][kk] = sum[jj][kk] / div;
      }
    }
    if (pin && wave_in) {
#pragma unroll
      for (int jj = 0; jj < CJ; jj++)
#pragma unroll
        for (int kk = 0; kk < CK; kk++) out.v[jj][kk] = val[jj][kk];
    } else {
#pragma unroll
      for (int jj = 0; jj < CJ; jj++)
#pragma unroll
        for (int kk = 0; kk < CK; kk++)
          out.v[jj][kk] = (pin && jin[jj] && kin[kk]) ? val[jj][kk] : rin[R].v[jj][kk];
    }
  };

  // Input planes are loaded PF iterations ahead of their use.  The iteration
  // count is padded to a multiple of UNR so the unrolled body has no exit
  // inside it: the ring rotation then costs no register moves.  The padding
  // iterations only produce planes past the segment, never stored.
  // 3D: a plane is CJ rows and the LDS exchange paces the loop; a shallow
  // queue keeps the registers for the rings.  The double-buffered radius-1
  // stars are the exception: every wave of the workgroup meets at one barrier
  // per plane, so with one plane in flight they would all wait on their loads
  // together; they queue the whole unrolled body ahead.
  constexpr int PF = prefetch_depth<UNR, CJ * CK * (int)(sizeof(T) / 4) *
                                             (DIM3 && !star_exchange_db<D>(BT) ? 3 : 1)>();
  // The loop body holds U iterations.  A loaded plane lives P + PF
  // iterations (queue, then input ring); when that exceeds U it crosses two
  // loop back-edges and the register allocator copies it (waiting for the
  // load) at the end of the body.  Radius-1 kernels of moderate body size
  // (BT * CJ * CK <= 36 cell-stages per plane) unroll 2*UNR iterations so
  // every value crosses at most one; longer bodies would exceed what the
  // compiler fully unrolls or cost occupancy, and the wider radii keep UNR.
  constexpr int U = R == 1 && UNR == P && P + PF > UNR && PF <= P && BT * CJ * CK <= 36 ? 2 * UNR : UNR;
  constexpr bool PF_ROT = U % PF == 0;  // queue slots rotate by renaming
  // radius-1 stars: edge rows of plane `pl` as stage slot `st`'s centre plane
  // of the next iteration, into exchange buffer `b`
  auto publish = [&](const Plane<T, CJ, CK>& pl, int st, int b) {
    T* lds = reinterpret_cast<T*>(an5d_lds) + (size_t)b * XBUF;
#pragma unroll
    for (int jj = 0; jj < CJ; jj++) {
      const int e = edge_index<CJ, XJ>(jj);
      if (e >= 0) store_run<T, CK>(lds + ((((size_t)st * NW + w) * ERS + e) * 64 + lane) * CK, pl.v[jj]);
    }
  };
  if constexpr (STARX_DB) {
    // iteration 0 reads buffer 0: the (warm-up) centre planes of the rings
#pragma unroll
    for (int st = 0; st < BT; st++) publish(SHARE ? scaled(ring[st][R + 1]) : ring[st][R + 1], st, 0);
  }
  const int nit_pad = (nit + U - 1) / U * U;
  Plane<T, CJ, CK> pre[PF];
#pragma unroll
  for (int f = 0; f < PF; f++) load_plane(pre[f], i0 - H + f);
  // the U iterations from n0
  auto block = [&](int n0) {
#pragma unroll
    for (int u = 0; u < U; u++) {
      const int n = n0 + u;
      const int p_in = i0 - H + n;
      const int slot = PF_ROT ? u % PF : 0;
      if constexpr (STARX_DB) {
        // the edge rows of iteration n were written during iteration n-1
        // into buffer n&1; one barrier, then each stage reads its rows
        // (buffer n&1 is written again only after the next barrier)
        __syncthreads();
        xrd = n & 1;
      } else if constexpr (STARX) {
        // Every stage's centre plane this iteration is slot R+1 of its input
        // ring before the rotation: publish their edge rows, one barrier,
        // read the neighbours' rows, a second barrier before the next write.
        T* lds = reinterpret_cast<T*>(an5d_lds);
        static_for<1, BT + 1>([&](auto sc) {
          constexpr int S = decltype(sc)::value;
          static_for<0, CJ>([&](auto jc) {
            constexpr int JJ = decltype(jc)::value;
            constexpr int E = edge_index<CJ, XJ>(JJ);
            if constexpr (E >= 0)
            {
              T run[CK];
#pragma unroll
              for (int kk = 0; kk < CK; kk++)
                run[kk] = SHARE ? cf * ring[S - 1][R + 1].v[JJ][kk] : ring[S - 1][R + 1].v[JJ][kk];
              store_run<T, CK>(lds + ((((size_t)(S - 1) * NW + w) * ERS + E) * 64 + lane) * CK, run);
            }
          });
        });
        __syncthreads();
        static_for<1, BT + 1>([&](auto sc) {
          constexpr int S = decltype(sc)::value;
          static_for<0, 2 * XJ>([&](auto hc) {
            constexpr int HI = decltype(hc)::value;
            constexpr int JR = HI < XJ ? HI - XJ : CJ + (HI - XJ);
            constexpr int WOFF = JR >= 0 ? JR / CJ : -((-JR + CJ - 1) / CJ);
            constexpr int RR = JR - WOFF * CJ;
            constexpr int E = edge_index<CJ, XJ>(RR);
            static_assert(E >= 0, "halo row must be published");
            int ws = w + WOFF;
            ws = ws < 0 ? 0 : (ws >= NW ? NW - 1 : ws);
            load_run<T, CK>(xh[S - 1][HI], lds + ((((size_t)(S - 1) * NW + ws) * ERS + E) * 64 + lane) * CK);
          });
        });
        __syncthreads();
      }
#pragma unroll
      for (int p = 0; p < P - 1; p++) ring[0][p] = ring[0][p + 1];
      ring[0][P - 1] = pre[slot];
      if constexpr (!PF_ROT) {
#pragma unroll
        for (int f = 0; f + 1 < PF; f++) pre[f] = pre[f + 1];
      }
      load_plane(pre[PF_ROT ? slot : PF - 1], p_in + PF);
      if constexpr (STARX_DB) publish(SHARE ? scaled(ring[0][P - 1]) : ring[0][P - 1], 0, (n + 1) & 1);
      static_for<1, BT + 1>([&](auto sc) {
        constexpr int S = decltype(sc)::value;
        Plane<T, CJ, CK> out;
        compute(ring[S - 1], out, p_in - S * R, sc);
        if constexpr (STARX_DB && S < BT) publish(SHARE ? scaled(out) : out, S, (n + 1) & 1);
        if constexpr (S < BT) {
          // owned planes of stage S: n in [H + S*R, H + S*R + sle)
          if constexpr (MIDS) {
            if ((a.mid_mask >> S) & 1u) {
              if (n >= H + S * R && n < H + S * R + sle) {
                const int m = __builtin_popcount(a.mid_mask & ((1u << S) - 1u));
                store_plane(out, p_in - S * R, a.mid + (long long)m * a.mid_stride);
              }
            }
          }
#pragma unroll
          for (int p = 0; p < P - 1; p++) ring[S][p] = ring[S][p + 1];
          ring[S][P - 1] = out;
        } else {
          if (n >= 2 * BT * R && n < nit) store_plane(out, p_in - BT * R, a.dst);
        }
      });
    }
  };
  for (int n0 = 0; n0 < nit_pad; n0 += U) block(n0);
}

// The division mode is fixed per launch: dispatch once, not per cell.
template <int ID, typename T, int BT, int CK, int CJ, int NW, bool EXACT, bool MIDS, int UNR, bool SHARE>
__device__ __forceinline__ void sweep_dispatch(const SweepArgs<T>& a) {
  if constexpr (!EXACT || !Def<ID>::linear) {
    sweep_body<ID, T, BT, CK, CJ, NW, EXACT, 0, UNR, MIDS, SHARE>(a);  // no division in these forms
  } else if constexpr (Def<ID>::divisor != 0) {
    if (a.has_div == 2) sweep_body<ID, T, BT, CK, CJ, NW, EXACT, 2, UNR, MIDS, SHARE>(a);
    else if (a.has_div == 1) sweep_body<ID, T, BT, CK, CJ, NW, EXACT, 1, UNR, MIDS, SHARE>(a);
    else sweep_body<ID, T, BT, CK, CJ, NW, EXACT, 0, UNR, MIDS, SHARE>(a);
  } else {
    if (a.has_div == 0) sweep_body<ID, T, BT, CK, CJ, NW, EXACT, 0, UNR, MIDS, SHARE>(a);
    else sweep_body<ID, T, BT, CK, CJ, NW, EXACT, -1, UNR, MIDS, SHARE>(a);
  }
}

template <int ID, typename T, int BT, int CK, int CJ, int NW, bool EXACT = true, bool MIDS = false,
          int UNR = default_unroll<Def<ID>::rad>(), bool SHARE = false>
__global__ __launch_bounds__(64 * NW) void sweep_kernel(SweepArgs<T> a) {
  sweep_dispatch<ID, T, BT, CK, CJ, NW, EXACT, MIDS, UNR, SHARE>(a);
}

// The same sweep with a register cap: at least WPE waves per SIMD, so at
// most 512 / WPE registers per lane (the compiler spills what does not fit).
// The cap trades spills for occupancy, as nvcc's --maxrregcount does for the
// reference's binaries (make.sh:29-31); the tuner measures both.
template <int ID, typename T, int BT, int CK, int CJ, int NW, bool EXACT, int WPE>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(WPE))) void sweep_kernel_cap(
    SweepArgs<T> a) {
  sweep_dispatch<ID, T, BT, CK, CJ, NW, EXACT, false, default_unroll<Def<ID>::rad>(), false>(a);
}

// One step without temporal blocking: every thread evaluates CK adjacent
// interior cells straight from global memory.  Used where the register
// pipeline does not fit (the radius-2..4 3D boxes) and as a baseline.
template <int ID, typename T, int CK>
__global__ __launch_bounds__(256) void direct_kernel(SweepArgs<T> a) {
  using D = Def<ID>;
  constexpr int R = D::rad;
  constexpr bool DIM3 = D::dim == 3;
  constexpr int NT = D::linear ? D::ntaps : 1;
  const int nk = a.D2 - 2 * R;
  const int nkt = (nk + CK - 1) / CK;
  const int nj = DIM3 ? a.D1 - 2 * R : 1;
  const long long gid = (long long)blockIdx.x * blockDim.x + threadIdx.x;
  const long long total = (long long)nkt * nj * (a.D0 - 2 * R);
  if (gid >= total) return;
  const int kt = (int)(gid % nkt);
  const long long rest = gid / nkt;
  const int j = DIM3 ? (int)(rest % nj) + R : 0;
  const int i = (int)(DIM3 ? rest / nj : rest) + R;
  const long long base = (long long)i * a.s0 + (DIM3 ? (long long)j * a.s1 : 0ll);
  T c[NT + 2];
#pragma unroll
  for (int q = 0; q < NT + 2; q++) c[q] = a.coef[q];
  const int has_div = a.has_div;
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    const int k = R + kt * CK + kk;
    if (k >= a.D2 - R) break;
    const T* p = a.src + base + k;
    MemGet<T> get{p, a.s0, a.s1};
    a.dst[base + k] = eval_point<D, T>(get, c, c[NT], has_div, a.gscale);
  }
}

}  // namespace an5d

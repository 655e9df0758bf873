// Temporally blocked adjoint sweep (2D), hand-written HIP for gfx950.
//
// Problem: u_{t+1} = S(u_t) on the interior (halo fixed), loss
// L = 1/2 sum_interior (u_T - target)^2.  Reverse mode over one time block
// [t0, t0+BT): given the forward checkpoint u_{t0} and lambda_{t0+BT}
// (= dL/du_{t0+BT} on the interior, zero on the halo), one launch
//   * recomputes u_{t0+1..t0+BT-1} from the checkpoint (forward stages in
//     working precision: folded coefficients and FMAs, gradient2d with the
//     hardware reciprocal square root in FP32; the states match the forward
//     pass to rounding, which only perturbs the Jacobians),
//   * runs BT adjoint stages lambda_t = J_t^T mu_{t+1}, mu = lambda masked to
//     the interior:  linear stencils: lambda_t(y) = sum_q s_q c_q/div *
//     mu_{t+1}(y - d_q) (the transposed taps);  gradient2d: the Jacobian of
//     C + scale/sqrt(eps + sum_m (C - N_m)^2) at u_t, transposed,
//   * accumulates the coefficient gradient sum_x mu_{t+1}(x) dF/dc(x, u_t)
//     for every t of the block over the cells this wave owns.
// All of it streams along dim 0 in registers like the forward sweep: forward
// stage m lags the checkpoint input by m*R planes; the adjoint pipeline lags
// it by DLT + a*R planes so that adjoint stage a (time t0+BT-a) finds the
// forward state u_{t0+BT-a} it needs (W planes either side) still in the
// stage's ring.  Fastest-dim neighbours come from DPP as in the forward.
// Tolerance-based (FP32 vs an FP64 oracle): adjoint arithmetic contracts to
// FMA; the checkpoints themselves come from the exact forward sweep.
#pragma once

#include "engine.hpp"
#include "sweep.hpp"

namespace an5d {

template <typename T>
struct AdjArgs {
  const T* __restrict__ U;     // checkpoint u_{t0}
  const T* __restrict__ Lin;   // lambda_{t0+BT}, zero on the halo
  T* __restrict__ Lout;        // lambda_{t0}: interior written
  const T* __restrict__ coef;  // forward coefficients, then the divisor (gradient2d: eps)
  double gscale;
  int has_div;
  double* __restrict__ gslot;  // per-wave coefficient-gradient sums: NG doubles per wave tile
  int D0, D2;
  long long s0;                // row pitch of U, Lin, Lout (elements)
  int ntk, nseg, sl, SK;
};

// Input rows in flight ahead of their use: the largest divisor of the ring
// length that fits ~28 VGPRs (the queue then rotates by renaming), else 2 with
// the ring length rounded up to even.
template <int LFMAX, int CK, int ESZ>
constexpr int adj_prefetch() {
  const int per = 2 * CK * ESZ / 4;  // VGPRs per queued plane (U and lambda rows)
  int best = 1;
  for (int d = 2; d <= LFMAX; d++)
    if (LFMAX % d == 0 && d * per <= 28) best = d;
  return best > 1 ? best : 2;
}

template <class D>
constexpr int adj_reach() { return D::linear ? D::rad : 2; }

__device__ __forceinline__ float rsqrt_t(float x) { return __builtin_amdgcn_rsqf(x); }
__device__ __forceinline__ double rsqrt_t(double x) { return 1.0 / __builtin_sqrt(x); }

// Cells a tile gives up on each side along the fastest dim (the garbage
// spread, rounded up to whole lanes so that every lane's CK cells are either
// all owned or all not).
template <int BT, int R, int W, int CK>
constexpr int adj_halo() { return ((2 * BT - 2) * R + W + CK - 1) / CK * CK; }

template <typename T, int CK, int XK>
__device__ __forceinline__ void extend_row(const T (&src)[CK], T (&x)[CK + 2 * XK]) {
#pragma unroll
  for (int kk = 0; kk < CK; kk++) x[kk + XK] = src[kk];
#pragma unroll
  for (int m = 1; m <= XK; m++) {
    x[XK - m] = from_left(src[CK - m]);
    x[XK + CK - 1 + m] = from_right(src[m - 1]);
  }
}

template <typename T>
__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// One wave tile gw (< ntk * nseg) of a reverse launch; the caller skips
// tiles past the end.
template <int ID, typename T, int BT, int CK>
__device__ __forceinline__ void adjoint_body(const AdjArgs<T>& a, int gw) {
  using D = Def<ID>;
  static_assert(D::dim == 2, "adjoint sweep is 2D");
  constexpr int R = D::rad;
  constexpr int P = 2 * R + 1;
  constexpr int W = adj_reach<D>();
  constexpr int DLT = W + (BT - 2) * R;  // adjoint lag behind the checkpoint input
  constexpr int HU = 2 * (BT - 1) * R + W;
  constexpr int HA = adj_halo<BT, R, W, CK>();
  constexpr int NT = D::linear ? D::ntaps : 1;
  constexpr int NG = D::linear ? D::ntaps : 2;  // gradient entries
  constexpr int XK = R;                          // fastest-dim extension of a plane
  constexpr int XU = D::linear ? R : 1;          // fastest-dim reach into forward states
  static_assert(CK >= R && CK >= XU, "a lane must hold at least R cells");
  static_assert(DLT >= 0, "bad lag");
  // forward ring of stage m holds LF(m) planes
  constexpr auto LF = [](int m) constexpr { return 2 * W + (2 * BT - 2 * m - 2) * R + 1; };
  constexpr int LFMAX = LF(0);

  const int lane = threadIdx.x & 63;
  const int tk = gw % a.ntk, seg = gw / a.ntk;
  const int kv0 = R + tk * a.SK;
  const int kl = kv0 - HA + lane * CK;
  const int i0 = R + seg * a.sl;
  const int sle = min(a.sl, a.D0 - R - i0);
  const int nit = sle + HU + DLT + BT * R;

  bool kin[CK];
  bool lane_in = true;
#pragma unroll
  for (int kk = 0; kk < CK; kk++) {
    const int k = kl + kk;
    kin[kk] = k >= R && k < a.D2 - R;
    lane_in = lane_in && kin[kk];
  }
  // every cell of the wave interior along the fastest dim: no per-cell masks
  const bool wave_in = __builtin_amdgcn_ballot_w64(!lane_in) == 0;
  // the lane's cells belong to this tile (lane-aligned: HA and SK are whole
  // lanes); in the last tile the owned lanes past the interior hold lambda = 0
  const bool lown = kl >= kv0 && kl < kv0 + a.SK;

  T c[NT + 2];
#pragma unroll
  for (int q = 0; q < NT + 2; q++) c[q] = a.coef[q];
  const T div = c[NT];
  const int has_div = a.has_div;
  const double gscale = a.gscale;
  // transposed taps: s_q c_q / div
  T ct[NT];
  if constexpr (D::linear) {
    static_for<0, NT>([&](auto qc) {
      constexpr int Q = decltype(qc)::value;
      constexpr TapDef tp = D::taps[Q];
      T v = tp.sign > 0 ? c[Q] : -c[Q];
      ct[Q] = has_div ? v / div : v;
    });
  }

  // All rings share one length L = LFMAX and the main loop is unrolled by L,
  // so rotating them is register renaming.  Forward stage m only reads its
  // last LF(m) slots and adjoint stages their last P: the other slots are
  // dead and cost nothing.  Rings start at zero and every load reads a real
  // grid cell, so every value the pipeline computes is finite and the
  // masked products (mu == 0 off the interior) stay exactly zero.
  constexpr int PF = adj_prefetch<LFMAX, CK, (int)sizeof(T)>();
  constexpr int L = (LFMAX + PF - 1) / PF * PF;
  T fr[BT][L][CK];  // forward rings (stage 0 = checkpoint input)
  T ar[BT][L][CK];  // adjoint rings (stage 0 = lambda input)
#pragma unroll
  for (int m = 0; m < BT; m++)
#pragma unroll
    for (int p = 0; p < L; p++)
#pragma unroll
      for (int kk = 0; kk < CK; kk++) {
        fr[m][p][kk] = T(0);
        ar[m][p][kk] = T(0);
      }
  double gsum[NG];
  T gpart[NG];
  // gradient2d: per adjoint stage, the partial lambda of the plane above the
  // one being finished and the S fluxes of its plane
  T nl_acc[BT][CK], nl_cs[BT][CK];
#pragma unroll
  for (int m = 0; m < BT; m++)
#pragma unroll
    for (int kk = 0; kk < CK; kk++) nl_acc[m][kk] = nl_cs[m][kk] = T(0);
#pragma unroll
  for (int g = 0; g < NG; g++) { gsum[g] = 0.0; gpart[g] = 0; }

  // Planes outside the grid read the workspace's zero padding rows and the
  // lanes beyond the row ends its zero padding cells, so every value the
  // pipeline computes is finite and the masked products (mu == 0 off the
  // interior) stay exactly zero.
  auto shift = [&](T (&ring)[L][CK], const T (&in)[CK]) {
#pragma unroll
    for (int p = 0; p < L - 1; p++)
#pragma unroll
      for (int kk = 0; kk < CK; kk++) ring[p][kk] = ring[p + 1][kk];
#pragma unroll
    for (int kk = 0; kk < CK; kk++) ring[L - 1][kk] = in[kk];
  };

  // inputs are fetched PF iterations ahead: iteration n consumes checkpoint
  // plane i0-HU+n and lambda plane i0-HU-DLT+n
  const long long s0 = a.s0;
  // uniform row pointers biased by -kAdjPad and a non-negative lane offset
  const unsigned koff = (unsigned)(kl + kAdjPad);
  const T* up = a.U + (long long)(i0 - HU) * s0 - kAdjPad;
  const T* lp = a.Lin + (long long)(i0 - HU - DLT) * s0 - kAdjPad;
  T preU[PF][CK], preL[PF][CK];
#pragma unroll
  for (int f = 0; f < PF; f++) {
    gload_run<false, T, CK>(preU[f], up + koff);
    gload_run<false, T, CK>(preL[f], lp + koff);
    up += s0;
    lp += s0;
  }
  const int nit_pad = (nit + L - 1) / L * L;
  for (int n0 = 0; n0 < nit_pad; n0 += L) {
#pragma unroll
    for (int u = 0; u < L; u++) {
      const int n = n0 + u;
      const int pU = i0 - HU + n;
      shift(fr[0], preU[u % PF]);
      shift(ar[0], preL[u % PF]);
      // unconditional (the padding rows absorb the overrun): a conditional
      // load would merge with the old slot value and wait for the load at once
      gload_run<false, T, CK>(preU[u % PF], up + koff);
      gload_run<false, T, CK>(preL[u % PF], lp + koff);
      up += s0;
      lp += s0;
      // ---- forward recompute (stages 1..BT-1) ------------------------------
      static_for<1, BT>([&](auto mc) {
        constexpr int M = decltype(mc)::value;
        const int q = pU - M * R;
        T xin[P][CK + 2 * XK];
        static_for<0, P>([&](auto pc) {
          constexpr int PP = decltype(pc)::value;
          extend_row<T, CK, XK>(fr[M - 1][L - P + PP], xin[PP]);
        });
        const bool pin = q >= R && q < a.D0 - R;
        T out[CK];
        static_for<0, CK>([&](auto kc) {
          constexpr int KK = decltype(kc)::value;
          T val;
          if constexpr (D::linear) {
            // folded coefficients s_q c_q / div, FMA chain
#pragma clang fp contract(fast)
            val = T(0);
            static_for<0, NT>([&](auto qc) {
              constexpr int Q = decltype(qc)::value;
              constexpr TapDef tp = D::taps[Q];
              val = fma_t(ct[Q], xin[R + tp.d0][KK + tp.d2 + XK], val);
            });
          } else {
            // gradient2d in working precision: the states feed only the
            // Jacobians, which the tolerance covers
            const T cc = xin[R][KK + XK];
            const T dn = cc - xin[R - 1][KK + XK], ds = cc - xin[R + 1][KK + XK];
            const T de = cc - xin[R][KK + XK + 1], dw = cc - xin[R][KK + XK - 1];
            val = cc + (T)gscale * rsqrt_t(c[0] + dn * dn + ds * ds + de * de + dw * dw);
          }
          out[KK] = val;
        });
        if (!(pin && wave_in)) {
#pragma unroll
          for (int kk = 0; kk < CK; kk++) out[kk] = (pin && kin[kk]) ? out[kk] : xin[R][kk + XK];
        }
        shift(fr[M], out);
      });
      // ---- adjoint stages ---------------------------------------------------
      static_for<1, BT + 1>([&](auto ac) {
        constexpr int A = decltype(ac)::value;
        constexpr int M = BT - A;          // forward stage paired with this adjoint stage
        constexpr int BM = L - LF(M);      // slot of plane qa - W in ring M
        const int qa = pU - DLT - A * R;
        const bool pin = qa >= R && qa < a.D0 - R;
        const bool own = qa >= i0 && qa < i0 + sle && n < nit;
        // mu planes qa-R..qa+R, row-extended
        T mx[P][CK + 2 * XK];
        static_for<0, P>([&](auto pc) {
          constexpr int PP = decltype(pc)::value;
          extend_row<T, CK, XK>(ar[A - 1][L - P + PP], mx[PP]);
        });
        T out[CK];
        if constexpr (D::linear) {
#pragma clang fp contract(fast)
          static_for<0, CK>([&](auto kc) {
            constexpr int KK = decltype(kc)::value;
            T acc = T(0);
            static_for<0, NT>([&](auto qc) {
              constexpr int Q = decltype(qc)::value;
              constexpr TapDef tp = D::taps[Q];
              acc = fma_t(ct[Q], mx[R - tp.d0][KK - tp.d2 + XK], acc);
            });
            out[KK] = acc;
          });
          {
            // sum_x mu(x) u_t(x + d_q) over owned cells of plane qa
            T ux[2 * R + 1][CK + 2 * XK];
            static_for<0, 2 * R + 1>([&](auto pc) {
              constexpr int PP = decltype(pc)::value;
              extend_row<T, CK, XK>(fr[M][BM + (W - R) + PP], ux[PP]);
            });
            // mu is 0 off the interior; lanes this tile does not own are
            // dropped when the partial sums are flushed
            if (own) {
              static_for<0, CK>([&](auto kc) {
                constexpr int KK = decltype(kc)::value;
                static_for<0, NT>([&](auto qc) {
                  constexpr int Q = decltype(qc)::value;
                  constexpr TapDef tp = D::taps[Q];
                  gpart[Q] = fma_t(mx[R][KK + XK], ux[R + tp.d0][KK + tp.d2 + XK], gpart[Q]);
                });
              });
            }
          }
        } else {
          // gradient2d: F(x) = C + scale * rs(x), rs = (eps + sum_m (C - N_m)^2)^-1/2
          //   dF/dC = 1 + f * sum_m (C - N_m), dF/dN_m = -f (C - N_m), f = -scale rs^3
          // Each x of plane qa+1 (the newest mu plane) sends mu(x) dF(x)/du(y)
          // to itself and its four neighbours y, from one Jacobian evaluation:
          //   lambda(qa) = [own terms of plane qa, its E/W fluxes and the S flux
          //                 of plane qa-1: kept from the previous iteration]
          //              + N flux of plane qa+1
#pragma clang fp contract(fast)
          T ux[3][CK + 2 * XU];  // u_t planes qa, qa+1, qa+2
          static_for<0, 3>([&](auto pc) {
            constexpr int PP = decltype(pc)::value;
            extend_row<T, CK, XU>(fr[M][BM + 2 + PP], ux[PP]);
          });
          const T eps = c[0];
          const T sc = (T)gscale;
          const bool own1 = qa + 1 >= i0 && qa + 1 < i0 + sle;
          T cN[CK], cS[CK], cE[CK], cW[CK], cC[CK];
          static_for<0, CK>([&](auto kc) {
            constexpr int KK = decltype(kc)::value;
            constexpr int K = KK + XU;
            const T cc = ux[1][K];
            const T dn = cc - ux[0][K], ds = cc - ux[2][K], de = cc - ux[1][K + 1], dw = cc - ux[1][K - 1];
            const T s2 = eps + dn * dn + ds * ds + de * de + dw * dw;
            const T rs = rsqrt_t(s2);
            const T f = -sc * rs * rs * rs;
            const T mu = mx[R + 1][KK + XK];
            const T mf = -mu * f;
            cN[KK] = mf * dn;  // to the plane above
            cS[KK] = mf * ds;  // to the plane below
            cE[KK] = mf * de;  // to x + (0,1)
            cW[KK] = mf * dw;  // to x - (0,1)
            cC[KK] = mu - mf * (dn + ds + de + dw);
            if (own1) {
              gpart[0] = gpart[0] + mu * (T(0.5) * f);  // d/d eps = -1/2 scale rs^3 = f/2
              gpart[1] = gpart[1] + mu * rs;           // d/d scale
            }
          });
          T cEx[CK + 2], cWx[CK + 2];
          extend_row<T, CK, 1>(cE, cEx);
          extend_row<T, CK, 1>(cW, cWx);
          static_for<0, CK>([&](auto kc) {
            constexpr int KK = decltype(kc)::value;
            out[KK] = nl_acc[A - 1][KK] + cN[KK];
            nl_acc[A - 1][KK] = cC[KK] + cEx[KK] + cWx[KK + 2] + nl_cs[A - 1][KK];
            nl_cs[A - 1][KK] = cS[KK];
          });
        }
        if (!(pin && wave_in)) {
#pragma unroll
          for (int kk = 0; kk < CK; kk++) out[kk] = (pin && kin[kk]) ? out[kk] : T(0);
        }
        if constexpr (A < BT) {
          shift(ar[A], out);
        } else {
          if (own && lown) gstore_run<false, T, CK>(a.Lout + (long long)qa * a.s0 + kl, out);
        }
      });
    }
    // flush the per-lane partial sums to FP64 once per unrolled body
#pragma unroll
    for (int g = 0; g < NG; g++) {
      gsum[g] += lown ? (double)gpart[g] : 0.0;
      gpart[g] = 0;
    }
  }
  static_for<0, NG>([&](auto gc) {
    constexpr int G = decltype(gc)::value;
    double v = wave_sum<T>(gsum[G] + (lown ? (double)gpart[G] : 0.0));
    if constexpr (D::linear) {
      // d u_{t+1} / d c_q = s_q u_t(x + d_q) / div
      v = (D::taps[G].sign > 0 ? v : -v) / (has_div ? (double)div : 1.0);
    }
    // each wave tile owns its slot (no atomics: thousands of waves ending
    // together on the same few addresses would serialise); a separate
    // reduction sums the slots in a fixed order
    if (lane == 0) a.gslot[(size_t)gw * NG + G] += v;
  });
}

template <int ID, typename T, int BT, int CK, int NW>
__global__ __launch_bounds__(64 * NW) void adjoint_kernel(AdjArgs<T> a) {
  const int w = __builtin_amdgcn_readfirstlane(threadIdx.x >> 6);
  const int gw = blockIdx.x * NW + w;
  if (gw >= a.ntk * a.nseg) return;
  adjoint_body<ID, T, BT, CK>(a, gw);
}

// lambda_T = (u_T - target) on the interior, 0 on the halo; the loss partial
// sum of every wave goes to its own slot (kSeedBlocks * 4 slots).
template <typename T>
__global__ __launch_bounds__(256) void adjoint_seed_kernel(const T* __restrict__ uT,
                                                           const T* __restrict__ target,
                                                           T* __restrict__ lam, double* __restrict__ lslot,
                                                           int D0, int D2, long long s0, long long st,
                                                           int R) {
  const long long total = (long long)D0 * D2;
  double part = 0.0;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    const int i = (int)(idx / D2), k = (int)(idx % D2);
    const long long x = (long long)i * s0 + k;
    const bool in = i >= R && i < D0 - R && k >= R && k < D2 - R;
    T d = in ? uT[x] - target[(long long)i * st + k] : T(0);
    lam[x] = d;
    part += 0.5 * (double)d * (double)d;
  }
  part = wave_sum<T>(part);
  if ((threadIdx.x & 63) == 0) lslot[blockIdx.x * 4 + (threadIdx.x >> 6)] = part;
}

// Copy the halo cells of `src` into `dst` (same layout), for the checkpoint
// slots: R full rows at the top and bottom, R cells at both ends of the
// other rows.  Only halo cells are visited.
template <typename T>
__global__ __launch_bounds__(256) void halo_copy_kernel(const T* __restrict__ src, T* __restrict__ dst,
                                                        int D0, int D2, long long s0, int R) {
  const long long rows_tb = 2LL * R * D2;               // top and bottom rows
  const long long sides = 2LL * R * (long long)(D0 - 2 * R);  // left/right cells of the other rows
  const long long total = rows_tb + sides;
  for (long long idx = (long long)blockIdx.x * blockDim.x + threadIdx.x; idx < total;
       idx += (long long)gridDim.x * blockDim.x) {
    int i, k;
    if (idx < rows_tb) {
      const int r = (int)(idx / D2);
      i = r < R ? r : D0 - 2 * R + r;
      k = (int)(idx % D2);
    } else {
      const long long j = idx - rows_tb;
      i = R + (int)(j / (2 * R));
      const int c = (int)(j % (2 * R));
      k = c < R ? c : D2 - 2 * R + c;
    }
    const long long x = (long long)i * s0 + k;
    dst[x] = src[x];
  }
}

}  // namespace an5d

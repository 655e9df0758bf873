// One step of a high-order 3D star per launch, streamed along dim 0 with the
// planes staged in LDS.
//
// The register pipeline of sweep_kernel keeps 2R+1 planes of every stage in
// VGPRs and recomputes R overlapping rows per tile edge; at radius 2-4 that
// either spills (FP64) or leaves one or two waves per SIMD.  Here a
// workgroup owns a tile of TJ = NW*CJ rows x TK = 64*CK cells of every plane
// of its dim-0 segment and nothing is recomputed:
//   - each plane's tile plus its R-cell halo ((TJ+2R) x (TK+2R) values) is
//     read from global memory once and kept in a ring of R+2 LDS slots: the
//     output plane i reads its in-plane taps from tile i and its dim-0 taps
//     i+1 .. i+R from the own cells of the next R tiles, while tile i+R+1 is
//     written into the slot that tile i-1 left; one barrier per plane;
//   - the dim-0 taps i-R .. i-1 come from a per-thread register queue of the
//     own cells of the last R planes (a star has no dim-0 tap off its own
//     column);
//   - the tiles of the next two planes are in flight in registers while a
//     plane is evaluated;
//   - lane l owns the cells k0 + l + 64*c (c < CK), so every global access
//     and every LDS read of a wave is 64 consecutive elements.
// Taps are evaluated in source order by eval_point / eval_fast, as in the
// other kernels, so exact mode is bitwise equal to the C loop.
#pragma once

#include "sweep.hpp"

namespace an5d {

// every tap off the centre plane lies on the cell's own dim-0 column
template <class D>
constexpr bool plane_taps_ok() {
  if (D::dim != 3 || !D::linear) return false;
  for (int q = 0; q < D::ntaps; q++)
    if (D::taps[q].d0 != 0 && (D::taps[q].d1 != 0 || D::taps[q].d2 != 0)) return false;
  return true;
}

template <typename T, int CK, int CJ, int NW, int R>
constexpr size_t plane_lds_bytes() {
  return (size_t)(R + 2) * (NW * CJ + 2 * R) * (64 * CK + 2 * R) * sizeof(T);
}

// Input of own cell (JJ, C) of output plane i: dim-0 taps i-R .. i-1 from
// the queue, the rest from the LDS tiles.  `t[m]` points at the cell's own
// position in the tile of plane i+m (m = 0..R).
template <class D, typename T, int CJ, int CK, int LK, int JJ, int C>
struct PlaneGet {
  const T (&qv)[D::rad][CJ][CK];
  const T* const (&t)[D::rad + 1];
  template <int A0, int A1, int A2>
  __device__ __forceinline__ T at() const {
    if constexpr (A0 < 0) return qv[D::rad + A0][JJ][C];
    else if constexpr (A0 > 0) return t[A0][0];
    else return t[0][A1 * LK + A2];
  }
};

template <int ID, typename T, int CK, int CJ, int NW, bool EXACT, int DM>
__device__ __forceinline__ void plane_body(const SweepArgs<T>& a) {
  using D = Def<ID>;
  static_assert(plane_taps_ok<D>(), "plane kernel: 3D linear stars only");
  constexpr int R = D::rad;
  constexpr int NT = D::ntaps;
  constexpr int TJ = NW * CJ, TK = 64 * CK;
  constexpr int LJ = TJ + 2 * R, LK = TK + 2 * R;
  constexpr int NTH = 64 * NW;
  constexpr int TILE = LJ * LK;
  constexpr int PER = (TILE + NTH - 1) / NTH;
  constexpr int NB = R + 2;  // LDS ring slots

  const int lane = threadIdx.x & 63;
  const int w = threadIdx.x >> 6;
  int b = xcd_tile(blockIdx.x, gridDim.x);
  const int tk = b % a.ntk;
  b /= a.ntk;
  const int tj = b % a.ntj;
  const int seg = b / a.ntj;
  const int k0 = R + tk * TK, j0 = R + tj * TJ;  // first output cell of the tile
  const int i0 = R + seg * a.sl;
  const int i1 = min(i0 + a.sl, a.D0 - R);  // exclusive
  if (i0 >= i1) return;                     // whole workgroup: uniform

  // tile element e = row * LK + col  <->  grid cell (j0 - R + row, k0 - R + col),
  // clamped into the grid (clamped cells feed no stored output)
  int toff[PER];
#pragma unroll
  for (int m = 0; m < PER; m++) {
    int e = (int)threadIdx.x + m * NTH;
    e = e < TILE ? e : TILE - 1;
    int gj = j0 - R + e / LK, gk = k0 - R + e % LK;
    gj = gj < a.D1 ? gj : a.D1 - 1;
    gk = gk < a.D2 ? gk : a.D2 - 1;
    toff[m] = gj * (int)a.s1 + gk;
  }
  int qoff[CJ][CK];
  bool own[CJ][CK];
#pragma unroll
  for (int jj = 0; jj < CJ; jj++)
#pragma unroll
    for (int c = 0; c < CK; c++) {
      const int j = j0 + w * CJ + jj, k = k0 + lane + 64 * c;
      own[jj][c] = j < a.D1 - R && k < a.D2 - R;
      qoff[jj][c] = (j < a.D1 ? j : a.D1 - 1) * (int)a.s1 + (k < a.D2 ? k : a.D2 - 1);
    }
  auto plane_ptr = [&](int p) {
    p = p < 0 ? 0 : (p >= a.D0 ? a.D0 - 1 : p);
    return a.src + (long long)p * a.s0;
  };

  T c[NT + 2];
#pragma unroll
  for (int q = 0; q < NT + 2; q++) c[q] = a.coef[q];
  const T div = c[NT];
  const int has_div = a.has_div;
  const double gscale = a.gscale;

  extern __shared__ __align__(16) unsigned char an5d_lds[];
  T* lds = reinterpret_cast<T*>(an5d_lds);
  // slot of plane p >= i0
  auto slot = [&](int p) { return lds + ((p - i0) % NB) * TILE; };

  auto load_tile = [&](T (&dst)[PER], int p) {
    const T* bp = plane_ptr(p);
#pragma unroll
    for (int m = 0; m < PER; m++) dst[m] = bp[toff[m]];
  };
  auto store_tile = [&](T* s, const T (&src)[PER]) {
#pragma unroll
    for (int m = 0; m < PER; m++) {
      const int e = (int)threadIdx.x + m * NTH;
      if (e < TILE) s[e] = src[m];
    }
  };

  // prologue: tiles of planes i0 .. i0+R into their slots, the queue with the
  // own cells of planes i0-R .. i0-1, then tiles i0+R+1 and i0+R+2 in flight
  T tvA[PER], tvB[PER];
#pragma unroll
  for (int m = 0; m <= R; m++) {
    load_tile(tvA, i0 + m);
    store_tile(slot(i0 + m), tvA);
  }
  T qv[R][CJ][CK];
#pragma unroll
  for (int s = 0; s < R; s++) {
    const T* bp = plane_ptr(i0 - R + s);
#pragma unroll
    for (int jj = 0; jj < CJ; jj++)
#pragma unroll
      for (int cc = 0; cc < CK; cc++) qv[s][jj][cc] = bp[qoff[jj][cc]];
  }
  load_tile(tvA, i0 + R + 1);
  load_tile(tvB, i0 + R + 2);
  __syncthreads();

  // A serves even steps, B odd ones, so a register buffer is refilled right
  // after it is consumed and never copied
  auto step = [&](int i, T (&tn)[PER]) {
    const T* t[R + 1];
#pragma unroll
    for (int m = 0; m <= R; m++) t[m] = slot(i + m);
    T* out = a.dst + (long long)i * a.s0;
    T centre[CJ][CK];
    static_for<0, CJ>([&](auto jc) {
      constexpr int JJ = decltype(jc)::value;
      static_for<0, CK>([&](auto kc) {
        constexpr int C = decltype(kc)::value;
        const int pos = (R + w * CJ + JJ) * LK + R + lane + 64 * C;
        const T* tc[R + 1];
#pragma unroll
        for (int m = 0; m <= R; m++) tc[m] = t[m] + pos;
        PlaneGet<D, T, CJ, CK, LK, JJ, C> get{qv, tc};
        T v;
        if constexpr (EXACT) v = eval_point<D, T, DM>(get, c, div, has_div, gscale);
        else v = eval_fast<D, T>(get, c);
        if (own[JJ][C]) out[qoff[JJ][C]] = v;
        centre[JJ][C] = tc[0][0];
      });
    });
    // tile i+R+1 into the slot of tile i-1, whose last readers finished
    // before the barrier that ended the previous step
    store_tile(slot(i + R + 1), tn);
#pragma unroll
    for (int s = 0; s < R - 1; s++)
#pragma unroll
      for (int jj = 0; jj < CJ; jj++)
#pragma unroll
        for (int cc = 0; cc < CK; cc++) qv[s][jj][cc] = qv[s + 1][jj][cc];
#pragma unroll
    for (int jj = 0; jj < CJ; jj++)
#pragma unroll
      for (int cc = 0; cc < CK; cc++) qv[R - 1][jj][cc] = centre[jj][cc];
    load_tile(tn, i + R + 3);
    __syncthreads();
  };
  for (int i = i0; i < i1; i += 2) {
    step(i, tvA);
    if (i + 1 < i1) step(i + 1, tvB);  // uniform over the workgroup
  }
}

// at least four waves per SIMD (<= 128 VGPRs): the plane loads need the
// occupancy to hide their latency
template <int ID, typename T, int CK, int CJ, int NW, bool EXACT>
__global__ __launch_bounds__(64 * NW) __attribute__((amdgpu_waves_per_eu(4))) void plane_kernel(
    SweepArgs<T> a) {
  if constexpr (!EXACT) {
    plane_body<ID, T, CK, CJ, NW, false, 0>(a);
  } else if constexpr (Def<ID>::divisor != 0) {
    if (a.has_div == 2) plane_body<ID, T, CK, CJ, NW, true, 2>(a);
    else if (a.has_div == 1) plane_body<ID, T, CK, CJ, NW, true, 1>(a);
    else plane_body<ID, T, CK, CJ, NW, true, 0>(a);
  } else {
    if (a.has_div == 0) plane_body<ID, T, CK, CJ, NW, true, 0>(a);
    else plane_body<ID, T, CK, CJ, NW, true, -1>(a);
  }
}

}  // namespace an5d

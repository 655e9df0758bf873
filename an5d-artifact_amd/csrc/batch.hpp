// Batched sweeps: many independent 2D grids advanced (and differentiated)
// together.
//
// Each instance's work is a chain of "levels"; one launch covers level l of
// every instance whose chain has one.  Its waves pull work units from a device
// counter, so small and large grids, cheap and expensive stencils, forward
// tiles of one instance and adjoint tiles of another share the chip in one
// launch.  Stream order between level launches is the only dependency (a unit
// of level l+1 reads what its instance's level-l units wrote), so no wave ever
// waits on another inside a launch.  The level sequence is captured once into
// a HIP graph.
//
// Unit types:
//   FWD    one wave tile of a forward launch (the sweep of sweep.hpp)
//   ADJ    one wave tile of a reverse block (the adjoint of adjoint.hpp)
//   PREP   rows of the initial state copied into checkpoint slot 0, with the
//          halo cells of those rows copied into every other slot (the sweeps
//          never write the halo); chunk 0 also clears the gradient slots
//   SEED   rows of lambda_T = u_T - target (0 on the halo) and a loss partial
//   GRAD   rows of lambda_0 copied into the caller's gradient grid
//   FINISH gradient and loss partials summed in a fixed order
#pragma once

#include "adjoint.hpp"

namespace an5d {

enum BatchUnitType { kUnitFwd = 0, kUnitAdj = 1, kUnitPrep = 2, kUnitSeed = 3, kUnitGrad = 4, kUnitFinish = 5 };

struct BatchRec {
  int type;     // BatchUnitType
  int kind;     // index into the batch kernel's stencil list
  int bt;       // steps of a FWD / ADJ record
  int has_div;  // division mode of the coefficient table
  int D0, D2;   // full extents (halo included)
  int R;        // radius
  int rows;     // PREP / SEED / GRAD: rows per unit
  long long s0;      // row pitch (elements) of the grids src / dst / aux (workspace or caller grid)
  long long s0_ref;  // PREP / SEED / GRAD: row pitch of the caller's grid
  int ntk, nseg, sl, SK;  // FWD / ADJ tiling
  const void* coef;
  const void* src;  // FWD input grid | ADJ checkpoint | PREP u0 | SEED u_T | GRAD lambda_0
  void* dst;        // FWD output grid | ADJ lambda out | PREP slot 0 | SEED lambda_T | GRAD caller's grad
  const void* aux;  // ADJ lambda in | SEED target
  double* slots;    // ADJ / PREP / FINISH gradient slots | SEED loss slots
  double* lslots;   // FINISH loss slots
  double* out;      // FINISH: ng gradient sums
  double* out_loss; // FINISH: loss
  long long nslots; // PREP / FINISH: gradient slots (ng doubles each)
  int ng, nl;       // FINISH: gradient entries, loss slots
  int nckpt;        // PREP: checkpoint slots
  long long ckpt_stride;  // PREP: elements between checkpoint slots
};

struct BatchUnit {
  int rec, tile;
};

// level kernel kinds (see batch_level_kernel)
constexpr int kLevelFwd = 0, kLevelFwdMisc = 1, kLevelAdj = 2;

// Stencils the batched kernels are built for (2D, FP32): forward CK = 4, the
// adjoint at its own FP32 shape (CK = 4 for radius <= 2).
constexpr int kBatchIds[] = {8 /* j2d5pt */, 9 /* j2d9pt */, 1 /* star2d2r */, 10 /* j2d9pt-gol */};
constexpr int kBatchKinds = 4;
constexpr int kBatchMaxBT = 5;
constexpr int kBatchCK = 4;
constexpr int kBatchNW = 4;
constexpr int kBatchAdjCK = 4;
// deepest adjoint block per kind: the level kernel holding every adjoint body
// stays within 256 VGPRs (two waves per SIMD); j2d5pt's bt = 5 body alone
// takes 246 and would push it past
constexpr int batch_adj_btmax(int kind) {
  return kind == 0 || kind == 3 ? 4 : 3;
}

#ifdef AN5D_BATCH_KERNELS  // kernels: defined in one translation unit only

template <int ID, int BT>
__device__ __forceinline__ void batch_fwd_tile(const BatchRec& r, int tile) {
  SweepArgs<float> a;
  a.src = static_cast<const float*>(r.src);
  a.dst = static_cast<float*>(r.dst);
  a.coef = static_cast<const float*>(r.coef);
  a.gscale = 1.0;
  a.has_div = r.has_div;
  a.D0 = r.D0;
  a.D1 = 1;
  a.D2 = r.D2;
  a.s0 = r.s0;
  a.s1 = 0;
  a.ntk = r.ntk;
  a.ntj = 1;
  a.nseg = r.nseg;
  a.sl = r.sl;
  a.SK = r.SK;
  a.SJ = 1;
  a.mid = nullptr;
  a.mid_stride = 0;
  a.mid_mask = 0;
  a.nt_store = 0;
  constexpr int UNR = default_unroll<Def<ID>::rad>();
  if constexpr (Def<ID>::divisor != 0) {
    if (r.has_div == 2) sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, 2, UNR>(a, tile);
    else sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, -1, UNR>(a, tile);
  } else {
    if (r.has_div == 0) sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, 0, UNR>(a, tile);
    else sweep_body<ID, float, BT, kBatchCK, 1, kBatchNW, true, -1, UNR>(a, tile);
  }
}

template <int K>
__device__ __forceinline__ void batch_fwd_dispatch(const BatchRec& r, int tile) {
  constexpr int ID = kBatchIds[K];
  switch (r.bt) {
    case 1: batch_fwd_tile<ID, 1>(r, tile); break;
    case 2: batch_fwd_tile<ID, 2>(r, tile); break;
    case 3: batch_fwd_tile<ID, 3>(r, tile); break;
    case 4: batch_fwd_tile<ID, 4>(r, tile); break;
    default: batch_fwd_tile<ID, 5>(r, tile); break;
  }
}

template <int ID, int BT>
__device__ __forceinline__ void batch_adj_tile(const BatchRec& r, int tile) {
  AdjArgs<float> a;
  a.U = static_cast<const float*>(r.src);
  a.Lin = static_cast<const float*>(r.aux);
  a.Lout = static_cast<float*>(r.dst);
  a.coef = static_cast<const float*>(r.coef);
  a.gscale = 1.0;
  a.has_div = r.has_div;
  a.gslot = r.slots;
  a.D0 = r.D0;
  a.D2 = r.D2;
  a.s0 = r.s0;
  a.ntk = r.ntk;
  a.nseg = r.nseg;
  a.sl = r.sl;
  a.SK = r.SK;
  adjoint_body<ID, float, BT, kBatchAdjCK>(a, tile);
}

template <int K>
__device__ __forceinline__ void batch_adj_dispatch(const BatchRec& r, int tile) {
  constexpr int ID = kBatchIds[K];
  constexpr int BTM = batch_adj_btmax(K);
  switch (r.bt) {
    case 1: batch_adj_tile<ID, 1>(r, tile); break;
    case 2: batch_adj_tile<ID, 2>(r, tile); break;
    case 3: batch_adj_tile<ID, 3>(r, tile); break;
    case 4: if constexpr (BTM >= 4) batch_adj_tile<ID, 4>(r, tile); break;
    default: if constexpr (BTM >= 5) batch_adj_tile<ID, 5>(r, tile); break;
  }
}

__device__ __forceinline__ double batch_wave_sum(double v) {
#pragma unroll
  for (int off = 32; off >= 1; off >>= 1) v += __shfl_xor(v, off, 64);
  return v;
}

// rows [c*rows, min(D0, (c+1)*rows)) of a row-chunk record
__device__ __forceinline__ void chunk_rows(const BatchRec& r, int c, int& i0, int& i1) {
  i0 = c * r.rows;
  i1 = min(r.D0, i0 + r.rows);
}

__device__ __forceinline__ void batch_prep(const BatchRec& r, int c) {
  const int lane = threadIdx.x & 63;
  const float* __restrict__ u0 = static_cast<const float*>(r.src);
  float* __restrict__ slot0 = static_cast<float*>(r.dst);
  int i0, i1;
  chunk_rows(r, c, i0, i1);
  for (int i = i0; i < i1; i++) {
    const float* src = u0 + (long long)i * r.s0_ref;
    float* d0 = slot0 + (long long)i * r.s0;
    for (int k = lane; k < r.D2; k += 64) d0[k] = src[k];
    const bool full = i < r.R || i >= r.D0 - r.R;
    for (int j = 1; j < r.nckpt; j++) {
      float* dj = d0 + (long long)j * r.ckpt_stride;
      if (full) {
        for (int k = lane; k < r.D2; k += 64) dj[k] = src[k];
      } else if (lane < 2 * r.R) {
        const int k = lane < r.R ? lane : r.D2 - 2 * r.R + lane;
        dj[k] = src[k];
      }
    }
  }
  if (c == 0)
    for (long long x = lane; x < r.nslots * r.ng; x += 64) r.slots[x] = 0.0;
}

__device__ __forceinline__ void batch_seed(const BatchRec& r, int c) {
  const int lane = threadIdx.x & 63;
  const float* __restrict__ uT = static_cast<const float*>(r.src);
  const float* __restrict__ tgt = static_cast<const float*>(r.aux);
  float* __restrict__ lam = static_cast<float*>(r.dst);
  int i0, i1;
  chunk_rows(r, c, i0, i1);
  double part = 0.0;
  for (int i = i0; i < i1; i++) {
    const bool irow = i >= r.R && i < r.D0 - r.R;
    for (int k = lane; k < r.D2; k += 64) {
      const long long x = (long long)i * r.s0 + k;
      const bool in = irow && k >= r.R && k < r.D2 - r.R;
      const float d = in ? uT[x] - tgt[(long long)i * r.s0_ref + k] : 0.0f;
      lam[x] = d;
      part += 0.5 * (double)d * (double)d;
    }
  }
  part = batch_wave_sum(part);
  if (lane == 0) r.slots[c] = part;
}

__device__ __forceinline__ void batch_grad(const BatchRec& r, int c) {
  const int lane = threadIdx.x & 63;
  const float* __restrict__ lam = static_cast<const float*>(r.src);
  float* __restrict__ g = static_cast<float*>(r.dst);
  int i0, i1;
  chunk_rows(r, c, i0, i1);
  for (int i = i0; i < i1; i++)
    for (int k = lane; k < r.D2; k += 64) g[(long long)i * r.s0_ref + k] = lam[(long long)i * r.s0 + k];
}

__device__ __forceinline__ void batch_finish(const BatchRec& r) {
  const int lane = threadIdx.x & 63;
  for (int b = 0; b < r.ng; b++) {
    double s = 0.0;
    for (long long x = lane; x < r.nslots; x += 64) s += r.slots[x * r.ng + b];
    s = batch_wave_sum(s);
    if (lane == 0) r.out[b] = s;
  }
  double s = 0.0;
  for (int x = lane; x < r.nl; x += 64) s += r.lslots[x];
  s = batch_wave_sum(s);
  if (lane == 0) *r.out_loss = s;
}

// Level kernels, each with its own register budget: KIND 0 runs FWD units
// (the forward batch), KIND 1 the FWD, PREP, SEED, GRAD and FINISH units of
// the forward + adjoint batch, KIND 2 its ADJ units.  Every wave leaves the
// loop once the counter passes nunits.
template <int KIND>
__global__ __launch_bounds__(64 * kBatchNW) void batch_level_kernel(const BatchRec* __restrict__ recs,
                                                                    const BatchUnit* __restrict__ units,
                                                                    int nunits, int* __restrict__ counter) {
  const int lane = threadIdx.x & 63;
  for (;;) {
    int u = 0;
    if (lane == 0) u = atomicAdd(counter, 1);
    u = __builtin_amdgcn_readlane(u, 0);
    if (u >= nunits) break;
    const BatchUnit wu = units[u];
    const BatchRec r = recs[wu.rec];
    if constexpr (KIND == kLevelAdj) {
      switch (r.kind) {
        case 0: batch_adj_dispatch<0>(r, wu.tile); break;
        case 1: batch_adj_dispatch<1>(r, wu.tile); break;
        case 2: batch_adj_dispatch<2>(r, wu.tile); break;
        default: batch_adj_dispatch<3>(r, wu.tile); break;
      }
    } else {
      if (r.type == kUnitFwd) {
        switch (r.kind) {
          case 0: batch_fwd_dispatch<0>(r, wu.tile); break;
          case 1: batch_fwd_dispatch<1>(r, wu.tile); break;
          case 2: batch_fwd_dispatch<2>(r, wu.tile); break;
          default: batch_fwd_dispatch<3>(r, wu.tile); break;
        }
      } else if constexpr (KIND == kLevelFwdMisc) {
        if (r.type == kUnitPrep) batch_prep(r, wu.tile);
        else if (r.type == kUnitSeed) batch_seed(r, wu.tile);
        else if (r.type == kUnitGrad) batch_grad(r, wu.tile);
        else batch_finish(r);
      }
    }
  }
}

__global__ void batch_reset_kernel(int* __restrict__ counters, int n) {
  for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) counters[i] = 0;
}

#endif  // AN5D_BATCH_KERNELS

}  // namespace an5d

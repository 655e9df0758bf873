// Adjoint entry points: checkpointed forward, seed, reverse block sweep.
// No counterpart in the reference (it computes no gradients).
#include <string.h>

#include <mutex>
#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "engine.hpp"
#include "internal.hpp"

namespace an5d {

#define AN5D_ADJ_DECL_2(id)                                  \
  void register_adjoint_##id##_0(std::vector<AdjVariant>&); \
  void register_adjoint_##id##_1(std::vector<AdjVariant>&);
#define AN5D_ADJ_DECL_3(id)
#define AN5D_ADJ_DECL(id, ident, dim, rad, fpp, n, div, lin) AN5D_ADJ_DECL_##dim(id)
AN5D_STENCILS(AN5D_ADJ_DECL)

std::vector<AdjVariant>& adj_registry() {
  static std::vector<AdjVariant> v = [] {
    std::vector<AdjVariant> out;
#define AN5D_ADJ_CALL_2(id)       \
  register_adjoint_##id##_0(out); \
  register_adjoint_##id##_1(out);
#define AN5D_ADJ_CALL_3(id)
#define AN5D_ADJ_CALL(id, ident, dim, rad, fpp, n, div, lin) AN5D_ADJ_CALL_##dim(id)
    AN5D_STENCILS(AN5D_ADJ_CALL)
    return out;
  }();
  return v;
}

static const AdjVariant* adj_variant(int stencil, int dtype) {
  for (const AdjVariant& v : adj_registry())
    if (v.stencil == stencil && v.dtype == dtype) return &v;
  return nullptr;
}

static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

// Workspace grids are padded: kAdjPad cells before column 0 and at least
// kAdjPad after the last one, kAdjRowPad rows above row 0 and below the last,
// so the adjoint kernel's loads need no bounds checks.  The padding is zero
// (the workspace is zeroed once) and is never written.
constexpr long long kPad = kAdjPad;
constexpr long long kRowPad = kAdjRowPad;

struct AdjPlan {
  const Variant* fwd = nullptr;
  const AdjVariant* adj = nullptr;
  int bt = 0, sl = 0;    // adjoint block size and segment length
  int btf = 0, fsl = 0;  // forward launches: steps per launch and sl
  std::vector<int> fsteps;  // forward launch schedule
  std::vector<int> blocks;  // adjoint blocks in time order (bt steps, the last may be shorter)
  Geom g;
  long long pitch = 0;  // elements per padded row
  size_t grid_bytes = 0, ws_bytes = 0;
  // after the NB+3 grids (checkpoints u at every block start and u_T, two
  // lambda buffers): per-wave gradient slots (nslots x ng doubles), then the
  // seed kernel's per-wave loss slots (kSeedBlocks * 4 doubles)
  long long nslots = 0;
  int ng = 0;
  size_t gslot_off = 0, lslot_off = 0;
};

static int adj_launch_plan(const AdjPlan& ap, int k, AdjLaunch& l, dim3& grid, dim3& block,
                           bool sizing = false);

// The forward pass runs at its own temporal block btf (the largest multiple
// of bt not above the tuned forward depth) and stores the state at every multiple of
// bt: a launch's output is the checkpoint where it ends, and the intermediate
// stages that reach a block boundary inside it store theirs.
static int plan_adjoint(const an5d_problem* p, const an5d_config* cfg, AdjPlan& ap) {
  std::string err;
  if (!make_geom(p, ap.g, err)) return fail(AN5D_E_INVALID, err);
  if (ap.g.dim != 2) return fail(AN5D_E_UNSUPPORTED, "the adjoint sweep supports 2D stencils");
  if (p->timesteps < 1) return fail(AN5D_E_INVALID, "adjoint needs timesteps >= 1");
  ap.adj = adj_variant(p->stencil, p->dtype);
  if (!ap.adj) return fail(AN5D_E_UNSUPPORTED, "no adjoint kernels for this stencil/dtype");
  auto vs = variants_of(p->stencil, p->dtype);
  if (cfg && cfg->variant >= (int)vs.size()) return fail(AN5D_E_INVALID, "variant index out of range");
  // forward kernel: the caller's variant, else the tuned/default forward choice
  an5d_config fc{0, cfg ? cfg->variant : -1, 0, AN5D_EXACT};
  int bt_pref = 0, sl_pref = 0;
  int rc = resolve(p, &fc, &ap.fwd, &bt_pref, &sl_pref);
  if (rc) return rc;
  int bt = cfg && cfg->bt > 0 ? cfg->bt : ap.adj->btmax;
  if (bt > ap.adj->btmax || bt > ap.fwd->btmax) {
    if (cfg && cfg->bt > 0) return fail(AN5D_E_UNSUPPORTED, "bt exceeds the adjoint or forward btmax");
    bt = ap.adj->btmax < ap.fwd->btmax ? ap.adj->btmax : ap.fwd->btmax;
  }
  if (ap.fwd->kind == kDirect) bt = 1;
  ap.bt = bt;
  // every launch ends on a block boundary (bt <= btmax, so m = 1 always fits);
  // the largest multiple of bt not above bt_pref: a deeper launch that also
  // stores intermediate states holds more registers than the tuned depth
  // (measured at 8192^2, T=256: j2d9pt bt=3 14.7 ms at m=1 against 15.3 at
  // m=2; j2d9pt-gol bt=4 11.0 against 12.0; at bt=3 / bt=2, where the
  // multiple stays within the tuned depth, m=2 won)
  int m = bt_pref / bt;
  if (m < 1) m = 1;
  while (m > 1 && (m * bt > ap.fwd->btmax || m * bt > kMaxMidBT)) m--;
  int btf = m * bt;
  if (ap.fwd->kind == kDirect) btf = 1;
  ap.btf = btf;
  ap.fsl = btf == bt_pref ? sl_pref : 0;
  // plain split (the states go to separate buffers, so the ping-pong parity
  // rule of the forward schedule does not apply)
  const int T = p->timesteps;
  ap.fsteps.assign(T / btf, btf);
  if (T % btf) ap.fsteps.push_back(T % btf);
  ap.blocks.clear();
  for (int t = 0; t < T; t += bt) ap.blocks.push_back(T - t < bt ? T - t : bt);
  ap.sl = cfg ? cfg->sl : 0;
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  ap.pitch = (ap.g.D2 + 2 * kPad + 63) / 64 * 64;
  ap.grid_bytes = ((size_t)(ap.g.D0 + 2 * kRowPad) * ap.pitch * esz + 255) / 256 * 256;
  // gradient slots: the most wave tiles any adjoint launch uses
  ap.ng = stencil_rt(p->stencil).linear ? stencil_rt(p->stencil).ntaps : 2;
  ap.nslots = 0;
  for (int k : ap.blocks) {
    AdjLaunch l{};
    dim3 grid, block;
    rc = adj_launch_plan(ap, k, l, grid, block, true);
    if (rc) return rc;
    const long long nt = (long long)l.ntk * l.nseg;
    if (nt > ap.nslots) ap.nslots = nt;
  }
  ap.gslot_off = ap.grid_bytes * (ap.blocks.size() + 1 + 2);
  ap.lslot_off = ap.gslot_off + ((size_t)ap.nslots * ap.ng * sizeof(double) + 255) / 256 * 256;
  ap.ws_bytes = ap.lslot_off + (size_t)kSeedBlocks * 4 * sizeof(double);
  return AN5D_OK;
}

// Zeroes n doubles (a fill kernel rather than a memset node, so the call
// sequence stays capturable in a graph).
__global__ __launch_bounds__(256) void zero_f64_kernel(double* __restrict__ p, long long n) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = 0.0;
}

// Block b < ng: gacc[b] = sum over the gradient slots of entry b; block ng:
// loss = sum of the loss slots.  Fixed summation order: the results are
// reproducible run to run.
__global__ __launch_bounds__(256) void adj_finish_kernel(const double* __restrict__ gslot,
                                                         long long nslots, int ng,
                                                         double* __restrict__ gacc,
                                                         const double* __restrict__ lslot, int nl,
                                                         double* __restrict__ loss) {
  __shared__ double red[256];
  const int b = blockIdx.x;
  double s = 0.0;
  if (b < ng) {
    for (long long i = threadIdx.x; i < nslots; i += 256) s += gslot[i * ng + b];
  } else {
    for (int i = threadIdx.x; i < nl; i += 256) s += lslot[i];
  }
  red[threadIdx.x] = s;
  __syncthreads();
  for (int o = 128; o > 0; o >>= 1) {
    if ((int)threadIdx.x < o) red[threadIdx.x] += red[threadIdx.x + o];
    __syncthreads();
  }
  if (threadIdx.x == 0) {
    if (b < ng) gacc[b] = red[0];
    else *loss = red[0];
  }
}

// Segments along dim 0: by default as many as fill the GPU's resident wave
// slots in one round (a second, partly filled round would leave most SIMDs
// idle), at least kMinSl planes each.  `sizing` gives the most segments any
// launch can use (the gradient-slot count) without touching the device.
constexpr long long kMinSl = 16;
static int adj_launch_plan(const AdjPlan& ap, int k, AdjLaunch& l, dim3& grid, dim3& block,
                           bool sizing) {
  const int R = ap.g.r, W = ap.adj->reach;
  const long long ck = ap.adj->ck;
  const long long HA = ((2LL * k - 2) * R + W + ck - 1) / ck * ck;  // whole lanes (adj_halo)
  const long long SK = 64LL * ap.adj->ck - 2 * HA;
  if (SK <= 0) return fail(AN5D_E_UNSUPPORTED, "adjoint block too deep for the tile");
  const long long nk = ap.g.D2 - 2 * R, n0 = ap.g.D0 - 2 * R;
  const long long ntk = cdiv(nk, SK);
  long long sl = ap.sl;
  if (sl <= 0) {
    long long slots = 0;
    if (!sizing) {
      const int occ = ap.adj->occupancy ? ap.adj->occupancy(k) : 0;
      slots = (long long)occ * device_cus() * ap.adj->nw;
    }
    if (sizing) {
      sl = kMinSl;
    } else if (slots > 0) {
      const long long nseg = slots / ntk > 1 ? slots / ntk : 1;
      sl = cdiv(n0, nseg);
    } else {
      sl = cdiv(n0, cdiv(4096, ntk));
    }
    if (sl < kMinSl) sl = kMinSl;
  }
  if (sl > n0) sl = n0;
  const long long nseg = cdiv(n0, sl);
  l.D0 = (int)ap.g.D0;
  l.D2 = (int)ap.g.D2;
  l.s0 = ap.pitch;
  // planes the kernel reads (see adjoint_kernel): checkpoint planes from
  // i0-HU on, lambda planes from i0-HU-DLT on (HU = 2(k-1)R+W,
  // DLT = W+(k-2)R), through i0+sle+DLT+k*R-1 plus the unrolled tail and the
  // prefetch queue (< 2 ring lengths, a ring length <= 2W+(2k-2)R+2); all of
  // them must stay inside the padding rows
  const long long HU = 2LL * (k - 1) * R + W, DLT = W + (long long)(k - 2) * R;
  const long long ring = 2LL * W + (2LL * k - 2) * R + 2;
  const long long lo = R - HU - (DLT > 0 ? DLT : 0);
  const long long hi = (ap.g.D0 - R) + (DLT > 0 ? DLT : 0) + (long long)k * R + 2 * ring;
  if (lo < -kRowPad || hi >= ap.g.D0 + kRowPad || 64LL * ap.adj->ck > kPad)
    return fail(AN5D_E_UNSUPPORTED, "adjoint block reaches beyond the workspace padding");
  l.ntk = (int)ntk;
  l.nseg = (int)nseg;
  l.sl = (int)sl;
  l.SK = (int)SK;
  block = dim3(64 * ap.adj->nw);
  grid = dim3((unsigned)cdiv(ntk * nseg, ap.adj->nw));
  return AN5D_OK;
}

}  // namespace an5d

using namespace an5d;

extern "C" int an5d_adjoint_workspace(const an5d_problem* p, const an5d_config* cfg, size_t* bytes) {
  AdjPlan ap;
  int rc = plan_adjoint(p, cfg, ap);
  if (rc) return rc;
  if (bytes) *bytes = ap.ws_bytes;
  // one padded grid per bt-step block: the footprint grows with T / bt.  A
  // workspace larger than the current device's memory can never be
  // allocated: say so here rather than as a failed allocation later (no
  // check without a device).
  size_t free_b = 0, total_b = 0;
  int dev = 0;
  if (hipGetDevice(&dev) == hipSuccess && hipMemGetInfo(&free_b, &total_b) == hipSuccess && total_b > 0 &&
      ap.ws_bytes > total_b)
    return fail(AN5D_E_ALLOC, "adjoint workspace of " + std::to_string(ap.ws_bytes >> 20) + " MiB (" +
                                  std::to_string(ap.blocks.size() + 3) +
                                  " padded grids: one per bt-step block) exceeds the device's " +
                                  std::to_string(total_b >> 20) + " MiB; use a larger bt or fewer steps");
  (void)hipGetLastError();
  return AN5D_OK;
}

namespace an5d {

// Enqueues the checkpointed forward pass, the seed, the reverse blocks and the
// gradient reduction of plan `ap` on `st`.  `ws` holds ap.ws_bytes of
// workspace whose grid padding is zero; `dcoef` is the device coefficient
// table built by build_coefs (has_div, gscale from the same call).
static int adjoint_enqueue(const AdjPlan& ap, const an5d_problem* p, const void* du0,
                           const void* dtarget, void* dgrad, double* dgcoef, double* dloss, void* ws,
                           const void* dcoef, int has_div, double gscale, hipStream_t st,
                           double* ms_fwd) {
  const Geom& g = ap.g;
  const int R = g.r;
  std::string err;
  int rc = AN5D_OK;
  const int L = (int)ap.blocks.size();  // checkpoint slot L holds u_T
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  char* base = static_cast<char*>(ws);
  auto ckpt = [&](int i) {
    return base + (size_t)i * ap.grid_bytes + (size_t)(kRowPad * ap.pitch + kPad) * esz;
  };
  char* lam[2] = {ckpt(L + 1), ckpt(L + 2)};
  double* gslot = reinterpret_cast<double*>(base + ap.gslot_off);
  double* lslot = reinterpret_cast<double*>(base + ap.lslot_off);
  const long long P = ap.pitch;
  hipError_t e;
  hipEvent_t ev0 = nullptr, ev1 = nullptr;
  if (ms_fwd) {
    (void)hipEventCreate(&ev0);
    (void)hipEventCreate(&ev1);
    (void)hipEventRecord(ev0, st);
  }
  // checkpoint 0 = u0; every other slot gets u0's halo (forward launches write interiors only)
  e = hipMemcpy2DAsync(ckpt(0), P * esz, du0, g.D2 * esz, g.D2 * esz, g.D0, hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy2DAsync(u0)");
  for (int i = 1; i <= L; i++) {
    e = ap.adj->halo(ckpt(0), ckpt(i), (int)g.D0, (int)g.D2, P, R, st);
    if (e != hipSuccess) return hip_fail(e, "halo copy");
  }
  // forward: launch j advances [t0, t0+kf); the state at every multiple of
  // bt inside it goes to that block's checkpoint
  const long long slot_elems = (long long)(ap.grid_bytes / esz);
  const int bt = ap.bt, T = p->timesteps;
  int t0 = 0;
  for (size_t j = 0; j < ap.fsteps.size(); j++) {
    const int kf = ap.fsteps[j], t1 = t0 + kf;
    LaunchPlan lp;
    if (!plan_launch(*ap.fwd, kf, g, ap.fsl, lp, err)) return fail(AN5D_E_UNSUPPORTED, err);
    if (t0 % bt != 0 || (t1 != T && t1 % bt != 0))
      return fail(AN5D_E_UNSUPPORTED, "forward launch does not end on a block boundary");
    lp.args.src = ckpt(t0 / bt);
    lp.args.dst = t1 == T ? ckpt(L) : ckpt(t1 / bt);
    lp.args.coef = dcoef;
    lp.args.gscale = gscale;
    lp.args.has_div = has_div;
    lp.args.s0 = P;
    lp.args.s1 = P;
    unsigned mask = 0;
    int bfirst = -1;
    for (int b = t0 / bt + 1; b * bt < t1; b++) {
      mask |= 1u << (b * bt - t0);
      if (bfirst < 0) bfirst = b;
    }
    if (mask) {
      lp.fn = ap.fwd->fn_mid[kf];
      if (!lp.fn) return fail(AN5D_E_UNSUPPORTED, "no forward kernel with intermediate outputs");
    }
    lp.args.mid = mask ? ckpt(bfirst) : nullptr;
    lp.args.mid_stride = slot_elems;
    lp.args.mid_mask = mask;
    lp.args.nt_store = 2 * ap.grid_bytes > kStreamBytes;
    e = lp.fn(lp.args, lp.grid, lp.block, lp.lds, st);
    if (e != hipSuccess) return hip_fail(e, "forward launch");
    t0 = t1;
  }
  if (ms_fwd) (void)hipEventRecord(ev1, st);
  // seed: lambda_T = u_T - target on the interior (0 on the halo), loss slots
  {
    const long long nz = (long long)(ap.ws_bytes - ap.gslot_off) / (long long)sizeof(double);
    hipLaunchKernelGGL(zero_f64_kernel, dim3(256), dim3(256), 0, st, gslot, nz);
    e = hipGetLastError();
    if (e != hipSuccess) return hip_fail(e, "slot clear");
  }
  e = ap.adj->seed(ckpt(L), dtarget, lam[0], lslot, (int)g.D0, (int)g.D2, P, g.D2, R, st);
  if (e != hipSuccess) return hip_fail(e, "adjoint seed");
  // reverse blocks
  int cur = 0;
  for (int i = L - 1; i >= 0; i--) {
    const int k = ap.blocks[i];
    AdjLaunch l{};
    dim3 grid, block;
    rc = adj_launch_plan(ap, k, l, grid, block);
    if (rc) return rc;
    l.U = ckpt(i);
    l.Lin = lam[cur];
    l.Lout = lam[cur ^ 1];
    l.coef = dcoef;
    l.gscale = gscale;
    l.has_div = has_div;
    l.gslot = gslot;
    if (!ap.adj->fn[k]) return fail(AN5D_E_UNSUPPORTED, "no adjoint kernel for this block size");
    e = ap.adj->fn[k](l, grid, block, st);
    if (e != hipSuccess) return hip_fail(e, "adjoint launch");
    cur ^= 1;
  }
  e = hipMemcpy2DAsync(dgrad, g.D2 * esz, lam[cur], P * esz, g.D2 * esz, g.D0,
                       hipMemcpyDeviceToDevice, st);
  if (e != hipSuccess) return hip_fail(e, "hipMemcpy2DAsync(grad)");
  hipLaunchKernelGGL(adj_finish_kernel, dim3(ap.ng + 1), dim3(256), 0, st, gslot, ap.nslots, ap.ng,
                     dgcoef, lslot, kSeedBlocks * 4, dloss);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "gradient reduction");
  if (ms_fwd) {
    float f = 0.f;
    (void)hipEventSynchronize(ev1);
    (void)hipEventElapsedTime(&f, ev0, ev1);
    *ms_fwd = f;
    (void)hipEventDestroy(ev0);
    (void)hipEventDestroy(ev1);
  }
  return AN5D_OK;
}

}  // namespace an5d

extern "C" int an5d_adjoint_device(const an5d_problem* p, const void* du0, const void* dtarget,
                                   void* dgrad, double* dgcoef, double* dloss, void* ws,
                                   size_t ws_bytes, const an5d_config* cfg, void* stream_v,
                                   double* ms_fwd) {
  AdjPlan ap;
  int rc = plan_adjoint(p, cfg, ap);
  if (rc) return rc;
  if (!du0 || !dtarget || !dgrad || !dgcoef || !dloss || !ws)
    return fail(AN5D_E_INVALID, "null device pointer");
  if (ws_bytes < ap.ws_bytes) return fail(AN5D_E_INVALID, "workspace too small");
  std::vector<unsigned char> cbytes;
  int has_div = 0;
  double gscale = 1.0;
  std::string err;
  if (!build_coefs(p, cbytes, has_div, gscale, err)) return fail(AN5D_E_INVALID, err);
  const void* dcoef = nullptr;
  rc = device_coefs(cbytes, &dcoef);
  if (rc) return rc;
  return adjoint_enqueue(ap, p, du0, dtarget, dgrad, dgcoef, dloss, ws, dcoef, has_div, gscale,
                         static_cast<hipStream_t>(stream_v), ms_fwd);
}

// ---- the adjoint without allocation ------------------------------------------
// The workspace and the coefficient table live in `scratch`, a device array of
// fixed size that every device holds from module load to process end.  In host
// code `scratch` names a host-side stand-in, never device memory: the device
// address comes from hipGetSymbolAddress, once per device, and only that
// address is passed to kernels.
namespace an5d {

constexpr size_t kNoAllocScratchBytes = 64ull << 20;
constexpr size_t kNoAllocCoefBytes = 1024;  // >= (81 taps + 2) * 8 bytes
static_assert(kNoAllocScratchBytes % 256 == 0, "scratch is carved in 256-byte units");

static __device__ __align__(256) unsigned char scratch[kNoAllocScratchBytes];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[kMaxDevices];

// Device address of this device's `scratch` (looked up on the first call on
// each device).
static int scratch_address(void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  if (dev < 0 || dev >= kMaxDevices) return fail(AN5D_E_UNSUPPORTED, "device index beyond the library's table");
  std::lock_guard<std::mutex> lk(g_scratch_mu);
  if (!g_scratch_ptr[dev]) {
    void* ptr = nullptr;
    e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
    if (e != hipSuccess) return hip_fail(e, "hipGetSymbolAddress(scratch)");
    if (!ptr) return fail(AN5D_E_HIP, "hipGetSymbolAddress(scratch) returned null");
    g_scratch_ptr[dev] = ptr;
  }
  *out = g_scratch_ptr[dev];
  return AN5D_OK;
}

// Zeroes n16 16-byte words (a fill kernel, not a memset node).
__global__ __launch_bounds__(256) void zero_words_kernel(uint4* __restrict__ p, long long n16) {
  for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n16;
       i += (long long)gridDim.x * blockDim.x)
    p[i] = make_uint4(0u, 0u, 0u, 0u);
}

// The coefficient table travels by value in the kernel arguments and is
// written to device memory by the stream itself.
struct CoefPack {
  unsigned char bytes[kNoAllocCoefBytes];
  int n;
};
__global__ __launch_bounds__(256) void stage_coefs_kernel(CoefPack pk, unsigned char* __restrict__ dst) {
  for (int i = threadIdx.x; i < pk.n; i += blockDim.x) dst[i] = pk.bytes[i];
}

}  // namespace an5d

extern "C" int an5d_adjoint_device_noalloc(const an5d_problem* p, const void* du0, const void* dtarget,
                                           void* dgrad, double* dgcoef, double* dloss,
                                           const an5d_config* cfg, void* stream_v) {
  AdjPlan ap;
  int rc = plan_adjoint(p, cfg, ap);
  if (rc) return rc;
  if (!du0 || !dtarget || !dgrad || !dgcoef || !dloss) return fail(AN5D_E_INVALID, "null device pointer");
  const size_t need = ap.ws_bytes + kNoAllocCoefBytes;
  if (need > kNoAllocScratchBytes)
    return fail(AN5D_E_INVALID, "problem needs " + std::to_string(need) + " bytes of adjoint workspace; " +
                                    "an5d_adjoint_device_noalloc holds at most " +
                                    std::to_string(kNoAllocScratchBytes) +
                                    " (use an5d_adjoint_device with a caller workspace)");
  std::vector<unsigned char> cbytes;
  int has_div = 0;
  double gscale = 1.0;
  std::string err;
  if (!build_coefs(p, cbytes, has_div, gscale, err)) return fail(AN5D_E_INVALID, err);
  if (cbytes.size() > kNoAllocCoefBytes) return fail(AN5D_E_INVALID, "coefficient table too large");
  void* ptr = nullptr;
  rc = scratch_address(&ptr);
  if (rc) return rc;
  hipStream_t st = static_cast<hipStream_t>(stream_v);
  unsigned char* base = static_cast<unsigned char*>(ptr);
  // the previous call (any geometry) left its data here: the grid padding and
  // the partial-sum slots must start at zero
  const long long n16 = (long long)(ap.ws_bytes / 16);
  long long blocks = (n16 + 255) / 256;
  if (blocks > 2048) blocks = 2048;
  if (blocks < 1) blocks = 1;
  hipLaunchKernelGGL(zero_words_kernel, dim3((unsigned)blocks), dim3(256), 0, st,
                     reinterpret_cast<uint4*>(base), n16);
  hipError_t e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "scratch clear");
  CoefPack pk{};
  memcpy(pk.bytes, cbytes.data(), cbytes.size());
  pk.n = (int)cbytes.size();
  unsigned char* dcoef = base + ap.ws_bytes;  // ws_bytes is a multiple of 256
  hipLaunchKernelGGL(stage_coefs_kernel, dim3(1), dim3(256), 0, st, pk, dcoef);
  e = hipGetLastError();
  if (e != hipSuccess) return hip_fail(e, "coefficient staging");
  return adjoint_enqueue(ap, p, du0, dtarget, dgrad, dgcoef, dloss, base, dcoef, has_div, gscale, st,
                         nullptr);
}

extern "C" int an5d_adjoint_host(const an5d_problem* p, const void* u0, const void* target,
                                 void* grad, double* gcoef, double* loss, const an5d_config* cfg,
                                 double* ms_fwd, double* ms_total) {
  AdjPlan ap;
  int rc = plan_adjoint(p, cfg, ap);
  if (rc) return rc;
  if (!u0 || !target || !grad || !gcoef || !loss) return fail(AN5D_E_INVALID, "null host pointer");
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  const size_t gb = (size_t)ap.g.area * esz;
  const int nc = ap.ng;
  void *du0 = nullptr, *dt = nullptr, *dg = nullptr, *ws = nullptr;
  double* dsc = nullptr;
  hipError_t e = hipMalloc(&du0, gb);
  if (e == hipSuccess) e = hipMalloc(&dt, gb);
  if (e == hipSuccess) e = hipMalloc(&dg, gb);
  if (e == hipSuccess) e = hipMalloc(&ws, ap.ws_bytes);
  if (e == hipSuccess) e = hipMemset(ws, 0, ap.ws_bytes);  // the workspace contract: zeroed once
  if (e == hipSuccess) e = hipMalloc(&dsc, sizeof(double) * (nc + 1));
  auto cleanup = [&]() {
    (void)hipFree(du0);
    (void)hipFree(dt);
    (void)hipFree(dg);
    (void)hipFree(ws);
    (void)hipFree(dsc);
  };
  if (e != hipSuccess) {
    cleanup();
    return fail(AN5D_E_ALLOC, std::string("hipMalloc(adjoint): ") + hipGetErrorString(e));
  }
  e = hipMemcpy(du0, u0, gb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipMemcpy(dt, target, gb, hipMemcpyHostToDevice);
  if (e == hipSuccess) e = hipDeviceSynchronize();
  if (e != hipSuccess) {
    cleanup();
    return hip_fail(e, "adjoint H2D");
  }
  const double t0 = sb_time_ms();
  double mf = 0.0;
  rc = an5d_adjoint_device(p, du0, dt, dg, dsc, dsc + nc, ws, ap.ws_bytes, cfg, nullptr, &mf);
  e = hipDeviceSynchronize();
  const double t1 = sb_time_ms();
  if (rc == AN5D_OK && e != hipSuccess) rc = hip_fail(e, "adjoint execution");
  if (rc == AN5D_OK) {
    e = hipMemcpy(grad, dg, gb, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(gcoef, dsc, sizeof(double) * nc, hipMemcpyDeviceToHost);
    if (e == hipSuccess) e = hipMemcpy(loss, dsc + nc, sizeof(double), hipMemcpyDeviceToHost);
    if (e != hipSuccess) rc = hip_fail(e, "adjoint D2H");
  }
  cleanup();
  if (ms_fwd) *ms_fwd = mf;
  if (ms_total) *ms_total = t1 - t0;
  return rc;
}

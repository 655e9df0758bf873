// C ABI of the engine: registry, launch schedule, device-buffer management
// and the forward entry points.  Host code only; kernels live in the
// per-stencil translation units built from fwd_inst.hip.
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

#include <map>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "engine.hpp"
#include "internal.hpp"
#include "stencil_defs.hpp"

namespace an5d {

// ---- errors -----------------------------------------------------------------
static thread_local std::string g_err;

int fail(int code, const std::string& msg) {
  g_err = msg;
  return code;
}

int hip_fail(hipError_t e, const char* what) {
  return fail(AN5D_E_HIP, std::string(what) + ": " + hipGetErrorString(e));
}

// ---- registry -----------------------------------------------------------------
#define AN5D_DECL_REG(id, ident, dim, rad, fpp, n, div, lin) \
  void register_variants_##id##_0(std::vector<Variant>&);     \
  void register_variants_##id##_1(std::vector<Variant>&);
AN5D_STENCILS(AN5D_DECL_REG)

std::vector<Variant>& registry() {
  static std::vector<Variant> v = [] {
    std::vector<Variant> out;
#define AN5D_CALL_REG(id, ident, dim, rad, fpp, n, div, lin) \
  register_variants_##id##_0(out);                            \
  register_variants_##id##_1(out);
    AN5D_STENCILS(AN5D_CALL_REG)
    return out;
  }();
  return v;
}

std::vector<const Variant*> variants_of(int stencil, int dtype) {
  std::vector<const Variant*> out;
  for (const Variant& v : registry())
    if (v.stencil == stencil && v.dtype == dtype) out.push_back(&v);
  return out;
}

const StencilRT& stencil_rt(int id) {
  static const std::vector<StencilRT> table = [] {
    std::vector<StencilRT> t;
#define AN5D_RT(id_, ident, dim_, rad_, fpp_, n_, div_, lin_)                       \
  {                                                                                  \
    StencilRT s;                                                                     \
    s.name = #ident;                                                                 \
    for (char& ch : s.name) if (ch == '_') ch = '-';                                 \
    s.dim = dim_; s.rad = rad_; s.fpp = fpp_; s.ntaps = n_; s.divisor = div_;        \
    s.linear = lin_;                                                                 \
    for (int q = 0; q < n_; q++) s.taps.push_back(Def<id_>::taps[q]);                \
    t.push_back(s);                                                                  \
  }
    AN5D_STENCILS(AN5D_RT)
    return t;
  }();
  return table[id];
}

// ---- geometry ---------------------------------------------------------------
bool make_geom(const an5d_problem* p, Geom& g, std::string& err) {
  if (!p) { err = "null problem"; return false; }
  if (p->stencil < 0 || p->stencil >= kNumStencils) { err = "unknown stencil id"; return false; }
  if (p->dtype != AN5D_FLOAT32 && p->dtype != AN5D_FLOAT64) { err = "bad dtype"; return false; }
  const StencilRT& s = stencil_rt(p->stencil);
  const int nd = s.dim;
  for (int d = 0; d < nd; d++)
    if (p->n[d] < 1) { err = "interior extents must be >= 1"; return false; }
  g.r = s.rad;
  g.dim = nd;
  g.D0 = p->n[0] + 2 * s.rad;
  if (nd == 2) { g.D1 = 1; g.D2 = p->n[1] + 2 * s.rad; }
  else { g.D1 = p->n[1] + 2 * s.rad; g.D2 = p->n[2] + 2 * s.rad; }
  if (g.D0 > (1LL << 30) || g.D1 > (1LL << 30) || g.D2 > (1LL << 30)) {
    err = "extent too large";
    return false;
  }
  g.area = g.D0 * g.D1 * g.D2;
  return true;
}

an5d_layout reference_layout(const Geom& g) {
  an5d_layout L;
  L.s1 = g.D2;
  L.s0 = g.D1 * g.D2;
  L.buf_stride = g.area;
  return L;
}

an5d_layout padded_layout(const Geom& g) {
  an5d_layout L;
  const long long pitch = (g.D2 + 63) / 64 * 64;  // 256-byte (FP32) rows
  L.s1 = pitch;
  L.s0 = g.D1 * pitch;
  L.buf_stride = g.D0 * L.s0;
  return L;
}

// ---- launch planning ----------------------------------------------------------
static long long cdiv(long long a, long long b) { return (a + b - 1) / b; }

int device_cus() {
  static std::mutex mu;
  static std::map<int, int> cache;
  int dev = 0, n = 0;
  if (hipGetDevice(&dev) != hipSuccess) return 0;
  std::lock_guard<std::mutex> lk(mu);
  auto it = cache.find(dev);
  if (it != cache.end()) return it->second;
  if (hipDeviceGetAttribute(&n, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) return 0;
  cache[dev] = n;
  return n;
}

// Default segment length along dim 0: as many segments as fill the GPU's
// resident workgroup slots in one round (a partly filled second round would
// leave most SIMDs idle); without a device to ask, about 8192 wave tiles.
int default_sl(const Variant& v, int bt, const Geom& g, long long ntk, long long ntj, int mode) {
  const long long n0 = g.D0 - 2 * g.r;
  const bool tiles3 = (v.kind == kSweep || v.kind == kPlane) && g.dim == 3;
  long long slots = 0;
  if ((v.kind == kSweep || v.kind == kPlane || v.kind == kRows) && bt >= 1 && bt <= kMaxBT && v.occ[bt]) {
    const int occ = v.occ[bt](mode);
    const int cus = occ > 0 ? device_cus() : 0;
    slots = (long long)occ * cus * (tiles3 ? 1 : v.nw);  // workgroups (3D) or waves (2D)
  }
  long long nseg;
  if (slots > 0) {
    const long long per_seg = tiles3 ? ntk * ntj : ntk;
    nseg = slots / per_seg > 1 ? slots / per_seg : 1;
  } else {
    long long tiles = ntk * ntj * (tiles3 ? v.nw : 1);
    nseg = cdiv(8192, tiles > 0 ? tiles : 1);
  }
  long long sl = cdiv(n0, nseg);
  long long minsl = 8LL * bt * g.r;
  if (sl < minsl) sl = minsl;
  if (sl > n0) sl = n0;
  if (sl < 1) sl = 1;
  return (int)sl;
}

bool plan_launch(const Variant& v, int bt, const Geom& g, int sl, LaunchPlan& lp,
                 std::string& err, int mode) {
  const int R = g.r;
  const long long n0 = g.D0 - 2 * R, nj = g.D1 - (g.dim == 3 ? 2 * R : 0), nk = g.D2 - 2 * R;
  LaunchArgs& a = lp.args;
  a.D0 = (int)g.D0; a.D1 = (int)g.D1; a.D2 = (int)g.D2;
  lp.fn = nullptr;
  if (bt >= 1 && bt <= v.btmax)
    lp.fn = mode == AN5D_FAST ? v.fn_fast[bt] : (mode == kLaunchShare ? v.fn_share[bt] : v.fn[bt]);
  if (!lp.fn) { err = "no kernel for bt=" + std::to_string(bt) + " in this variant"; return false; }
  lp.lds = (bt >= 1 && bt <= kMaxBT) ? v.lds[bt] : 0;
  if (v.kind == kDirect) {
    const long long total = cdiv(nk, v.ck) * nj * n0;
    lp.block = dim3(256);
    lp.grid = dim3((unsigned)cdiv(total, 256));
    a.ntk = a.ntj = a.nseg = a.sl = a.SK = a.SJ = 0;
    if (cdiv(total, 256) >= (1LL << 31)) { err = "grid too large"; return false; }
    return true;
  }
  if (v.kind == kPlane) {
    // tiles of nw*cj rows x 64*ck cells, halos staged, nothing recomputed
    if (g.dim != 3) { err = "the plane kernel is 3D only"; return false; }
    const long long TK = 64LL * v.ck, TJ = (long long)v.nw * v.cj;
    const long long ntk = cdiv(nk, TK), ntj = cdiv(nj, TJ);
    if (sl <= 0) sl = default_sl(v, bt, g, ntk, ntj, mode);
    const long long nseg = cdiv(n0, sl);
    a.ntk = (int)ntk; a.ntj = (int)ntj; a.nseg = (int)nseg; a.sl = sl;
    a.SK = (int)TK; a.SJ = (int)TJ;
    lp.block = dim3(64 * v.nw);
    const long long blocks = ntk * ntj * nseg;
    if (blocks >= (1LL << 31)) { err = "grid too large"; return false; }
    lp.grid = dim3((unsigned)blocks);
    return true;
  }
  // the sweep kernels address a plane (3D) or a row (2D) with 32-bit byte
  // offsets from a per-plane base
  const long long elem = v.dtype == kF64 ? 8 : 4;
  if ((g.dim == 3 ? g.D1 * g.D2 : g.D2) * elem >= (1LL << 31)) {
    err = "plane of 2 GiB or more: the register-streamed kernels address planes with 32-bit offsets";
    return false;
  }
  // the row partial-sum kernels read their tile edges themselves: no overlap
  const long long SK = v.kind == kRows ? 64LL * v.ck : 64LL * v.ck - 2LL * bt * R;
  if (SK <= 0) { err = "bt too deep for this variant's tile width"; return false; }
  const long long ntk = cdiv(nk, SK);
  long long SJ = 1, ntj = 1;
  if (g.dim == 3) {
    SJ = (long long)v.nw * v.cj - 2LL * bt * R;
    if (SJ <= 0) { err = "bt too deep for this variant's tile height"; return false; }
    ntj = cdiv(nj, SJ);
  }
  if (sl <= 0) sl = default_sl(v, bt, g, ntk, ntj, mode);
  const long long nseg = cdiv(n0, sl);
  a.ntk = (int)ntk; a.ntj = (int)ntj; a.nseg = (int)nseg; a.sl = sl;
  a.SK = (int)SK; a.SJ = (int)SJ;
  lp.block = dim3(64 * v.nw);
  long long blocks = (g.dim == 3) ? ntk * ntj * nseg : cdiv(ntk * nseg, v.nw);
  if (blocks >= (1LL << 31)) { err = "grid too large"; return false; }
  lp.grid = dim3((unsigned)blocks);
  return true;
}

// ---- coefficient buffers -------------------------------------------------------
// Host image of the kernel's coefficient array: ntaps values (T), the
// divisor and RN(1/divisor); for gradient2d {eps, 0, 0}.  has_div: 0 none,
// 1 IEEE division, 2 the corrected reciprocal product (only for the source
// divisors 118 and 159, for which it was checked against IEEE division).
bool build_coefs(const an5d_problem* p, std::vector<unsigned char>& bytes, int& has_div,
                 double& gscale, std::string& err, int mode) {
  const StencilRT& s = stencil_rt(p->stencil);
  const bool f64 = p->dtype == AN5D_FLOAT64;
  const int nt = s.linear ? s.ntaps : 1;
  std::vector<double> vals(nt + 2, 0.0);
  gscale = 1.0;
  has_div = 0;
  if (s.linear) {
    for (int q = 0; q < s.ntaps; q++)
      vals[q] = p->coefs ? p->coefs[q] : (double)s.taps[q].coef;
    double d = 0.0;
    if (p->divisor < 0) d = s.divisor;
    else if (p->divisor > 0) d = p->divisor;
    if (d != 0.0) {
      if (!(d == d) || d == INFINITY) { err = "divisor must be finite"; return false; }
      has_div = (d == 118.0 || d == 159.0) ? 2 : 1;
      vals[nt] = d;
    }
  } else {
    vals[0] = p->coefs ? p->coefs[0] : (double)0.0001f;
    gscale = p->coefs ? p->coefs[1] : 1.0;
    if (!(gscale == gscale)) { err = "scale is NaN"; return false; }
  }
  if (mode == AN5D_FAST && s.linear) {
    // fold signs and the divisor into the coefficients: sum_q fma(c'_q, v_q, .)
    const double d = has_div ? vals[nt] : 1.0;
    for (int q = 0; q < s.ntaps; q++) vals[q] = (s.taps[q].sign > 0 ? vals[q] : -vals[q]) / d;
    has_div = 0;
    vals[nt] = 0.0;
  }
  const size_t esz = f64 ? 8 : 4;
  bytes.resize((nt + 2) * esz);
  for (int q = 0; q <= nt; q++) {
    if (f64) {
      double v = vals[q];
      memcpy(&bytes[q * 8], &v, 8);
    } else {
      float v = (float)vals[q];
      memcpy(&bytes[q * 4], &v, 4);
    }
  }
  if (f64) {
    double rc = has_div ? 1.0 / vals[nt] : 0.0;
    memcpy(&bytes[(nt + 1) * 8], &rc, 8);
  } else {
    float rc = has_div ? 1.0f / (float)vals[nt] : 0.0f;
    memcpy(&bytes[(nt + 1) * 4], &rc, 4);
  }
  return true;
}

// The shared-product form applies when the stencil is a radius-1 star (a
// centre and its 2*dim face taps) and the coefficient table gives every face
// tap the same value in the working precision (bit for bit: each product
// c * u is then one rounded value, whichever sum it enters).
bool coefs_shareable(const an5d_problem* p, const std::vector<unsigned char>& bytes) {
  const StencilRT& s = stencil_rt(p->stencil);
  if (!s.linear || s.rad != 1 || s.ntaps != 2 * s.dim + 1) return false;
  int centre = -1;
  for (int q = 0; q < s.ntaps; q++) {
    const TapDef& t = s.taps[q];
    if (t.d1 != 0 && t.d2 != 0) return false;
    if (t.d0 == 0 && t.d1 == 0 && t.d2 == 0) centre = q;
  }
  if (centre < 0) return false;
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  const int ref = centre == 0 ? 1 : 0;
  for (int q = 0; q < s.ntaps; q++)
    if (q != centre && memcmp(&bytes[q * esz], &bytes[ref * esz], esz) != 0) return false;
  return true;
}

struct DeviceCache {
  std::mutex mu;
  std::map<std::pair<int, std::string>, void*> coefs;  // (device, bytes) -> device copy
  std::map<int, std::pair<void*, size_t>> grids;       // device -> staging grid buffer
  std::map<int, std::mutex> grid_users;                // device -> holder of its staging grid
};
static DeviceCache& cache() {
  static DeviceCache* c = new DeviceCache();
  return *c;
}

int device_coefs(const std::vector<unsigned char>& bytes, const void** out) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  DeviceCache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  std::string key(bytes.begin(), bytes.end());
  auto it = c.coefs.find({dev, key});
  if (it != c.coefs.end()) {
    *out = it->second;
    return AN5D_OK;
  }
  void* d = nullptr;
  e = hipMalloc(&d, bytes.size());
  if (e != hipSuccess) return hip_fail(e, "hipMalloc(coefficients)");
  e = hipMemcpy(d, bytes.data(), bytes.size(), hipMemcpyHostToDevice);
  if (e != hipSuccess) { (void)hipFree(d); return hip_fail(e, "hipMemcpy(coefficients)"); }
  c.coefs[{dev, key}] = d;
  *out = d;
  return AN5D_OK;
}

int device_grid(size_t bytes, void** out, std::unique_lock<std::mutex>& hold) {
  int dev = 0;
  hipError_t e = hipGetDevice(&dev);
  if (e != hipSuccess) return hip_fail(e, "hipGetDevice");
  DeviceCache& c = cache();
  std::mutex* user;
  {
    std::lock_guard<std::mutex> lk(c.mu);
    user = &c.grid_users[dev];  // map nodes stay put
  }
  // one host call at a time uses (and may regrow) a device's staging grid
  hold = std::unique_lock<std::mutex>(*user);
  std::lock_guard<std::mutex> lk(c.mu);
  auto& slot = c.grids[dev];
  if (slot.second < bytes) {
    if (slot.first) (void)hipFree(slot.first);
    slot = {nullptr, 0};
    void* d = nullptr;
    e = hipMalloc(&d, bytes);
    if (e != hipSuccess) return fail(AN5D_E_ALLOC, std::string("hipMalloc(grid): ") + hipGetErrorString(e));
    slot = {d, bytes};
  }
  *out = slot.first;
  return AN5D_OK;
}

// ---- variant / bt resolution ---------------------------------------------------
int default_bt(const StencilRT& s, int dtype) {
  if (s.dim == 2) return s.rad == 1 ? 4 : 2;
  return s.rad == 1 ? 2 : 1;
}

int resolve(const an5d_problem* p, const an5d_config* cfg, const Variant** vout, int* btout,
            int* slout) {
  auto vs = variants_of(p->stencil, p->dtype);
  if (vs.empty()) return fail(AN5D_E_UNSUPPORTED, "no kernels compiled for this stencil/dtype");
  const StencilRT& s = stencil_rt(p->stencil);
  int want_bt = cfg ? cfg->bt : 0;
  int want_v = cfg ? cfg->variant : -1;
  int sl = cfg ? cfg->sl : 0;
  const int mode = cfg ? cfg->mode : AN5D_EXACT;
  if (want_v >= (int)vs.size()) return fail(AN5D_E_INVALID, "variant index out of range");
  if (want_bt < 0 || want_bt > kMaxBT) return fail(AN5D_E_INVALID, "bt out of range");
  const Variant* v = nullptr;
  if (want_v >= 0) {
    v = vs[want_v];
  } else {
    TunedChoice tc;
    if (tuned_lookup(p, mode, &tc) && tc.variant < (int)vs.size() &&
        (want_bt == 0 || want_bt <= vs[tc.variant]->btmax)) {
      v = vs[tc.variant];
      if (want_bt == 0) want_bt = tc.bt;
      if (sl == 0) sl = tc.sl;
    } else {
      v = vs[0];  // first sweep shape, or the direct kernel when none exists
    }
  }
  int bt = want_bt ? want_bt : default_bt(s, p->dtype);
  if (bt > v->btmax) {
    if (want_bt && cfg && cfg->bt) return fail(AN5D_E_UNSUPPORTED, "bt=" + std::to_string(bt) + " exceeds the variant's btmax=" + std::to_string(v->btmax));
    bt = v->btmax;
  }
  *vout = v;
  *btout = bt;
  *slout = sl;
  return AN5D_OK;
}

}  // namespace an5d

using namespace an5d;

extern "C" int an5d_resolve_config(const an5d_problem* p, const an5d_config* cfg,
                                   an5d_config* out) {
  if (!p || !out) return fail(AN5D_E_INVALID, "null argument");
  Geom g;
  std::string err;
  if (!make_geom(p, g, err)) return fail(AN5D_E_INVALID, err);
  const Variant* v = nullptr;
  int bt = 0, sl = 0;
  const int mode = cfg ? cfg->mode : AN5D_EXACT;
  an5d_config rcfg = cfg ? *cfg : an5d_config{0, -1, 0, AN5D_EXACT};
  if (mode == AN5D_EXACT_GENERAL) rcfg.mode = AN5D_EXACT;  // tiled like exact mode
  int rc = resolve(p, &rcfg, &v, &bt, &sl);
  if (rc) return rc;
  LaunchPlan lp;
  if (!plan_launch(*v, bt, g, sl, lp, err, rcfg.mode)) return fail(AN5D_E_UNSUPPORTED, err);
  auto vs = variants_of(p->stencil, p->dtype);
  int vi = 0;
  while (vi < (int)vs.size() && vs[vi] != v) vi++;
  *out = an5d_config{bt, vi, v->kind == kDirect ? 0 : lp.args.sl, mode};
  return AN5D_OK;
}

// ---- schedule -------------------------------------------------------------------
extern "C" int an5d_schedule(int T, int bt, int* steps, int max_launches) {
  if (T < 0 || bt < 1) return fail(AN5D_E_INVALID, "schedule needs T >= 0 and bt >= 1");
  std::vector<int> out;
  const int full = T / bt, r = T % bt;
  const int nat = (T + bt - 1) / bt;  // launches of the plain split
  const bool mismatch = (T % 2) != (nat % 2);
  int main_count = full;
  if (mismatch && r < 2) main_count -= 1;  // __c0Padr (host.cu:72)
  for (int c = 0; c < main_count; c++) out.push_back(bt);
  if (mismatch) {
    if (r == 0) {
      out.push_back((bt + 1) / 2);
      out.push_back(bt / 2);
    } else if (r == 1) {
      const int b = (bt + 1) / 3;
      out.push_back(bt + 1 - 2 * b);
      out.push_back(b);
      out.push_back(b);
    } else {
      out.push_back((r + 1) / 2);
      out.push_back(r / 2);
    }
  } else if (r) {
    out.push_back(r);
  }
  if (steps) {
    if ((int)out.size() > max_launches) return fail(AN5D_E_INVALID, "steps[] too small");
    for (size_t c = 0; c < out.size(); c++) steps[c] = out[c];
  }
  return (int)out.size();
}

// ---- registry queries ---------------------------------------------------------
extern "C" int an5d_stencil_count(void) { return kNumStencils; }

extern "C" int an5d_stencil_id(const char* name) {
  if (!name) return fail(AN5D_E_INVALID, "null name");
  for (int i = 0; i < kNumStencils; i++)
    if (stencil_rt(i).name == name) return i;
  return fail(AN5D_E_INVALID, std::string("unknown stencil: ") + name);
}

extern "C" int an5d_stencil_get_info(int id, an5d_stencil_info* out) {
  if (id < 0 || id >= kNumStencils || !out) return fail(AN5D_E_INVALID, "bad stencil id");
  const StencilRT& s = stencil_rt(id);
  memset(out, 0, sizeof(*out));
  snprintf(out->name, sizeof(out->name), "%s", s.name.c_str());
  out->dim = s.dim; out->rad = s.rad; out->fpp = s.fpp; out->ntaps = s.ntaps;
  out->divisor = s.divisor; out->linear = s.linear;
  return AN5D_OK;
}

extern "C" int an5d_stencil_taps(int id, int* offsets, double* coefs, int max_taps) {
  if (id < 0 || id >= kNumStencils) return fail(AN5D_E_INVALID, "bad stencil id");
  const StencilRT& s = stencil_rt(id);
  if (max_taps < s.ntaps) return fail(AN5D_E_INVALID, "buffer too small");
  for (int q = 0; q < s.ntaps; q++) {
    if (offsets) {
      offsets[4 * q] = s.taps[q].sign;
      offsets[4 * q + 1] = s.taps[q].d0;
      offsets[4 * q + 2] = s.taps[q].d1;
      offsets[4 * q + 3] = s.taps[q].d2;
    }
    if (coefs) coefs[q] = s.taps[q].coef;
  }
  return s.ntaps;
}

extern "C" int an5d_variant_count(int stencil, int dtype) {
  if (stencil < 0 || stencil >= kNumStencils) return fail(AN5D_E_INVALID, "bad stencil id");
  return (int)variants_of(stencil, dtype).size();
}

extern "C" int an5d_variant_get_info(int stencil, int dtype, int variant, an5d_variant_info* out) {
  auto vs = variants_of(stencil, dtype);
  if (variant < 0 || variant >= (int)vs.size() || !out) return fail(AN5D_E_INVALID, "bad variant");
  const Variant* v = vs[variant];
  out->kind = v->kind; out->ck = v->ck; out->cj = v->cj; out->nw = v->nw;
  out->btmax = v->btmax; out->lds_bytes = (long long)v->lds[v->btmax];
  out->regcap = v->regcap;
  return AN5D_OK;
}

// ---- forward ------------------------------------------------------------------------
namespace an5d {

int run_schedule(const an5d_problem* p, void* dA, const an5d_layout* layout,
                 const an5d_config* cfg, int steps_only, int c0_first, hipStream_t st) {
  Geom g;
  std::string err;
  if (!make_geom(p, g, err)) return fail(AN5D_E_INVALID, err);
  if (!dA) return fail(AN5D_E_INVALID, "null device pointer");
  an5d_layout L = layout ? *layout : reference_layout(g);
  if (L.s1 < g.D2 || (g.dim == 3 && L.s0 < g.D1 * L.s1) || (g.dim == 2 && L.s0 < g.D2) ||
      L.buf_stride < g.D0 * L.s0)
    return fail(AN5D_E_INVALID, "layout strides smaller than the grid");
  const int mode = cfg ? cfg->mode : AN5D_EXACT;
  if (mode != AN5D_EXACT && mode != AN5D_FAST && mode != AN5D_EXACT_GENERAL)
    return fail(AN5D_E_INVALID, "bad mode");
  // AN5D_EXACT_GENERAL tiles like AN5D_EXACT (same table entries)
  an5d_config rcfg = cfg ? *cfg : an5d_config{0, -1, 0, AN5D_EXACT};
  if (mode == AN5D_EXACT_GENERAL) rcfg.mode = AN5D_EXACT;
  const Variant* v = nullptr;
  int bt = 0, sl = 0;
  int rc = resolve(p, &rcfg, &v, &bt, &sl);
  if (rc) return rc;
  std::vector<unsigned char> cbytes;
  int has_div = 0;
  double gscale = 1.0;
  if (!build_coefs(p, cbytes, has_div, gscale, err, rcfg.mode)) return fail(AN5D_E_INVALID, err);
  // launch mode per step count: the shared-product form where it applies
  // (bitwise the same) and its kernel does not spill
  const bool share = mode == AN5D_EXACT && v->kind == kSweep && coefs_shareable(p, cbytes);
  auto launch_mode = [&](int k) {
    if (share && v->fn_share[k] && v->share_scratch[k] && v->share_scratch[k]() == 0) return kLaunchShare;
    return rcfg.mode;
  };
  const void* dcoef = nullptr;
  rc = device_coefs(cbytes, &dcoef);
  if (rc) return rc;
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  std::vector<int> steps;
  if (steps_only > 0) {
    steps.push_back(steps_only);
  } else {
    int n = an5d_schedule(p->timesteps, v->kind == kDirect ? 1 : bt, nullptr, 0);
    if (n < 0) return n;
    steps.resize(n);
    an5d_schedule(p->timesteps, v->kind == kDirect ? 1 : bt, steps.data(), n);
  }
  int c0 = c0_first;
  // one plan per distinct step count (a schedule has at most three)
  LaunchPlan plans[kMaxBT + 1];
  bool planned[kMaxBT + 1] = {};
  for (int k : steps) {
    if (k < 1 || k > kMaxBT) return fail(AN5D_E_INVALID, "bad step count");
    if (!planned[k]) {
      if (!plan_launch(*v, k, g, sl, plans[k], err, launch_mode(k))) return fail(AN5D_E_UNSUPPORTED, err);
      planned[k] = true;
    }
    LaunchPlan lp = plans[k];
    char* base = static_cast<char*>(dA);
    lp.args.src = base + (size_t)(c0 % 2) * L.buf_stride * esz;
    lp.args.dst = base + (size_t)((c0 + 1) % 2) * L.buf_stride * esz;
    lp.args.coef = dcoef;
    lp.args.gscale = gscale;
    lp.args.has_div = has_div;
    lp.args.s0 = L.s0;
    lp.args.s1 = L.s1;
    lp.args.nt_store = (size_t)2 * L.buf_stride * esz > kStreamBytes;
    hipError_t e = lp.fn(lp.args, lp.grid, lp.block, lp.lds, st);
    if (e != hipSuccess) return hip_fail(e, "kernel launch");
    c0++;
  }
  return AN5D_OK;
}

double sb_time_ms() {
  struct timespec tp;
  clock_gettime(CLOCK_MONOTONIC_RAW, &tp);
  return (tp.tv_sec + tp.tv_nsec * 1.0e-9) * 1.0e3;
}

}  // namespace an5d

extern "C" int an5d_forward_device(const an5d_problem* p, void* dA, const an5d_layout* layout,
                                   const an5d_config* cfg, void* stream) {
  if (!p) return fail(AN5D_E_INVALID, "null problem");
  if (p->timesteps < 0) return fail(AN5D_E_INVALID, "timesteps must be >= 0");
  if (p->timesteps == 0) return AN5D_OK;
  return run_schedule(p, dA, layout, cfg, 0, 0, static_cast<hipStream_t>(stream));
}

extern "C" int an5d_launch(const an5d_problem* p, void* dA, const an5d_layout* layout,
                           const an5d_config* cfg, int steps, int c0, void* stream) {
  if (steps < 1 || c0 < 0) return fail(AN5D_E_INVALID, "steps must be >= 1 and c0 >= 0");
  return run_schedule(p, dA, layout, cfg, steps, c0, static_cast<hipStream_t>(stream));
}

extern "C" int an5d_forward_host(const an5d_problem* p, void* A, const an5d_config* cfg,
                                 double* ms) {
  Geom g;
  std::string err;
  if (!make_geom(p, g, err)) return fail(AN5D_E_INVALID, err);
  if (!A) return fail(AN5D_E_INVALID, "null host array");
  if (ms) *ms = 0.0;
  if (p->timesteps < 0) return fail(AN5D_E_INVALID, "timesteps must be >= 0");
  if (p->timesteps == 0) return AN5D_OK;
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  an5d_layout L = padded_layout(g);
  void* dA = nullptr;
  std::unique_lock<std::mutex> hold;  // the staging grid is ours until we return
  int rc = device_grid((size_t)2 * L.buf_stride * esz, &dA, hold);
  if (rc) return rc;
  const size_t rows = (size_t)g.D0 * g.D1;
  hipError_t e;
  for (int b = 0; b < 2; b++) {
    e = hipMemcpy2D(static_cast<char*>(dA) + (size_t)b * L.buf_stride * esz, L.s1 * esz,
                    static_cast<char*>(A) + (size_t)b * g.area * esz, g.D2 * esz, g.D2 * esz, rows,
                    hipMemcpyHostToDevice);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy2D H2D");
  }
  e = hipDeviceSynchronize();
  if (e != hipSuccess) return hip_fail(e, "hipDeviceSynchronize");
  const double t0 = sb_time_ms();
  rc = run_schedule(p, dA, &L, cfg, 0, 0, nullptr);
  if (rc) return rc;
  e = hipDeviceSynchronize();
  const double t1 = sb_time_ms();
  if (e != hipSuccess) return hip_fail(e, "kernel execution");
  for (int b = 0; b < 2; b++) {
    e = hipMemcpy2D(static_cast<char*>(A) + (size_t)b * g.area * esz, g.D2 * esz,
                    static_cast<char*>(dA) + (size_t)b * L.buf_stride * esz, L.s1 * esz, g.D2 * esz,
                    rows, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy2D D2H");
  }
  if (ms) *ms = t1 - t0;
  return AN5D_OK;
}

extern "C" int an5d_kernel_stencil(int stencil, int dtype, void* A1, int compsize, int timestep,
                                   int scop, double* ms) {
  if (stencil < 0 || stencil >= kNumStencils) return fail(AN5D_E_INVALID, "bad stencil id");
  const StencilRT& s = stencil_rt(stencil);
  an5d_problem p;
  memset(&p, 0, sizeof(p));
  p.stencil = stencil;
  p.dtype = dtype;
  for (int d = 0; d < s.dim; d++) p.n[d] = compsize;
  p.timesteps = timestep;
  p.coefs = nullptr;
  p.divisor = -1.0;
  if (ms) *ms = 0.0;
  // host.cu:17: nothing to do unless dimsize >= 2r+1 and timestep >= 1
  if ((long long)compsize + 2 * s.rad < 2 * s.rad + 1 || timestep < 1) return AN5D_OK;
  if (scop) {
    // tiling overrides for run drivers: AN5D_BT, AN5D_VARIANT, AN5D_SL
    an5d_config cfg{0, -1, 0, AN5D_EXACT};
    if (const char* e = getenv("AN5D_BT")) cfg.bt = atoi(e);
    if (const char* e = getenv("AN5D_VARIANT")) cfg.variant = atoi(e);
    if (const char* e = getenv("AN5D_SL")) cfg.sl = atoi(e);
    if (const char* e = getenv("AN5D_MODE")) cfg.mode = atoi(e);
    return an5d_forward_host(&p, A1, &cfg, ms);
  }
  return an5d_cpu_forward(&p, A1, 0, ms);
}

extern "C" int an5d_release(void) {
  DeviceCache& c = cache();
  std::lock_guard<std::mutex> lk(c.mu);
  for (auto& kv : c.coefs) (void)hipFree(kv.second);
  c.coefs.clear();
  for (auto& kv : c.grids)
    if (kv.second.first) (void)hipFree(kv.second.first);
  c.grids.clear();
  return AN5D_OK;
}

extern "C" const char* an5d_last_error(void) { return g_err.c_str(); }

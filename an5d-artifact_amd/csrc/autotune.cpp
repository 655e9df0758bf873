// Autotuner over kernel shape x bt x sl (and mode) on the device.
//
// Replaces the reference's offline selection (Excel model "Parameter Sweep"
// -> top-5 CSV -> compiled/nvcc_compile_top5.sh -> compiled/run.sh): the
// analytic model below ranks every candidate, the best-ranked ones are timed
// on the GPU with HIP events, and the winner is kept in a table that
// an5d_forward_* and an5d_kernel_stencil consult.  The model keeps the
// reference model's structure (Model_V100.xlsm "Float Results":
// time = max(compute, memory) per launch, times an occupancy factor;
// perf = T * cells / time) with gfx950 numbers and this engine's tiling:
//   compute = updates * bt * overlap * ops/update / VALU lane-op rate
//   memory  = cells * sizeof(T) * (read overlap + 1) / HBM bandwidth
// where overlap = (tile width / owned width) along each tiled dimension and
// (sl + 2*bt*R) / sl along the streamed one.
#include <dlfcn.h>
#include <math.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include <algorithm>
#include <cmath>
#include <mutex>
#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "internal.hpp"

namespace an5d {

namespace {

constexpr double kValuLaneOps = 256.0 * 4 * 16 * 2.4e9;  // non-packed VALU lane-ops/s
constexpr double kHbmBytes = 8.0e12 * 0.75;                // sustained share of 8 TB/s
constexpr double kLaunchSec = 4e-6;

struct Entry {
  int stencil, dtype, mode;
  long long n[3];
  int variant, bt, sl;
  double gstencil;
};

struct Table {
  std::mutex mu;
  std::vector<Entry> rows;
  bool loaded = false;
};

Table& table() {
  static Table* t = new Table();
  return *t;
}

std::string default_path() {
  if (const char* e = getenv("AN5D_TUNE_FILE")) return e;
  Dl_info info;
  if (dladdr(reinterpret_cast<void*>(&default_path), &info) && info.dli_fname) {
    std::string p = info.dli_fname;
    size_t slash = p.rfind('/');
    if (slash != std::string::npos) {
      p = p.substr(0, slash);  // .../lib
      slash = p.rfind('/');
      if (slash != std::string::npos) return p.substr(0, slash) + "/tuned.tsv";
    }
  }
  return "tuned.tsv";
}

int load_locked(Table& t, const std::string& path, bool quiet) {
  FILE* f = fopen(path.c_str(), "r");
  if (!f) return quiet ? AN5D_OK : fail(AN5D_E_INVALID, "cannot open " + path);
  char line[512];
  while (fgets(line, sizeof(line), f)) {
    if (line[0] == '#' || line[0] == '\n') continue;
    char name[64];
    Entry e{};
    int nf = sscanf(line, "%63s %d %d %lld %lld %lld %d %d %d %lf", name, &e.dtype, &e.mode, &e.n[0],
                    &e.n[1], &e.n[2], &e.variant, &e.bt, &e.sl, &e.gstencil);
    if (nf != 10) continue;
    e.stencil = -1;
    for (int i = 0; i < kNumStencils; i++)
      if (stencil_rt(i).name == name) e.stencil = i;
    if (e.stencil < 0) continue;
    t.rows.push_back(e);
  }
  fclose(f);
  return AN5D_OK;
}

void ensure_loaded(Table& t) {
  if (t.loaded) return;
  t.loaded = true;
  load_locked(t, default_path(), true);
}

void upsert(Table& t, const Entry& e) {
  for (Entry& r : t.rows)
    if (r.stencil == e.stencil && r.dtype == e.dtype && r.mode == e.mode && r.n[0] == e.n[0] &&
        r.n[1] == e.n[1] && r.n[2] == e.n[2]) {
      r = e;
      return;
    }
  t.rows.push_back(e);
}

double ops_per_update(const StencilRT& s, const Variant& v, int mode, bool has_div) {
  double ops;
  if (!s.linear) ops = 45.0;
  else if (mode == AN5D_FAST) ops = s.ntaps;
  else ops = 2.0 * s.ntaps - 1 + (has_div ? 6.0 : 0.0);
  if (v.kind == kSweep || v.kind == kRows) {
    // DPP halo per row and the middle-dim exchange, per cell
    ops += 2.0 * s.rad / v.ck;
    if (s.dim == 3) ops += 2.0 * s.rad * 2.0 / v.cj;
  } else if (v.kind == kPlane) {
    ops += 2.0 * s.rad;  // the dim-0 queue shifts by one plane per step
  } else {
    ops += 0.5 * s.ntaps;  // address arithmetic of the direct kernel
  }
  return ops;
}

}  // namespace

// Analytic estimate of GStencil/s for one configuration (the model prior).
double model_gstencil(const an5d_problem* p, const Variant& v, int bt, int sl, int mode) {
  Geom g;
  std::string err;
  if (!make_geom(p, g, err)) return 0.0;
  LaunchPlan lp;
  if (!plan_launch(v, bt, g, sl, lp, err, mode)) return 0.0;
  const StencilRT& s = stencil_rt(p->stencil);
  const double esz = p->dtype == AN5D_FLOAT64 ? 8.0 : 4.0;
  const double n0 = g.D0 - 2.0 * g.r, nj = g.dim == 3 ? g.D1 - 2.0 * g.r : 1.0, nk = g.D2 - 2.0 * g.r;
  const double cells = n0 * nj * nk;
  double ovk = 1, ovj = 1, ovi = 1, waves;
  if (v.kind == kSweep || v.kind == kRows) {
    ovk = lp.args.ntk * 64.0 * v.ck / nk;
    if (g.dim == 3) ovj = lp.args.ntj * (double)v.nw * v.cj / nj;
    ovi = (lp.args.nseg * (double)lp.args.sl + 2.0 * bt * g.r * lp.args.nseg) / n0;
    waves = (double)lp.args.ntk * lp.args.ntj * lp.args.nseg * (g.dim == 3 ? v.nw : 1);
  } else if (v.kind == kPlane) {
    ovk = lp.args.ntk * 64.0 * v.ck / nk;
    ovj = lp.args.ntj * (double)v.nw * v.cj / nj;
    waves = (double)lp.args.ntk * lp.args.ntj * lp.args.nseg * v.nw;
  } else {
    waves = cells / v.ck / 64.0;
  }
  const bool has_div = s.divisor != 0 || p->divisor > 0;
  const double ops = cells * bt * ovk * ovj * ovi * ops_per_update(s, v, mode, has_div);
  const double tc = ops / kValuLaneOps;
  const double tm = cells * esz * (ovk * ovj * ovi + 1.0) / kHbmBytes;
  // fewer than ~4 waves per SIMD leave latency exposed
  const double fill = std::min(1.0, waves / (4.0 * 1024.0));
  const double t = std::max(tc, tm) / std::max(fill, 0.05) + kLaunchSec;
  return cells * bt / t / 1e9;
}

bool tuned_lookup(const an5d_problem* p, int mode, TunedChoice* out) {
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.mu);
  ensure_loaded(t);
  const Entry* best = nullptr;
  double best_d = 1e300;
  const double lc = log((double)p->n[0] * std::max(1LL, p->n[1]) * std::max(1LL, p->n[2]));
  for (const Entry& r : t.rows) {
    if (r.stencil != p->stencil || r.dtype != p->dtype || r.mode != mode) continue;
    const double d = fabs(log((double)r.n[0] * std::max(1LL, r.n[1]) * std::max(1LL, r.n[2])) - lc);
    if (d < best_d) {
      best_d = d;
      best = &r;
    }
  }
  if (!best || best_d > log(8.0)) return false;  // only reuse within 8x of the size
  out->variant = best->variant;
  out->bt = best->bt;
  out->sl = (best->n[0] == p->n[0] && best->n[1] == p->n[1] && best->n[2] == p->n[2]) ? best->sl : 0;
  return true;
}

}  // namespace an5d

using namespace an5d;

extern "C" int an5d_model_estimate(const an5d_problem* p, const an5d_config* cfg, double* gstencil) {
  if (!p || !cfg || !gstencil) return fail(AN5D_E_INVALID, "null argument");
  auto vs = variants_of(p->stencil, p->dtype);
  if (cfg->variant < 0 || cfg->variant >= (int)vs.size()) return fail(AN5D_E_INVALID, "bad variant");
  *gstencil = model_gstencil(p, *vs[cfg->variant], cfg->bt, cfg->sl, cfg->mode);
  return AN5D_OK;
}

extern "C" int an5d_autotune(const an5d_problem* p, void* dA, const an5d_layout* layout, int mode,
                             int max_measured, an5d_config* best, double* best_gstencil) {
  Geom g;
  std::string err;
  if (!make_geom(p, g, err)) return fail(AN5D_E_INVALID, err);
  if (mode != AN5D_EXACT && mode != AN5D_FAST) return fail(AN5D_E_INVALID, "bad mode");
  if (max_measured <= 0) max_measured = 24;
  auto vs = variants_of(p->stencil, p->dtype);
  if (vs.empty()) return fail(AN5D_E_UNSUPPORTED, "no kernels for this stencil/dtype");
  struct Cand {
    int variant, bt, sl;
    double model, measured;
  };
  std::vector<Cand> cands;
  const long long n0 = g.D0 - 2 * g.r;
  const int sls[] = {0, 16, 32, 64, 128, 256};
  for (int vi = 0; vi < (int)vs.size(); vi++) {
    const Variant& v = *vs[vi];
    for (int bt = 1; bt <= v.btmax; bt++) {
      for (int sl : sls) {
        if (v.kind == kDirect && sl != 0) continue;
        if (sl > n0) continue;
        LaunchPlan lp;
        if (!plan_launch(v, bt, g, sl, lp, err, mode)) continue;
        cands.push_back(Cand{vi, bt, sl, model_gstencil(p, v, bt, sl, mode), 0.0});
      }
    }
  }
  if (cands.empty()) return fail(AN5D_E_UNSUPPORTED, "no valid configuration");
  std::sort(cands.begin(), cands.end(), [](const Cand& a, const Cand& b) { return a.model > b.model; });
  if ((int)cands.size() > max_measured) cands.resize(max_measured);

  // device buffers: the caller's, or a scratch grid of uniform [0,1) values
  const size_t esz = p->dtype == AN5D_FLOAT64 ? 8 : 4;
  an5d_layout L = layout ? *layout : reference_layout(g);
  void* buf = dA;
  void* owned = nullptr;
  hipError_t e;
  if (!buf) {
    std::vector<unsigned char> host((size_t)2 * g.area * esz);
    int rc = an5d_init_grid(p->dtype, host.data(), g.area, n0, 1);
    if (rc) return rc;
    e = hipMalloc(&owned, host.size());
    if (e != hipSuccess) return fail(AN5D_E_ALLOC, "hipMalloc(autotune grid)");
    e = hipMemcpy(owned, host.data(), host.size(), hipMemcpyHostToDevice);
    if (e != hipSuccess) { (void)hipFree(owned); return hip_fail(e, "hipMemcpy(autotune grid)"); }
    buf = owned;
    L = reference_layout(g);
  }
  // Candidates are timed on the field the run would hold halfway through
  // (prob->timesteps / 2 steps, at most 512, advanced first with the model's
  // favourite): with the source coefficients many fields decay, and a field
  // that has settled into zeros with thin subnormal bands at the boundary
  // (where the exact division takes its IEEE path) favours other tilings
  // than the random initial field.  Every candidate starts from a copy of
  // that field, so all of them see the same data.
  void* snap = nullptr;
  const size_t grid_bytes = (size_t)2 * L.buf_stride * esz;
  // a caller's grid is put back as it was when the search ends
  std::vector<unsigned char> caller_copy;
  if (!owned) {
    try {
      caller_copy.resize(grid_bytes);
    } catch (...) {
      return fail(AN5D_E_ALLOC, "no host memory to keep the caller's grid during the search");
    }
    e = hipMemcpy(caller_copy.data(), buf, grid_bytes, hipMemcpyDeviceToHost);
    if (e != hipSuccess) return hip_fail(e, "hipMemcpy(autotune keep grid)");
  }
  {
    int warm = p->timesteps / 2;
    if (warm > 512) warm = 512;
    warm -= warm % 2;  // keep buffer 0 current
    if (warm > 0) {
      an5d_problem w = *p;
      w.timesteps = warm;
      an5d_config wc{cands[0].bt, cands[0].variant, cands[0].sl, mode};
      int rc0 = an5d_forward_device(&w, buf, &L, &wc, nullptr);
      if (rc0) {
        if (owned) (void)hipFree(owned);
        if (!caller_copy.empty())
          (void)hipMemcpy(buf, caller_copy.data(), grid_bytes, hipMemcpyHostToDevice);
        return rc0;
      }
    }
    // without room for the copy the candidates run on in sequence
    if (hipMalloc(&snap, grid_bytes) != hipSuccess) {
      (void)hipGetLastError();
      snap = nullptr;
    } else if (hipMemcpy(snap, buf, grid_bytes, hipMemcpyDeviceToDevice) != hipSuccess) {
      (void)hipFree(snap);
      snap = nullptr;
    }
  }
  hipEvent_t e0, e1;
  (void)hipEventCreate(&e0);
  (void)hipEventCreate(&e1);
  const double cells = (double)(g.D0 - 2 * g.r) * (g.dim == 3 ? g.D1 - 2 * g.r : 1) * (g.D2 - 2 * g.r);
  int rc = AN5D_OK;
  for (Cand& c : cands) {
    an5d_config cfg{c.bt, c.variant, c.sl, mode};
    an5d_problem q = *p;
    q.timesteps = c.bt;
    if (snap) {
      e = hipMemcpyAsync(buf, snap, grid_bytes, hipMemcpyDeviceToDevice, nullptr);
      if (e != hipSuccess) { rc = hip_fail(e, "autotune restore"); break; }
    }
    // warm up, then time repeated single launches (buffers alternate)
    rc = an5d_launch(&q, buf, &L, &cfg, c.bt, 0, nullptr);
    if (rc) break;
    int reps = 2;
    double sec = 0.0;
    for (int attempt = 0; attempt < 6; attempt++) {
      (void)hipEventRecord(e0, nullptr);
      for (int r = 0; r < reps && rc == AN5D_OK; r++) rc = an5d_launch(&q, buf, &L, &cfg, c.bt, r, nullptr);
      (void)hipEventRecord(e1, nullptr);
      e = hipEventSynchronize(e1);
      if (rc || e != hipSuccess) break;
      float ms = 0.f;
      (void)hipEventElapsedTime(&ms, e0, e1);
      sec = ms / 1e3;
      if (sec > 2e-3) break;
      reps *= 2;
    }
    if (rc) break;
    if (e != hipSuccess) { rc = hip_fail(e, "autotune timing"); break; }
    c.measured = cells * c.bt * reps / sec / 1e9;
  }
  (void)hipEventDestroy(e0);
  (void)hipEventDestroy(e1);
  if (snap) (void)hipFree(snap);
  if (owned) (void)hipFree(owned);
  if (!caller_copy.empty()) {
    e = hipMemcpy(buf, caller_copy.data(), grid_bytes, hipMemcpyHostToDevice);
    if (e != hipSuccess && !rc) rc = hip_fail(e, "hipMemcpy(autotune restore grid)");
  }
  if (rc) return rc;
  const Cand* win = &cands[0];
  for (const Cand& c : cands)
    if (c.measured > win->measured) win = &c;
  Entry ent{};
  ent.stencil = p->stencil;
  ent.dtype = p->dtype;
  ent.mode = mode;
  for (int d = 0; d < 3; d++) ent.n[d] = d < g.dim ? p->n[d] : 0;
  ent.variant = win->variant;
  ent.bt = win->bt;
  ent.sl = win->sl;
  ent.gstencil = win->measured;
  {
    Table& t = table();
    std::lock_guard<std::mutex> lk(t.mu);
    ensure_loaded(t);
    upsert(t, ent);
  }
  if (best) *best = an5d_config{win->bt, win->variant, win->sl, mode};
  if (best_gstencil) *best_gstencil = win->measured;
  return AN5D_OK;
}

extern "C" int an5d_tune_load(const char* path) {
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.mu);
  t.loaded = true;
  return load_locked(t, path ? path : default_path(), false);
}

extern "C" int an5d_tune_save(const char* path) {
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.mu);
  ensure_loaded(t);
  const std::string p = path ? path : default_path();
  FILE* f = fopen(p.c_str(), "w");
  if (!f) return fail(AN5D_E_INVALID, "cannot write " + p);
  fprintf(f, "# stencil dtype mode n0 n1 n2 variant bt sl gstencil_per_s regcap (autotuner results; regcap:\n"
             "# registers per lane the chosen kernel was compiled to, 0 = no cap; informational)\n");
  for (const Entry& e : t.rows) {
    auto vs = variants_of(e.stencil, e.dtype);
    const int cap = e.variant >= 0 && e.variant < (int)vs.size() && e.bt * 2 > vs[e.variant]->btmax
                        ? vs[e.variant]->regcap
                        : 0;
    fprintf(f, "%s %d %d %lld %lld %lld %d %d %d %.1f %d\n", stencil_rt(e.stencil).name.c_str(), e.dtype,
            e.mode, e.n[0], e.n[1], e.n[2], e.variant, e.bt, e.sl, e.gstencil, cap);
  }
  fclose(f);
  return AN5D_OK;
}

extern "C" int an5d_tune_clear(void) {
  Table& t = table();
  std::lock_guard<std::mutex> lk(t.mu);
  t.rows.clear();
  t.loaded = true;
  return AN5D_OK;
}

// Batched forward sweep over many independent 2D grids: host side.
// No counterpart in the reference, which runs one grid per binary.
#include <string.h>

#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "batch.hpp"
#include "engine.hpp"
#include "internal.hpp"

namespace an5d {

// batch_inst.hip
hipError_t batch_launch_level(const BatchRec* recs, const BatchUnit* units, int nunits, int* counter,
                              int grid, hipStream_t st);
hipError_t batch_launch_reset(int* counters, int n, hipStream_t st);
int batch_kind_of(int stencil);
int batch_max_bt();
int batch_ck();

}  // namespace an5d

using namespace an5d;

struct an5d_batch {
  int n = 0;
  int levels = 0;
  std::vector<BatchRec*> d_recs;
  std::vector<BatchUnit*> d_units;
  std::vector<int> nunits;
  int* d_counters = nullptr;
  hipGraph_t graph = nullptr;
  hipGraphExec_t exec = nullptr;
  hipStream_t cap = nullptr;
  long long total_units = 0;
};

static void batch_free(an5d_batch* b) {
  if (!b) return;
  for (auto* p : b->d_recs) (void)hipFree(p);
  for (auto* p : b->d_units) (void)hipFree(p);
  (void)hipFree(b->d_counters);
  if (b->exec) (void)hipGraphExecDestroy(b->exec);
  if (b->graph) (void)hipGraphDestroy(b->graph);
  if (b->cap) (void)hipStreamDestroy(b->cap);
  delete b;
}

extern "C" int an5d_batch_create(an5d_batch** out, int n, const an5d_problem* probs,
                                 void* const* dA, const an5d_config* cfgs) {
  if (!out || n < 1 || !probs || !dA) return fail(AN5D_E_INVALID, "batch needs n >= 1 problems and buffers");
  *out = nullptr;
  an5d_batch* b = new an5d_batch();
  b->n = n;
  struct Inst {
    Geom g;
    std::vector<int> steps;
    const void* coef;
    int has_div, kind, sl;
  };
  std::vector<Inst> inst(n);
  std::string err;
  for (int i = 0; i < n; i++) {
    const an5d_problem* p = &probs[i];
    Inst& I = inst[i];
    if (!make_geom(p, I.g, err)) { batch_free(b); return fail(AN5D_E_INVALID, "instance " + std::to_string(i) + ": " + err); }
    if (I.g.dim != 2 || p->dtype != AN5D_FLOAT32) {
      batch_free(b);
      return fail(AN5D_E_UNSUPPORTED, "batched sweeps take 2D FP32 problems");
    }
    I.kind = batch_kind_of(p->stencil);
    if (I.kind < 0) {
      batch_free(b);
      return fail(AN5D_E_UNSUPPORTED, "stencil " + stencil_rt(p->stencil).name + " is not in the batched kernel");
    }
    if (!dA[i] || p->timesteps < 0) { batch_free(b); return fail(AN5D_E_INVALID, "bad buffer or timesteps"); }
    std::vector<unsigned char> cb;
    double gs = 1.0;
    if (!build_coefs(p, cb, I.has_div, gs, err)) { batch_free(b); return fail(AN5D_E_INVALID, err); }
    int rc = device_coefs(cb, &I.coef);
    if (rc) { batch_free(b); return rc; }
    int bt = (cfgs && cfgs[i].bt > 0) ? cfgs[i].bt : batch_max_bt();
    if (bt > batch_max_bt()) { batch_free(b); return fail(AN5D_E_UNSUPPORTED, "batched bt is at most 4"); }
    I.sl = (cfgs && cfgs[i].sl > 0) ? cfgs[i].sl : 64;
    int nl = an5d_schedule(p->timesteps, bt, nullptr, 0);
    I.steps.resize(nl);
    an5d_schedule(p->timesteps, bt, I.steps.data(), nl);
    if (nl > b->levels) b->levels = nl;
  }
  hipError_t e = hipMalloc(&b->d_counters, sizeof(int) * (b->levels > 0 ? b->levels : 1));
  if (e != hipSuccess) { batch_free(b); return fail(AN5D_E_ALLOC, "hipMalloc(batch counters)"); }
  const int ck = batch_ck();
  for (int l = 0; l < b->levels; l++) {
    std::vector<BatchRec> recs;
    std::vector<BatchUnit> units;
    for (int i = 0; i < n; i++) {
      Inst& I = inst[i];
      if (l >= (int)I.steps.size()) continue;
      const int k = I.steps[l];
      const int R = I.g.r;
      const long long nk = I.g.D2 - 2 * R, n0 = I.g.D0 - 2 * R;
      const long long SK = 64LL * ck - 2LL * k * R;
      const long long ntk = (nk + SK - 1) / SK;
      const long long sl = I.sl < n0 ? I.sl : n0;
      const long long nseg = (n0 + sl - 1) / sl;
      const size_t gbytes = (size_t)I.g.area * 4;
      BatchRec r;
      r.src = static_cast<char*>(dA[i]) + (size_t)(l % 2) * gbytes;
      r.dst = static_cast<char*>(dA[i]) + (size_t)((l + 1) % 2) * gbytes;
      r.coef = I.coef;
      r.has_div = I.has_div;
      r.kind = I.kind;
      r.bt = k;
      r.D0 = (int)I.g.D0;
      r.D2 = (int)I.g.D2;
      r.s0 = I.g.D2;
      r.ntk = (int)ntk;
      r.nseg = (int)nseg;
      r.sl = (int)sl;
      r.SK = (int)SK;
      const int rid = (int)recs.size();
      recs.push_back(r);
      for (long long t = 0; t < ntk * nseg; t++) units.push_back(BatchUnit{rid, (int)t});
    }
    BatchRec* dr = nullptr;
    BatchUnit* du = nullptr;
    e = hipMalloc(&dr, sizeof(BatchRec) * recs.size());
    if (e == hipSuccess) e = hipMalloc(&du, sizeof(BatchUnit) * units.size());
    if (e == hipSuccess) e = hipMemcpy(dr, recs.data(), sizeof(BatchRec) * recs.size(), hipMemcpyHostToDevice);
    if (e == hipSuccess) e = hipMemcpy(du, units.data(), sizeof(BatchUnit) * units.size(), hipMemcpyHostToDevice);
    b->d_recs.push_back(dr);
    b->d_units.push_back(du);
    b->nunits.push_back((int)units.size());
    b->total_units += (long long)units.size();
    if (e != hipSuccess) { batch_free(b); return fail(AN5D_E_ALLOC, "batch tables"); }
  }
  *out = b;
  return AN5D_OK;
}

// Enqueue every level on `stream`: the first call captures the reset kernel
// and the level launches into a HIP graph, later calls replay it.
extern "C" int an5d_batch_run(an5d_batch* b, void* stream) {
  if (!b) return fail(AN5D_E_INVALID, "null batch");
  hipStream_t st = static_cast<hipStream_t>(stream);
  hipError_t e;
  if (!b->exec) {
    e = hipStreamCreateWithFlags(&b->cap, hipStreamNonBlocking);
    if (e != hipSuccess) return hip_fail(e, "hipStreamCreate");
    e = hipStreamBeginCapture(b->cap, hipStreamCaptureModeThreadLocal);
    if (e != hipSuccess) return hip_fail(e, "hipStreamBeginCapture");
    hipError_t le = batch_launch_reset(b->d_counters, b->levels, b->cap);
    for (int l = 0; l < b->levels && le == hipSuccess; l++) {
      const int waves = b->nunits[l];
      int grid = (waves + 3) / 4;
      if (grid > 2048) grid = 2048;
      if (grid < 1) grid = 1;
      le = batch_launch_level(b->d_recs[l], b->d_units[l], b->nunits[l], b->d_counters + l, grid, b->cap);
    }
    e = hipStreamEndCapture(b->cap, &b->graph);
    if (le != hipSuccess) return hip_fail(le, "batch level launch");
    if (e != hipSuccess) return hip_fail(e, "hipStreamEndCapture");
    e = hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0);
    if (e != hipSuccess) return hip_fail(e, "hipGraphInstantiate");
  }
  e = hipGraphLaunch(b->exec, st);
  if (e != hipSuccess) return hip_fail(e, "hipGraphLaunch");
  return AN5D_OK;
}

extern "C" int an5d_batch_info(const an5d_batch* b, int* levels, long long* units) {
  if (!b) return fail(AN5D_E_INVALID, "null batch");
  if (levels) *levels = b->levels;
  if (units) *units = b->total_units;
  return AN5D_OK;
}

extern "C" int an5d_batch_destroy(an5d_batch* b) {
  batch_free(b);
  return AN5D_OK;
}

// Declarations shared by the host translation units of liban5d.
#pragma once

#include <mutex>
#include <string>
#include <vector>

#include "../../include/an5d.h"
#include "engine.hpp"
#include "stencil_defs.hpp"

namespace an5d {

struct StencilRT {
  std::string name;
  int dim, rad, fpp, ntaps, divisor;
  bool linear;
  std::vector<TapDef> taps;
};

struct Geom {
  int dim, r;
  long long D0, D1, D2, area;  // full extents (2D: D1 == 1)
};

struct LaunchPlan {
  LaunchFn fn = nullptr;
  LaunchArgs args{};
  dim3 grid, block;
  size_t lds = 0;
};

struct TunedChoice {
  int variant, bt, sl;
};

int fail(int code, const std::string& msg);
int hip_fail(hipError_t e, const char* what);
const StencilRT& stencil_rt(int id);
std::vector<const Variant*> variants_of(int stencil, int dtype);
bool make_geom(const an5d_problem* p, Geom& g, std::string& err);
an5d_layout reference_layout(const Geom& g);
an5d_layout padded_layout(const Geom& g);
int resolve(const an5d_problem* p, const an5d_config* cfg, const Variant** vout, int* btout,
            int* slout);
// plan_launch's mode: AN5D_EXACT, AN5D_FAST, or kLaunchShare (exact, the
// shared-product kernels)
constexpr int kLaunchShare = 2;
bool plan_launch(const Variant& v, int bt, const Geom& g, int sl, LaunchPlan& lp,
                 std::string& err, int mode = 0);
bool coefs_shareable(const an5d_problem* p, const std::vector<unsigned char>& bytes);
bool build_coefs(const an5d_problem* p, std::vector<unsigned char>& bytes, int& has_div,
                 double& gscale, std::string& err, int mode = 0);
int device_coefs(const std::vector<unsigned char>& bytes, const void** out);
// The current device's staging grid of at least `bytes`; `hold` keeps other
// host calls on that device from using or regrowing it until released.
int device_grid(size_t bytes, void** out, std::unique_lock<std::mutex>& hold);
int run_schedule(const an5d_problem* p, void* dA, const an5d_layout* layout,
                 const an5d_config* cfg, int steps_only, int c0_first, hipStream_t st);
double sb_time_ms();
int device_cus();  // compute units of the current device (0 if unknown)

// autotuner table (autotune.cpp)
bool tuned_lookup(const an5d_problem* p, int mode, TunedChoice* out);
double model_gstencil(const an5d_problem* p, const Variant& v, int bt, int sl, int mode);

}  // namespace an5d

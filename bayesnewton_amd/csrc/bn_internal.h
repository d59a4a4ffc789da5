// Library-internal declarations shared by the HIP translation units.
#pragma once
#include <string>

#include <hip/hip_runtime.h>

#include "../../include/bn_hip.h"
#include "bn_lanes.h"
#include "bn_plan_host.h"

struct bn_plan {
  bn::PlanHost host;
  bn::PlanView view;
  void* dev = nullptr;  // one device allocation holding every table
};

namespace bn {
// Kernel ids of the optional per-launch timing (bn_timing_*).
enum KernelId : int32_t {
  KID_FILTER_REDUCE = 0, KID_FSCAN_TILE, KID_FSCAN_SERIES, KID_FILTER_PASS, KID_SMOOTHER_REDUCE,
  KID_SSCAN_TILE, KID_SSCAN_SERIES, KID_SMOOTHER_PASS, KID_SERIES_SUM, KID_ST_CONDITIONAL,
  KID_ST_SITES, KID_ST_FILTER, KID_ST_SMOOTHER, KID_ST_SUM, KID_SMOOTHER_MBF, KID_MF_PSEUDO, KID_MF_FILTER,
  KID_MF_SDIAG, KID_MF_BLOCKS, KID_MF_SMOOTHER, KID_IH_TIED, KID_IH_DARE, KID_IH_SINV, KID_IH_FILTER,
  KID_IH_SMOOTHER, KID_COUNT
};
// Brackets the launches of its scope with events on `s` when timing is enabled.
struct TimedLaunch {
  int32_t kid;
  hipStream_t st;
  hipEvent_t a = nullptr;
  TimedLaunch(int32_t k, hipStream_t s);
  ~TimedLaunch();
};
#define BN_TIMED(kid, st) ::bn::TimedLaunch bn_tl_##kid(kid, st)

// Records the thread's last error message and returns `code`.
int set_error(int code, const std::string& msg);
// Per-chunk FP64 scratch arrays (which = 0, 1) inside a workspace sized by
// bn_workspace_bytes for this plan and kernel.
double* site_partials(const bn_plan* plan, const bn_kernel_desc* kern, void* workspace, int which);
}  // namespace bn

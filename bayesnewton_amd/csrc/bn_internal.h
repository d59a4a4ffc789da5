// Library-internal declarations shared by the HIP translation units.
#pragma once
#include <string>

#include "../../include/bn_hip.h"
#include "bn_lanes.h"
#include "bn_plan_host.h"

struct bn_plan {
  bn::PlanHost host;
  bn::PlanView view;
  void* dev = nullptr;  // one device allocation holding every table
};

namespace bn {
// Records the thread's last error message and returns `code`.
int set_error(int code, const std::string& msg);
// Per-chunk FP64 scratch arrays (which = 0, 1) inside a workspace sized by
// bn_workspace_bytes for this plan and kernel.
double* site_partials(const bn_plan* plan, const bn_kernel_desc* kern, void* workspace, int which);
}  // namespace bn

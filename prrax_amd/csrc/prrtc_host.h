// prrtc_host.h -- host-side helpers shared by the library's translation units.
#pragma once
#include <hip/hip_runtime.h>

#include <string>

#include "prrtc_amd.h"
#include "prrtc_common.h"

namespace prrtc {

struct DevInfo {
    int ok = 0;
    int cus = 0;
    int max_lds = 0;
    float wall_khz = 0.0f;
};

extern thread_local std::string g_err;

// Records msg as the thread's last error and returns code.
int fail_with(int code, const std::string& msg);
// Cached device attributes (CU count, LDS per work-group, wall-clock rate).
DevInfo& dev_info(int dev);
RobotDev robot_dev(const prrtc_robot& r);
WorldDev world_dev(const prrtc_world& w, bool per_query);
int check_robot(const prrtc_robot* r);
int check_world(const prrtc_world* w);
// LDS layout of a check kernel's model with nt threads (shrinking nt until it fits), and the
// offset of its scan scratch.
int model_lds(const prrtc_robot& r, const prrtc_world& w, int& nt, int max_lds, ModelLayout& ml, size_t& lds,
              bool per_lane);
size_t scan_offset(const ModelLayout& ml);

}  // namespace prrtc

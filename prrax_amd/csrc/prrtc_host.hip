// prrtc_host.hip -- host-side helpers shared by the library's entry points: error reporting,
// cached device attributes, argument checks and C-ABI -> kernel-argument conversion.
#include <hip/hip_runtime.h>

#include <string>

#include "prrtc_amd.h"
#include "prrtc_host.h"

namespace prrtc {

thread_local std::string g_err;

int fail_with(int code, const std::string& msg) {
    g_err = msg;
    return code;
}

DevInfo& dev_info(int dev) {
    static DevInfo info[64];
    DevInfo& d = info[dev & 63];
    if (!d.ok) {
        int v = 0;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess) v = 1;
        d.cus = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeMaxSharedMemoryPerBlock, dev) != hipSuccess) v = 65536;
        d.max_lds = v;
        if (hipDeviceGetAttribute(&v, hipDeviceAttributeWallClockRate, dev) != hipSuccess) v = 100000;
        d.wall_khz = static_cast<float>(v > 0 ? v : 100000);
        d.ok = 1;
    }
    return d;
}

RobotDev robot_dev(const prrtc_robot& r) {
    RobotDev d;
    d.twists = r.fk_twists; d.parent_tf = r.fk_parent_tf; d.parent_idx = r.fk_parent_idx; d.act_idx = r.fk_act_idx;
    d.mimic_mul = r.fk_mimic_mul; d.mimic_off = r.fk_mimic_off; d.mimic_act_idx = r.fk_mimic_act_idx;
    d.topo_inv = r.fk_topo_inv; d.n_joints = r.n_joints; d.sphere_link_idx = r.sphere_link_idx;
    d.sphere_local = r.sphere_local; d.sphere_radius = r.sphere_radius; d.n_spheres = r.n_robot_spheres;
    return d;
}

WorldDev world_dev(const prrtc_world& w, bool per_query) {
    WorldDev d;
    d.ws = w.world_spheres; d.wc = w.world_capsules; d.wb = w.world_boxes; d.wh = w.world_halfspaces; d.sp = w.self_pairs;
    d.n_ws = w.n_world_spheres; d.n_wc = w.n_world_capsules; d.n_wb = w.n_world_boxes; d.n_wh = w.n_world_halfspaces;
    d.n_sp = w.n_self_pairs;
    d.s_ws = per_query ? 4LL * d.n_ws : 0; d.s_wc = per_query ? 7LL * d.n_wc : 0; d.s_wb = per_query ? 15LL * d.n_wb : 0;
    d.s_wh = per_query ? 6LL * d.n_wh : 0; d.s_sp = per_query ? 2LL * d.n_sp : 0;
    d.cnt[0] = per_query ? w.world_spheres_count : nullptr;
    d.cnt[1] = per_query ? w.world_capsules_count : nullptr;
    d.cnt[2] = per_query ? w.world_boxes_count : nullptr;
    d.cnt[3] = per_query ? w.world_halfspaces_count : nullptr;
    d.cnt[4] = per_query ? w.self_pairs_count : nullptr;
    return d;
}

int check_robot(const prrtc_robot* r) {
    if (!r) return fail_with(PRRTC_E_INVALID, "robot is null");
    if (r->n_joints < 0 || r->n_joints > PRRTC_MAX_JOINTS)
        return fail_with(PRRTC_E_INVALID, "n_joints must be in [0, 64]");
    if (r->n_robot_spheres < 0 || r->n_robot_spheres > 65535)
        return fail_with(PRRTC_E_INVALID, "n_robot_spheres must be in [0, 65535]");
    if (r->n_joints > 0 && (!r->fk_twists || !r->fk_parent_tf || !r->fk_parent_idx || !r->fk_act_idx ||
                            !r->fk_mimic_mul || !r->fk_mimic_off || !r->fk_mimic_act_idx || !r->fk_topo_inv))
        return fail_with(PRRTC_E_INVALID, "FK arrays must be non-null when n_joints > 0");
    if (r->n_robot_spheres > 0 && (!r->sphere_link_idx || !r->sphere_local || !r->sphere_radius))
        return fail_with(PRRTC_E_INVALID, "sphere arrays must be non-null when n_robot_spheres > 0");
    return PRRTC_OK;
}

int check_world(const prrtc_world* w) {
    if (!w) return fail_with(PRRTC_E_INVALID, "world is null");
    if (w->n_world_spheres < 0 || w->n_world_capsules < 0 || w->n_world_boxes < 0 || w->n_world_halfspaces < 0 ||
        w->n_self_pairs < 0)
        return fail_with(PRRTC_E_INVALID, "negative world count");
    if ((w->n_world_spheres && !w->world_spheres) || (w->n_world_capsules && !w->world_capsules) ||
        (w->n_world_boxes && !w->world_boxes) || (w->n_world_halfspaces && !w->world_halfspaces) ||
        (w->n_self_pairs && !w->self_pairs))
        return fail_with(PRRTC_E_INVALID, "world array null with nonzero count");
    // The coarse-checker arrays: sizes consistent, pointers present where sizes are nonzero.
    if (w->n_coarse_spheres < 0 || w->n_fine_link_offsets < 0 || w->n_coarse_self_pairs < 0 ||
        w->n_self_pair_ranges < 0)
        return fail_with(PRRTC_E_INVALID, "negative coarse-array count");
    if (w->n_coarse_spheres &&
        (!w->coarse_sphere_link_idx || !w->coarse_sphere_local || !w->coarse_sphere_radius))
        return fail_with(PRRTC_E_INVALID, "coarse sphere arrays null with nonzero count");
    if ((w->n_fine_link_offsets && !w->fine_link_offsets) || (w->n_coarse_self_pairs && !w->coarse_self_pairs) ||
        (w->n_self_pair_ranges && !w->self_pair_ranges))
        return fail_with(PRRTC_E_INVALID, "coarse array null with nonzero count");
    return PRRTC_OK;
}

}  // namespace prrtc

extern "C" const char* prrtc_last_error(void) { return prrtc::g_err.c_str(); }

// PRRTC_SRC_HASH: the build's hash of the library's source files (prrax_amd/build.py).
#ifndef PRRTC_SRC_HASH
#define PRRTC_SRC_HASH "unknown"
#endif

extern "C" const char* prrtc_version(void) { return "prrtc_amd 0.3.0 (gfx950) src " PRRTC_SRC_HASH; }

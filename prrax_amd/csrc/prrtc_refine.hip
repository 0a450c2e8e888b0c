// prrtc_refine.hip -- differentiable sphere-penetration cost through FK, and path refinement.
//
// No reference counterpart (the reference has no cost gradient; its nearest relatives are the
// margin predicate, prrtc_planner.cu:314-387, and pyroffi's absent collision-cost kernel called
// from benchmark_collision_checkers.py:11-15).  Cost of one configuration q:
//
//   C(q) = sum over robot spheres i, world spheres w of softplus_b(r_i + r_w - |c_i(q) - o_w|),
//   softplus_b(x) = log(1 + exp(b x)) / b,
//
// with c_i(q) the sphere centre through FK.  The adjoint goes through the kinematic tree
// without storing per-link transforms: FK is recomputed in the gradient kernel, and
//   dC/dc_i = -sum_w sigmoid(b x_iw) (c_i - o_w) / |c_i - o_w|            (force g_i)
//   F_j = sum of g_i over spheres on joint j's subtree, M_j = sum of c_i x g_i over the same,
//   dC/dtheta_j = w_j . (M_j - t_j x F_j) + (R_j v_j) . F_j,
// where (R_j, t_j) is joint j's world frame and (v_j, omega_j) its twist (w_j = R_j omega_j):
// moving theta_j moves every point x of the subtree by w_j x (x - t_j) + R_j v_j.  Joint
// gradients then flow to the configuration through the act / mimic maps of the FK.
//
// Layout: a wave takes a batch of nb configurations (FK one per lane), then the nb * n_spheres
// (configuration, sphere) items, 64 per wave step, configuration-minor (item k: sphere k / nb,
// configuration k % nb), so every lane is busy and the items of one configuration sit in the
// lanes congruent mod nb: a butterfly over lane offsets nb, 2 nb, ..., 32 sums them, in a fixed
// order (results are deterministic).  Each item loops over the world spheres (LDS broadcast);
// pairs beyond reach are screened out in squared form.  The adjoint needs, per joint j of the
// configuration, sum over spheres i on j's subtree of g_i . (w_j x (c_i - t_j) + R_j v_j): each
// item adds its term for every joint in its frame's ancestor mask, with the joint's world axis,
// origin and translation direction precomputed per (configuration, joint) in the wave's LDS
// area, and the per-joint sums go to LDS by the lane that holds the configuration.  FP32.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>

#include "prrtc_amd.h"
#include "prrtc_common.h"
#include "prrtc_host.h"

namespace prrtc {

// Algorithmic flops (FMA = 2) per robot-world sphere pair: the screen 10 (difference 3, squared
// norm 5, reach and its square 2) for every pair; for pairs inside the reach the distance 5
// (square root as d^2 / d, penetration 2, scale 1 ... ); for pairs with z >= kFarZ the softplus 5
// more (exp, log1p, max, add, scale) and the gradient 11 more (sigmoid 3, 1/d and scale 2,
// accumulate 6); per (sphere, joint) term of the adjoint 17 (lever arm 3, axis cross 6, plus
// translation 3, dot 5); per joint 30 for its axis and translation direction.
constexpr unsigned kFlopsPairScreen = 10, kFlopsPairDist = 5, kFlopsPairCost = 5, kFlopsPairGrad = 11;
constexpr unsigned kFlopsAdjTerm = 17, kFlopsJointAdj = 30;
// A pair with z = beta (r_i + r_w - d) < kFarZ contributes below exp(kFarZ) ~ 1e-13 to the cost
// and its derivative: skipped (z < kFarZ  <=>  d > r_i + r_w - kFarZ / beta).
constexpr float kFarZ = -30.0f;
constexpr int kMaxCostBatch = 16;  // configurations per wave batch (nb, a power of two)

struct CostArgs {
    RobotDev robot;
    WorldDev world;
    ModelLayout ml;
    size_t off_scan, off_anc, off_jw;
    int jw_stride;  // float4 per wave in the joint area: 3 * nb * n_joints, then nb * n_joints sums
    int nb, log_nb;
    const float* configs;
    int n, dim;
    float beta;
    float* out_cost;
    float* out_grad;  // [n, dim] or null
    unsigned long long* flops;
};

__device__ __forceinline__ float wave_sum(float v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Sum over the lanes congruent to this one mod nb (offsets nb .. 32 of the butterfly).
__device__ __forceinline__ float class_sum(float v, int nb) {
    for (int o = nb; o < kWave; o <<= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// log(1 + e) for e in [0, 1]: below 2^-6 the series e - e^2/2 + e^3/3 - e^4/4 (truncation below
// e^5/5, relative 2e-9), above it log2(1 + e) ln 2 (the rounding of 1 + e is then below 4e-6
// of the result); both are computed and one is selected, so the lanes do not diverge.
__device__ __forceinline__ float log1p_unit(float e) {
    const float ser = e * __builtin_fmaf(e, __builtin_fmaf(e, __builtin_fmaf(e, -0.25f, 0x1.555556p-2f), -0.5f), 1.0f);
    const float lg = __builtin_amdgcn_logf(1.0f + e) * 0x1.62e430p-1f;
    return e < 0x1p-6f ? ser : lg;
}

template <bool GRAD>
__global__ __launch_bounds__(256) void sphere_cost_kernel(CostArgs a) {
    extern __shared__ __align__(16) char lds[];
    int* scan = reinterpret_cast<int*>(lds + a.off_scan);
    Model M;
    stage_model(M, lds, a.ml, a.robot, a.world, 0, a.dim, 0, 0.0f, scan);
    const int lane = threadIdx.x & (kWave - 1), wid = threadIdx.x / kWave;
    const int nj = M.n_joints, nrs = M.n_spheres, nws = M.n_ws, nb = a.nb, lnb = a.log_nb;
    // ancestor masks: bit k of anc[j] set iff joint k lies on the path from the root to joint j
    unsigned long long* anc = reinterpret_cast<unsigned long long*>(lds + a.off_anc);
    if (GRAD && threadIdx.x == 0) {
        for (int j = 0; j < nj; ++j) {
            const int p = M.joints[j].parent;
            anc[j] = (p >= 0 ? anc[p] : 0ull) | (1ull << j);
        }
    }
    __syncthreads();
    const WaveArea W = wave_area(lds, a.ml, nj);
    LaneScratch L;
    L.q = W.q + lane;
    L.T = W.T + lane;
    L.nt = kFBP;
    float4* JW = reinterpret_cast<float4*>(lds + a.off_jw) + static_cast<size_t>(wid) * a.jw_stride;
    float* JS = reinterpret_cast<float*>(JW + 3 * nb * nj);  // per (configuration, joint) sums
    const float beta = a.beta, inv_beta = 1.0f / a.beta, reach = -kFarZ / a.beta;
    const int wpb = blockDim.x / kWave;
    const int total_waves = gridDim.x * wpb;
    unsigned long long n_act = 0, n_in = 0, n_near = 0, n_adj = 0;  // this lane's counts
    const int nbat = (a.n + nb - 1) / nb;
    const int items = nrs << lnb;
    for (int b = blockIdx.x * wpb + wid; b < nbat; b += total_waves) {
        const int c0 = b * nb;
        const int ncfg = min(nb, a.n - c0);
        if (lane < ncfg) {
            for (int d = 0; d < a.dim; ++d) L.q[d * kFBP] = a.configs[static_cast<size_t>(c0 + lane) * a.dim + d];
            if (nj > 0) fk_lane(M, L);
        }
        wave_sync();
        if (GRAD) {
            // (configuration, joint): world axis w (|omega| R a), origin t, translation R v
            for (int t = lane; t < nb * nj; t += kWave) {
                const int c = t / nj, j = t - c * nj;
                float4 w = make_float4(0.0f, 0.0f, 0.0f, 0.0f), o = w, u = w;
                const JointRec& J = M.joints[j];
                if (c < ncfg && J.type != kFixed) {
                    Mat3 R;
                    float tx, ty, tz;
                    frame_from(W.T + c, kFBP, j, R, tx, ty, tz);
                    o = make_float4(tx, ty, tz, 0.0f);
                    if (J.type == kRot || J.type == kScrew) {
                        const float3 v = mat_apply(R, 0.0f, 0.0f, 0.0f, J.wn * J.ax, J.wn * J.ay, J.wn * J.az);
                        w = make_float4(v.x, v.y, v.z, 0.0f);
                    }
                    if (J.type == kPrism || J.type == kScrew) {
                        const float3 v = mat_apply(R, 0.0f, 0.0f, 0.0f, J.vx, J.vy, J.vz);
                        u = make_float4(v.x, v.y, v.z, 0.0f);
                    }
                }
                JW[3 * t] = w;
                JW[3 * t + 1] = o;
                JW[3 * t + 2] = u;
                JS[t] = 0.0f;
            }
            wave_sync();
        }
        float cost_c = 0.0f;  // lane c < nb: cost of configuration c0 + c
        unsigned b_act = 0, b_in = 0, b_near = 0, b_adj = 0;
        for (int k0 = 0; k0 < items; k0 += kWave) {
            const int k = k0 + lane;
            const int s = k >> lnb, c = k & (nb - 1);
            float cost = 0.0f, gx = 0.0f, gy = 0.0f, gz = 0.0f, cx = 0.0f, cy = 0.0f, cz = 0.0f;
            unsigned long long am = 0ull;
            if (k < items && c < ncfg) {
                const int js = M.sphere_js[s];
                if (js >= 0) {
                    ++b_act;
                    if (GRAD) am = anc[js];
                    Mat3 R;
                    float tx, ty, tz;
                    frame_from(W.T + c, kFBP, js, R, tx, ty, tz);
                    const float4 sp = M.spheres[s];
                    const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
                    cx = p.x; cy = p.y; cz = p.z;
                    const float lim0 = sp.w + reach;
                    for (int w = 0; w < nws; ++w) {
                        const float4 o = M.wsph[w];
                        const float dx = cx - o.x, dy = cy - o.y, dz = cz - o.z;
                        const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
                        // beyond reach: d > lim (1e-5 relative slack for the rounding; a negative
                        // reach lets only coincident centres through, which the z test rejects)
                        const float lim = fmaxf(lim0 + o.w, 0.0f);
                        if (!(d2 <= (lim * lim) * 1.00001f)) continue;
                        ++b_in;
                        const float rs = d2 > 0.0f ? __builtin_amdgcn_rsqf(d2) : 0.0f;
                        const float d = d2 * rs;
                        const float z = beta * ((sp.w + o.w) - d);
                        if (z < kFarZ) continue;
                        ++b_near;
                        // softplus(z) = max(z, 0) + log1p(e), sigmoid(z) from the same e = exp(-|z|)
                        const float e = __builtin_amdgcn_exp2f(-fabsf(z) * 0x1.715476p+0f);
                        cost += (fmaxf(z, 0.0f) + log1p_unit(e)) * inv_beta;
                        if (GRAD) {
                            const float sg = (z >= 0.0f ? 1.0f : e) * __builtin_amdgcn_rcpf(1.0f + e);
                            const float kk = -sg * rs;  // 0 when the centres coincide
                            gx = __builtin_fmaf(kk, dx, gx);
                            gy = __builtin_fmaf(kk, dy, gy);
                            gz = __builtin_fmaf(kk, dz, gz);
                        }
                    }
                }
            }
            cost_c += class_sum(cost, nb);
            if (GRAD) {
                // this item's term for every joint of its frame's ancestor mask
                for (int j = 0; j < nj; ++j) {
                    if (M.joints[j].type == kFixed) continue;  // wave-uniform
                    float v = 0.0f;
                    if ((am >> j) & 1ull) {
                        ++b_adj;
                        const float4* e = JW + 3 * (c * nj + j);
                        const float4 w = e[0], o = e[1], u = e[2];
                        const float lx = cx - o.x, ly = cy - o.y, lz = cz - o.z;
                        const float vx = __builtin_fmaf(w.y, lz, -(w.z * ly)) + u.x;
                        const float vy = __builtin_fmaf(w.z, lx, -(w.x * lz)) + u.y;
                        const float vz = __builtin_fmaf(w.x, ly, -(w.y * lx)) + u.z;
                        v = __builtin_fmaf(gz, vz, __builtin_fmaf(gy, vy, gx * vx));
                    }
                    v = class_sum(v, nb);
                    if (lane < nb) JS[lane * nj + j] += v;
                }
            }
        }
        n_act += b_act;
        n_in += b_in;
        n_near += b_near;
        n_adj += b_adj;
        if (lane < ncfg) a.out_cost[c0 + lane] = cost_c;
        if (GRAD) {
            wave_sync();
            // joint sums to the configuration through the act / mimic maps, lanes over (c, d)
            for (int t = lane; t < nb * a.dim; t += kWave) {
                const int c = t / a.dim, d = t - c * a.dim;
                if (c >= ncfg) continue;
                float gd = 0.0f;
                for (int j = 0; j < nj; ++j) {
                    const JointRec& J = M.joints[j];
                    if (J.mimic >= 0) {
                        if (J.mimic == d) gd = __builtin_fmaf(J.mul, JS[c * nj + j], gd);
                    } else if (J.act == d) {
                        gd += JS[c * nj + j];
                    }
                }
                a.out_grad[static_cast<size_t>(c0 + c) * a.dim + d] = gd;
            }
        }
        wave_sync();
    }
    if (a.flops) {
        // FK and adjoint set-up per configuration, then the per-item and per-pair counts
        const unsigned long long mine =
            n_act * (kFlopsXf + static_cast<unsigned long long>(nws) * kFlopsPairScreen) + n_in * kFlopsPairDist +
            n_near * (kFlopsPairCost + (GRAD ? kFlopsPairGrad : 0u)) + n_adj * kFlopsAdjTerm;
        unsigned long long tot = mine;
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) tot += __shfl_xor(tot, o, kWave);
        if (lane == 0) {
            unsigned long long cfgs = 0;
            for (int b = blockIdx.x * wpb + wid; b < nbat; b += total_waves) cfgs += min(nb, a.n - b * nb);
            tot += cfgs * (static_cast<unsigned long long>(kFlopsJoint) * nj +
                           (GRAD ? static_cast<unsigned long long>(kFlopsJointAdj) * nj : 0ull));
            if (tot) global_add(a.flops, tot);
        }
    }
}

// Jacobi step of every interior waypoint: candidate = clamp(q - step * (dC/dq + dS/dq)), with
// S = w * sum_k |q_{k+1} - q_k|^2, so dS/dq_k = 2 w (2 q_k - q_{k-1} - q_{k+1}); end points stay.
__global__ void refine_update_kernel(const float* paths, const float* grad, int P, int L, int dim, float step,
                                     float smooth, const float* lo, const float* hi, float* cand) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(P) * L * dim;
    if (t >= total) return;
    const int d = static_cast<int>(t % dim);
    const int k = static_cast<int>((t / dim) % L);
    const float q = paths[t];
    if (k == 0 || k == L - 1) {
        cand[t] = q;
        return;
    }
    const float qm = paths[t - dim], qp = paths[t + dim];
    const float gs = 2.0f * smooth * ((2.0f * q - qm) - qp);
    float c = q - step * (grad[t] + gs);
    if (lo && hi) c = fminf(fmaxf(c, lo[d]), hi[d]);
    cand[t] = c;
}

// Acceptance of the moves.  A waypoint k may move to its candidate only if the candidate and
// kEdgeRes - 1 interior points of each of its two edges (to the current neighbours, at t = j /
// kEdgeRes) pass the exact predicate, so a refined path is free wherever the path it came from
// was checked.  Waypoints of one parity move together (their neighbours, of the other parity,
// stay put), the even ones first, then the odd ones against the moved even ones.
constexpr int kEdgeRes = 16;
constexpr int kMovePoints = 2 * kEdgeRes - 1;  // the candidate + both edges' interior points

// Interior waypoints of parity par per path: k = 1 + par + 2 j, k <= L - 2.
__host__ __device__ inline int parity_slots(int L, int par) { return (L - 2 - par + 1) / 2; }

// Configurations that decide the moves of parity par: [P][slots][kMovePoints][dim], slots =
// parity_slots(L, 0) (the larger count; unused slots of the odd phase hold the path's start).
__global__ void refine_move_points_kernel(const float* paths, const float* cand, int P, int L, int dim, int par,
                                          int slots, float* out) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    const long long total = static_cast<long long>(P) * slots * kMovePoints;
    if (t >= total) return;
    const int c = static_cast<int>(t % kMovePoints);
    const int j = static_cast<int>((t / kMovePoints) % slots);
    const int p = static_cast<int>(t / (static_cast<long long>(kMovePoints) * slots));
    const int k = 1 + par + 2 * j;
    float* o = out + t * dim;
    const float* row = paths + static_cast<size_t>(p) * L * dim;
    if (k > L - 2) {
        for (int d = 0; d < dim; ++d) o[d] = row[d];
        return;
    }
    const float* ck = cand + (static_cast<size_t>(p) * L + k) * dim;
    const float* a = row + static_cast<size_t>(k - 1) * dim;
    const float* b = row + static_cast<size_t>(k + 1) * dim;
    if (c == 0) {
        for (int d = 0; d < dim; ++d) o[d] = ck[d];
    } else if (c < kEdgeRes) {  // edge (a, candidate) at t = c / kEdgeRes
        const float tf = static_cast<float>(c) / static_cast<float>(kEdgeRes);
        for (int d = 0; d < dim; ++d) o[d] = a[d] + tf * (ck[d] - a[d]);
    } else {  // edge (candidate, b) at t = (c - kEdgeRes + 1) / kEdgeRes
        const float tf = static_cast<float>(c - kEdgeRes + 1) / static_cast<float>(kEdgeRes);
        for (int d = 0; d < dim; ++d) o[d] = ck[d] + tf * (b[d] - ck[d]);
    }
}

// Moves waypoint k of parity par to its candidate where none of its kMovePoints points collides.
__global__ void refine_accept_kernel(float* paths, const float* cand, const uint8_t* hit, int P, int L, int dim,
                                     int par, int slots) {
    const long long t = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x;
    if (t >= static_cast<long long>(P) * slots) return;
    const int j = static_cast<int>(t % slots), p = static_cast<int>(t / slots);
    const int k = 1 + par + 2 * j;
    if (k > L - 2) return;
    const uint8_t* h = hit + t * kMovePoints;
    for (int c = 0; c < kMovePoints; ++c)
        if (h[c]) return;
    const size_t off = (static_cast<size_t>(p) * L + k) * dim;
    for (int d = 0; d < dim; ++d) paths[off + d] = cand[off + d];
}

// Total cost of all paths: sum of the per-waypoint collision costs plus the smoothness term.
// One work-group, fixed summation order (deterministic).
__global__ __launch_bounds__(1024) void refine_total_kernel(const float* paths, const float* cost, int P, int L,
                                                            int dim, float smooth, float* out) {
    __shared__ float red[1024];
    const int tid = threadIdx.x;
    float s = 0.0f;
    const long long n = static_cast<long long>(P) * L;
    for (long long i = tid; i < n; i += blockDim.x) {
        s += cost[i];
        const int k = static_cast<int>(i % L);
        if (k + 1 < L) {
            float e = 0.0f;
            for (int d = 0; d < dim; ++d) {
                const float df = paths[(i + 1) * dim + d] - paths[i * dim + d];
                e = __builtin_fmaf(df, df, e);
            }
            s = __builtin_fmaf(smooth, e, s);
        }
    }
    red[tid] = s;
    __syncthreads();
    for (int o = blockDim.x / 2; o > 0; o >>= 1) {
        if (tid < o) red[tid] += red[tid + o];
        __syncthreads();
    }
    if (tid == 0) *out = red[0];
}

struct CostPlan {
    CostArgs a;
    size_t lds;
    int grid, nt;
};

int plan_cost(const prrtc_robot* robot, const float* world_spheres, int nws, int n, int dim, float beta,
              CostPlan& cp) {
    int rc = check_robot(robot);
    if (rc) return rc;
    if (n < 0 || dim < 1 || dim > PRRTC_MAX_DIM) return fail_with(PRRTC_E_INVALID, "bad n or dim");
    if (nws < 0 || (nws > 0 && !world_spheres)) return fail_with(PRRTC_E_INVALID, "bad world spheres");
    if (!(beta > 0.0f)) return fail_with(PRRTC_E_INVALID, "beta must be > 0");
    if (robot->n_robot_spheres > 0 && (!robot->sphere_link_idx || !robot->sphere_local || !robot->sphere_radius))
        return fail_with(PRRTC_E_INVALID, "robot sphere arrays are null");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_with(PRRTC_E_HIP, "hipGetDevice failed");
    const DevInfo& di = dev_info(dev);
    prrtc_world w;
    std::memset(&w, 0, sizeof(w));
    w.world_spheres = world_spheres;
    w.n_world_spheres = nws;
    CostArgs& a = cp.a;
    std::memset(&a, 0, sizeof(a));
    const int nj = robot->n_joints, nrs = robot->n_robot_spheres;
    // Configurations per wave batch: the largest power of two <= 16 that still gives every SIMD
    // 4 batches, at least 4 (FK runs on nb lanes, so smaller batches repeat it for too few
    // configurations).  32768 configurations of a 7-DOF arm: nb 8 takes 0.159 ms on MI355X,
    // nb 4 0.204 ms (FK repeated for fewer configurations), nb 16 0.231 ms (too few waves).
    int nb = kMaxCostBatch;
    while (nb > 4 && (static_cast<long long>(n) + nb - 1) / nb < 16ll * di.cus) nb /= 2;
    int nt = 256;
    for (;;) {
        a.ml = model_layout(nj, nrs, 0, nws, 0, 0, 0, nt / kWave, kFBP);
        a.nb = nb;
        a.log_nb = __builtin_ctz(static_cast<unsigned>(nb));
        a.jw_stride = 3 * nb * nj + (nb * nj + 3) / 4;
        a.off_anc = align16(a.ml.end);
        a.off_jw = align16(a.off_anc + sizeof(unsigned long long) * nj);
        a.off_scan = align16(a.off_jw + sizeof(float4) * a.jw_stride * (nt / kWave));
        cp.lds = align16(a.off_scan + sizeof(int) * 20);
        if (cp.lds <= static_cast<size_t>(di.max_lds)) break;
        if (nb > 4) nb /= 2;
        else if (nt > kWave) nt -= kWave;
        else break;
    }
    if (cp.lds > static_cast<size_t>(di.max_lds))
        return fail_with(PRRTC_E_LDS, "robot + world spheres exceed the work-group LDS");
    a.robot = robot_dev(*robot);
    a.world = world_dev(w, false);
    a.n = n;
    a.dim = dim;
    a.beta = beta;
    cp.nt = nt;
    const int batches = (n + nb - 1) / nb;
    cp.grid = std::max(1, std::min((batches + nt / kWave - 1) / (nt / kWave), 8 * di.cus));
    return PRRTC_OK;
}

hipError_t launch_cost(const CostPlan& cp, hipStream_t s) {
    const void* fn = cp.a.out_grad ? reinterpret_cast<const void*>(&sphere_cost_kernel<true>)
                                   : reinterpret_cast<const void*>(&sphere_cost_kernel<false>);
    hipError_t e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(cp.lds));
    if (e != hipSuccess) return e;
    if (cp.a.out_grad) hipLaunchKernelGGL(sphere_cost_kernel<true>, dim3(cp.grid), dim3(cp.nt), cp.lds, s, cp.a);
    else hipLaunchKernelGGL(sphere_cost_kernel<false>, dim3(cp.grid), dim3(cp.nt), cp.lds, s, cp.a);
    return hipGetLastError();
}

struct RefineWs {
    size_t cand, grad, cost, pts, hit, total;
    long long n_pts;
};

RefineWs refine_ws(int P, int L, int dim) {
    RefineWs w;
    const long long n = static_cast<long long>(P) * L;
    w.n_pts = static_cast<long long>(P) * std::max(parity_slots(L, 0), 0) * kMovePoints;
    size_t o = 0;
    w.cand = o; o = align16(o + sizeof(float) * n * dim);
    w.grad = o; o = align16(o + sizeof(float) * n * dim);
    w.cost = o; o = align16(o + sizeof(float) * n);
    w.pts = o; o = align16(o + sizeof(float) * static_cast<size_t>(w.n_pts) * dim);
    w.hit = o; o = align16(o + static_cast<size_t>(w.n_pts));
    w.total = o + 256;
    return w;
}

}  // namespace prrtc

using namespace prrtc;

extern "C" {

int prrtc_sphere_cost(prrtc_stream_t stream, const float* configs, int32_t n, int32_t dim, const prrtc_robot* robot,
                      const float* world_spheres, int32_t n_world_spheres, float beta, float* out_cost,
                      float* out_grad, uint64_t* work_flops) {
    CostPlan cp;
    int rc = plan_cost(robot, world_spheres, n_world_spheres, n, dim, beta, cp);
    if (rc) return rc;
    if (n == 0) return PRRTC_OK;
    if (!configs || !out_cost) return fail_with(PRRTC_E_INVALID, "configs / out_cost is null");
    cp.a.configs = configs;
    cp.a.out_cost = out_cost;
    cp.a.out_grad = out_grad;
    cp.a.flops = reinterpret_cast<unsigned long long*>(work_flops);
    const hipError_t e = launch_cost(cp, reinterpret_cast<hipStream_t>(stream));
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    return PRRTC_OK;
}

size_t prrtc_refine_workspace_bytes(int32_t n_paths, int32_t n_waypoints, int32_t dim) {
    if (n_paths < 0 || n_waypoints < 2 || dim < 1) return 0;
    return refine_ws(n_paths, n_waypoints, dim).total;
}

int prrtc_refine_paths(prrtc_stream_t stream, float* paths, int32_t n_paths, int32_t n_waypoints, int32_t dim,
                       const float* min_vals, const float* max_vals, const prrtc_robot* robot,
                       const prrtc_world* world, int32_t collision_mode, float collision_margin, float beta,
                       float smooth_weight, float step, int32_t iterations, float* out_cost, void* workspace,
                       size_t workspace_bytes, uint64_t* work_flops) {
    if (n_paths < 0 || n_waypoints < 2 || iterations < 0)
        return fail_with(PRRTC_E_INVALID, "need n_paths >= 0, n_waypoints >= 2, iterations >= 0");
    if (!world) return fail_with(PRRTC_E_INVALID, "world is null");
    if (!(step >= 0.0f) || !(smooth_weight >= 0.0f)) return fail_with(PRRTC_E_INVALID, "step and weight must be >= 0");
    if ((min_vals == nullptr) != (max_vals == nullptr)) return fail_with(PRRTC_E_INVALID, "give both limits or none");
    int rc = check_world(world);
    if (rc) return rc;
    const long long n = static_cast<long long>(n_paths) * n_waypoints;
    if (n > (1ll << 30)) return fail_with(PRRTC_E_INVALID, "too many waypoints");
    CostPlan cp;
    rc = plan_cost(robot, world->world_spheres, world->n_world_spheres, static_cast<int>(n), dim, beta, cp);
    if (rc) return rc;
    if (n == 0) return PRRTC_OK;
    if (!paths) return fail_with(PRRTC_E_INVALID, "paths is null");
    const RefineWs wl = refine_ws(n_paths, n_waypoints, dim);
    if (!workspace || workspace_bytes < wl.total) return fail_with(PRRTC_E_WORKSPACE, "workspace too small");
    if (wl.n_pts > (1ll << 31) - 1) return fail_with(PRRTC_E_INVALID, "too many move checks per step");
    char* ws = static_cast<char*>(workspace);
    float* cand = reinterpret_cast<float*>(ws + wl.cand);
    float* grad = reinterpret_cast<float*>(ws + wl.grad);
    float* cost = reinterpret_cast<float*>(ws + wl.cost);
    float* pts = reinterpret_cast<float*>(ws + wl.pts);
    uint8_t* hit = reinterpret_cast<uint8_t*>(ws + wl.hit);
    const int slots = parity_slots(n_waypoints, 0);
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    cp.a.configs = paths;
    cp.a.out_cost = cost;
    cp.a.flops = reinterpret_cast<unsigned long long*>(work_flops);
    const long long elems = n * dim;
    const int eblocks = static_cast<int>((elems + 255) / 256);
    for (int it = 0; it <= iterations; ++it) {
        const bool last = it == iterations;
        cp.a.out_grad = last ? nullptr : grad;
        hipError_t e = launch_cost(cp, s);
        if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
        if (out_cost) {
            hipLaunchKernelGGL(refine_total_kernel, dim3(1), dim3(1024), 0, s, paths, cost, n_paths, n_waypoints, dim,
                               smooth_weight, out_cost + it);
        }
        if (last) break;
        hipLaunchKernelGGL(refine_update_kernel, dim3(eblocks), dim3(256), 0, s, paths, grad, n_paths, n_waypoints,
                           dim, step, smooth_weight, min_vals, max_vals, cand);
        if (slots <= 0) continue;  // two waypoints: nothing moves
        for (int par = 0; par < 2; ++par) {
            const int pblocks = static_cast<int>((wl.n_pts + 255) / 256);
            hipLaunchKernelGGL(refine_move_points_kernel, dim3(pblocks), dim3(256), 0, s, paths, cand, n_paths,
                               n_waypoints, dim, par, slots, pts);
            rc = prrtc_check_configs(stream, pts, static_cast<int32_t>(wl.n_pts), dim, robot, world, collision_mode,
                                     collision_margin, hit, nullptr, nullptr, work_flops);
            if (rc) return rc;
            const long long moves = static_cast<long long>(n_paths) * slots;
            hipLaunchKernelGGL(refine_accept_kernel, dim3(static_cast<int>((moves + 255) / 256)), dim3(256), 0, s,
                               paths, cand, hit, n_paths, n_waypoints, dim, par, slots);
            e = hipGetLastError();
            if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
        }
    }
    return PRRTC_OK;
}

}  // extern "C"

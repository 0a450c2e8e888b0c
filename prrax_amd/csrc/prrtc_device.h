// prrtc_device.h -- device-side arithmetic for the MI355X pRRTC planner (gfx950, wave64).
//
// Forward kinematics (pyroffi fk_single, called at cuda-rrtc/CUDA/prrtc_planner.cu:509-522),
// the SE3 point transform (apply_se3_point, prrtc_planner.cu:246) and the primitive collision
// tests (sphere_*_dist, prrtc_planner.cu:250-304) are written here from the mathematics.
//
// Floating-point contract: this file is compiled with -ffp-contract=off and every fused
// multiply-add is an explicit __builtin_fmaf, so each value is one fixed sequence of IEEE
// operations.  sin/cos come from the polynomial below rather than the device library.  The
// CPU oracle restates the same sequence independently; planner trees are compared bit for bit.
#pragma once
#include <hip/hip_runtime.h>
#include <float.h>
#include <stdint.h>

// Bounds checks and observed queue depths for the CPU build of these functions (a host stand-in
// for the HIP header defines them there); on the device they expand to nothing.
#ifndef PRRTC_SIMT_ASSERT
#define PRRTC_SIMT_ASSERT(cond) ((void)0)
#endif
#ifndef PRRTC_SIMT_OBSERVE
#define PRRTC_SIMT_OBSERVE(slot, value) ((void)0)
#endif

namespace prrtc {

constexpr int kMaxJoints = 64;
constexpr int kMaxDim = 16;
constexpr int kWave = 64;

enum JointType : int { kFixed = 0, kRot = 1, kPrism = 2, kScrew = 3 };

// ---------------------------------------------------------------------------------- sin/cos
// Cody-Waite reduction by pi/2 (three-part split) and Taylor polynomials on [-pi/4, pi/4];
// coefficients are (-1)^k/n! rounded to float.
__device__ __forceinline__ void sincos_pr(float x, float& s_out, float& c_out) {
    const float k = __builtin_rintf(x * 0x1.45f306p-1f);
    float r = __builtin_fmaf(-k, 0x1.921fb6p+0f, x);
    r = __builtin_fmaf(-k, -0x1.777a5cp-25f, r);
    r = __builtin_fmaf(-k, -0x1.ee59dap-50f, r);
    const float r2 = r * r;
    float ps = -0x1.ae6456p-26f;
    ps = __builtin_fmaf(ps, r2, 0x1.71de3ap-19f);
    ps = __builtin_fmaf(ps, r2, -0x1.a01a02p-13f);
    ps = __builtin_fmaf(ps, r2, 0x1.111112p-7f);
    ps = __builtin_fmaf(ps, r2, -0x1.555556p-3f);
    const float sn = __builtin_fmaf(r * r2, ps, r);
    float pc = 0x1.1eed8ep-29f;
    pc = __builtin_fmaf(pc, r2, -0x1.27e4fcp-22f);
    pc = __builtin_fmaf(pc, r2, 0x1.a01a02p-16f);
    pc = __builtin_fmaf(pc, r2, -0x1.6c16c2p-10f);
    pc = __builtin_fmaf(pc, r2, 0x1.555556p-5f);
    const float cs = __builtin_fmaf(r2, __builtin_fmaf(r2, pc, -0.5f), 1.0f);
    const int q = static_cast<int>(k) & 3;
    const float a = (q & 1) ? cs : sn;
    const float b = (q & 1) ? sn : cs;
    s_out = (q & 2) ? -a : a;
    c_out = ((q + 1) & 2) ? -b : b;
}

// ---------------------------------------------------------------------------------- SE3
struct Quat { float w, x, y, z; };
struct Mat3 { float m0, m1, m2, m3, m4, m5, m6, m7, m8; };

__device__ __forceinline__ Quat qmul(const Quat& a, const Quat& b) {
    Quat o;
    float t;
    t = a.w * b.w; t = __builtin_fmaf(-a.x, b.x, t); t = __builtin_fmaf(-a.y, b.y, t); t = __builtin_fmaf(-a.z, b.z, t); o.w = t;
    t = a.w * b.x; t = __builtin_fmaf(a.x, b.w, t); t = __builtin_fmaf(a.y, b.z, t); t = __builtin_fmaf(-a.z, b.y, t); o.x = t;
    t = a.w * b.y; t = __builtin_fmaf(-a.x, b.z, t); t = __builtin_fmaf(a.y, b.w, t); t = __builtin_fmaf(a.z, b.x, t); o.y = t;
    t = a.w * b.z; t = __builtin_fmaf(a.x, b.y, t); t = __builtin_fmaf(-a.y, b.x, t); t = __builtin_fmaf(a.z, b.w, t); o.z = t;
    return o;
}

__device__ __forceinline__ Mat3 quat_mat(const Quat& q) {
    const float x2 = q.x + q.x, y2 = q.y + q.y, z2 = q.z + q.z;
    const float xx = q.x * x2, yy = q.y * y2, zz = q.z * z2;
    const float xy = q.x * y2, xz = q.x * z2, yz = q.y * z2;
    const float wx = q.w * x2, wy = q.w * y2, wz = q.w * z2;
    Mat3 R;
    R.m0 = 1.0f - (yy + zz); R.m1 = xy - wz;          R.m2 = xz + wy;
    R.m3 = xy + wz;          R.m4 = 1.0f - (xx + zz); R.m5 = yz - wx;
    R.m6 = xz - wy;          R.m7 = yz + wx;          R.m8 = 1.0f - (xx + yy);
    return R;
}

__device__ __forceinline__ float3 mat_apply(const Mat3& R, float tx, float ty, float tz, float px, float py, float pz) {
    float3 o;
    o.x = __builtin_fmaf(R.m2, pz, __builtin_fmaf(R.m1, py, __builtin_fmaf(R.m0, px, tx)));
    o.y = __builtin_fmaf(R.m5, pz, __builtin_fmaf(R.m4, py, __builtin_fmaf(R.m3, px, ty)));
    o.z = __builtin_fmaf(R.m8, pz, __builtin_fmaf(R.m7, py, __builtin_fmaf(R.m6, px, tz)));
    return o;
}

// ---------------------------------------------------------------------------------- model
// Joint record staged in LDS once per query (topological order of the context arrays).
struct JointRec {
    int type, parent, act, mimic;
    float mul, off, wn, ax, ay, az, vx, vy, vz;
    float pw, px, py, pz, ptx, pty, ptz;
};

// Classify a twist (v, omega) and derive its unit axis; same decisions as the oracle.
__device__ __forceinline__ void joint_prep(const float* tw, JointRec& j) {
    const float vx = tw[0], vy = tw[1], vz = tw[2], wx = tw[3], wy = tw[4], wz = tw[5];
    const float wn2 = __builtin_fmaf(wz, wz, __builtin_fmaf(wy, wy, wx * wx));
    const float vn2 = __builtin_fmaf(vz, vz, __builtin_fmaf(vy, vy, vx * vx));
    j.vx = vx; j.vy = vy; j.vz = vz;
    j.ax = 0.0f; j.ay = 0.0f; j.az = 0.0f; j.wn = 0.0f;
    if (wn2 == 0.0f && vn2 == 0.0f) {
        j.type = kFixed;
    } else if (wn2 == 0.0f) {
        j.type = kPrism;
    } else {
        const float wn = __builtin_sqrtf(wn2);
        j.wn = wn;
        j.ax = wx / wn; j.ay = wy / wn; j.az = wz / wn;
        j.type = (vn2 == 0.0f) ? kRot : kScrew;
    }
}

// World primitives in LDS, in the form the tests consume (derived values use the same
// operation sequence as the oracle's inline computation).
struct CapsuleL { float p1x, p1y, p1z, ax, ay, az, inv, r; };
// `ortho` = 1 when the axes are orthonormal (within 1e-4): only such boxes may be culled.
struct BoxL { float cx, cy, cz, a0x, a0y, a0z, a1x, a1y, a1z, a2x, a2y, a2z, h0, h1, h2, ortho; };

__device__ __forceinline__ CapsuleL make_capsule(const float* c) {
    CapsuleL o;
    o.p1x = c[0]; o.p1y = c[1]; o.p1z = c[2];
    o.ax = c[3] - c[0]; o.ay = c[4] - c[1]; o.az = c[5] - c[2];
    const float l2 = __builtin_fmaf(o.az, o.az, __builtin_fmaf(o.ay, o.ay, o.ax * o.ax));
    o.inv = (l2 > 0.0f) ? 1.0f / l2 : 0.0f;
    o.r = c[6];
    return o;
}

__device__ __forceinline__ float4 make_halfspace(const float* h) {
    const float off = __builtin_fmaf(h[2], h[5], __builtin_fmaf(h[1], h[4], h[0] * h[3]));
    return make_float4(h[0], h[1], h[2], off);
}

// Robot + world as seen by one work-group (all pointers into LDS).
struct Model {
    const JointRec* joints;
    const float4* spheres;   // robot spheres: local xyz, radius
    const int* sphere_js;    // sorted-joint index of each robot sphere, -1 if skipped
    const unsigned* pairs;   // valid self pairs, i | j << 16 (robot spheres < 65536)
    const float4* wsph;
    const CapsuleL* wcap;
    const BoxL* wbox;
    const float4* whs;
    int n_joints, n_spheres, n_pairs, n_ws, n_wc, n_wb, n_wh;
    int enabled;
    int chain;      // 1: a serial chain of revolute joints, joint j actuated, parent j - 1 (fk_lane_chain)
    int mode;       // 0 binary (also the coarse modes: same verdict), 1 margin
    float margin;
    // bounding sphere of each frame's robot spheres (groups in frame order)
    const float4* gcoarse;   // frame-local centre, radius = max(|p - c| + |r|)
    const int* gframe;       // sorted-joint index of the group's frame
    const int* goff;         // members of group g: gmem[goff[g] .. goff[g + 1])
    const int* gmem;
    const int* fgroup;       // group of each sorted frame, -1 for a frame without spheres
    int n_groups;
    char* cull_base;         // per-wave cull areas (see wave_culled)
    int cull_stride;
};

// ---------------------------------------------------------------------------------- tests
// Binary tests: the reference's "dist <= 0" (prrtc_planner.cu:250-304) without square roots.  A
// sphere or capsule whose radius sum rr = r + R is >= 0 is hit iff d^2 <= rr^2, never when
// rr < 0 (sqrt(d^2) - rr > 0): both at once as d^2 <= rr |rr| (|.| is a free operand modifier);
// a box with sr >= 0 iff the squared outside distance is <= sr^2 (0 when the centre is inside),
// with sr < 0 iff the centre lies deeper than |sr| inside (max q <= sr); a halfspace iff
// n.(c - p) <= sr.
__device__ __forceinline__ bool hit_sphere(float sx, float sy, float sz, float sr, const float4 o) {
    const float dx = sx - o.x, dy = sy - o.y, dz = sz - o.z;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float rr = sr + o.w;
    return d2 <= rr * fabsf(rr);
}

__device__ __forceinline__ float capsule_d2(float sx, float sy, float sz, const CapsuleL& c) {
    const float wx = sx - c.p1x, wy = sy - c.p1y, wz = sz - c.p1z;
    float t = __builtin_fmaf(wz, c.az, __builtin_fmaf(wy, c.ay, wx * c.ax)) * c.inv;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float dx = __builtin_fmaf(-t, c.ax, wx), dy = __builtin_fmaf(-t, c.ay, wy), dz = __builtin_fmaf(-t, c.az, wz);
    return __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
}

__device__ __forceinline__ bool hit_capsule(float sx, float sy, float sz, float sr, const CapsuleL& c) {
    const float rr = sr + c.r;
    return capsule_d2(sx, sy, sz, c) <= rr * fabsf(rr);
}

__device__ __forceinline__ void box_q(float sx, float sy, float sz, const BoxL& b, float& q0, float& q1, float& q2) {
    const float wx = sx - b.cx, wy = sy - b.cy, wz = sz - b.cz;
    const float l0 = __builtin_fmaf(wz, b.a0z, __builtin_fmaf(wy, b.a0y, wx * b.a0x));
    const float l1 = __builtin_fmaf(wz, b.a1z, __builtin_fmaf(wy, b.a1y, wx * b.a1x));
    const float l2 = __builtin_fmaf(wz, b.a2z, __builtin_fmaf(wy, b.a2y, wx * b.a2x));
    q0 = fabsf(l0) - b.h0;
    q1 = fabsf(l1) - b.h1;
    q2 = fabsf(l2) - b.h2;
}

// Squared outside distance of a box; qmax = the largest q (<= 0 inside).
__device__ __forceinline__ float box_d2(float sx, float sy, float sz, const BoxL& b, float& qmax) {
    float q0, q1, q2;
    box_q(sx, sy, sz, b, q0, q1, q2);
    const float m0 = fmaxf(q0, 0.0f), m1 = fmaxf(q1, 0.0f), m2 = fmaxf(q2, 0.0f);
    qmax = fmaxf(q0, fmaxf(q1, q2));
    return __builtin_fmaf(m2, m2, __builtin_fmaf(m1, m1, m0 * m0));
}

__device__ __forceinline__ bool hit_box(float sx, float sy, float sz, float sr, const BoxL& b) {
    float qmax;
    const float d2 = box_d2(sx, sy, sz, b, qmax);
    return (sr >= 0.0f) ? (d2 <= sr * sr) : (qmax <= sr);
}

__device__ __forceinline__ bool hit_halfspace(float sx, float sy, float sz, float sr, const float4 h) {
    const float s = __builtin_fmaf(h.z, sz, __builtin_fmaf(h.y, sy, h.x * sx)) - h.w;
    return s <= sr;
}

// Margin tests: the reference's "dist < margin" (prrtc_planner.cu:203-206, 314-387), also without
// square roots.  With L = r + R + margin, a sphere or capsule is near iff L > 0 and d^2 < L^2
// (d^2 < L |L|); a box (L = sr + margin) iff L > 0 and outside^2 < L^2 when the centre is outside
// (outside^2 > 0), else iff min(max q, 0) < L; a halfspace iff n.(c - p) - sr < margin.
__device__ __forceinline__ bool near_sphere(float sx, float sy, float sz, float sr, const float4 o, float margin) {
    const float dx = sx - o.x, dy = sy - o.y, dz = sz - o.z;
    const float d2 = __builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx));
    const float L = (sr + o.w) + margin;
    return d2 < L * fabsf(L);
}

__device__ __forceinline__ bool near_capsule(float sx, float sy, float sz, float sr, const CapsuleL& c, float margin) {
    const float L = (sr + c.r) + margin;
    return capsule_d2(sx, sy, sz, c) < L * fabsf(L);
}

__device__ __forceinline__ bool near_box(float sx, float sy, float sz, float sr, const BoxL& b, float margin) {
    float qmax;
    const float d2 = box_d2(sx, sy, sz, b, qmax);
    const float L = sr + margin;
    return (d2 > 0.0f) ? (L > 0.0f && d2 < L * L) : (fminf(qmax, 0.0f) < L);
}

__device__ __forceinline__ bool near_halfspace(float sx, float sy, float sz, float sr, const float4 h, float margin) {
    return (__builtin_fmaf(h.z, sz, __builtin_fmaf(h.y, sy, h.x * sx)) - h.w) - sr < margin;
}

// Signed distances (the minimum-distance value of the flat sweep).  The square root is the
// hardware's (v_sqrt_f32, within 1 ulp), not the correctly rounded one: no verdict depends on
// these values.
__device__ __forceinline__ float fast_sqrt(float x) { return __builtin_amdgcn_sqrtf(x); }

__device__ __forceinline__ float dist_sphere(float sx, float sy, float sz, float sr, const float4 o) {
    const float dx = sx - o.x, dy = sy - o.y, dz = sz - o.z;
    return fast_sqrt(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx))) - (sr + o.w);
}

// Signed distances of a point (a robot sphere's centre) to the obstacle surfaces; the flat sweep
// subtracts the robot sphere's radius once from their minimum.
__device__ __forceinline__ float gap_sphere(float sx, float sy, float sz, const float4 o) {
    const float dx = sx - o.x, dy = sy - o.y, dz = sz - o.z;
    return fast_sqrt(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx))) - o.w;
}

__device__ __forceinline__ float gap_box(float sx, float sy, float sz, const BoxL& b) {
    float qmax;
    const float outside = fast_sqrt(box_d2(sx, sy, sz, b, qmax));
    return outside + fminf(qmax, 0.0f);
}

// ---------------------------------------------------------------------------------- FK
// Per-lane scratch lives in LDS in lane-minor layout: element c of lane l at [c * nt + l], so
// a wave's 64 lanes touch 64 consecutive words (no bank conflicts).  q holds the config
// (kMaxDim slots), T holds 7 floats per joint (w, x, y, z, tx, ty, tz).
struct LaneScratch {
    float* q;
    float* T;
    int nt;
};

// fk_lane for a serial chain of revolute joints (Model::chain): joint 0 has no parent, joint j's
// parent is j - 1, every joint is actuated without mimic.  The arithmetic is fk_lane's for such
// joints, operation for operation, without its per-joint branches on type, parent and mimic.
__device__ __forceinline__ void fk_lane_chain(const Model& M, const LaneScratch& L) {
    const int nt = L.nt, nj = M.n_joints;
    Quat pq = {1.0f, 0.0f, 0.0f, 0.0f};
    float ptx = 0.0f, pty = 0.0f, ptz = 0.0f;
    JointRec Jn = M.joints[0];
    for (int j = 0; j < nj; ++j) {
        const JointRec J = Jn;
        if (j + 1 < nj) Jn = M.joints[j + 1];
        const float theta = L.q[J.act * nt];
        const Quat P = {J.pw, J.px, J.py, J.pz};
        Quat aq;
        float atx, aty, atz;
        if (j == 0) {
            aq = P; atx = J.ptx; aty = J.pty; atz = J.ptz;
        } else {
            aq = qmul(pq, P);
            const float3 t = mat_apply(quat_mat(pq), ptx, pty, ptz, J.ptx, J.pty, J.ptz);
            atx = t.x; aty = t.y; atz = t.z;
        }
        float s, c;
        sincos_pr(0.5f * (J.wn * theta), s, c);
        const Quat E = {c, s * J.ax, s * J.ay, s * J.az};
        const Quat oq = qmul(aq, E);
        float* To = L.T + (j * 7) * nt;
        To[0] = oq.w; To[nt] = oq.x; To[2 * nt] = oq.y; To[3 * nt] = oq.z;
        To[4 * nt] = atx; To[5 * nt] = aty; To[6 * nt] = atz;
        pq = oq; ptx = atx; pty = aty; ptz = atz;
    }
}

__device__ __forceinline__ void fk_lane(const Model& M, const LaneScratch& L) {
    if (M.chain) {
        fk_lane_chain(M, L);
        return;
    }
    const int nt = L.nt;
    Quat pq = {1.0f, 0.0f, 0.0f, 0.0f};
    float ptx = 0.0f, pty = 0.0f, ptz = 0.0f;
    // The next joint's record is loaded before this joint's frame is stored, so its reads are
    // not ordered behind the stores and overlap this joint's arithmetic.
    JointRec Jn{};
    if (M.n_joints > 0) Jn = M.joints[0];
    for (int j = 0; j < M.n_joints; ++j) {
        const JointRec J = Jn;
        if (j + 1 < M.n_joints) Jn = M.joints[j + 1];
        float theta;
        if (J.mimic >= 0) theta = __builtin_fmaf(L.q[J.mimic * nt], J.mul, J.off);
        else if (J.act >= 0) theta = L.q[J.act * nt];
        else theta = 0.0f;
        Quat aq;
        float atx, aty, atz;
        const Quat P = {J.pw, J.px, J.py, J.pz};
        if (J.parent < 0) {
            aq = P; atx = J.ptx; aty = J.pty; atz = J.ptz;
        } else {
            Quat tq;
            float ttx, tty, ttz;
            if (J.parent == j - 1) {
                tq = pq; ttx = ptx; tty = pty; ttz = ptz;
            } else {
                const float* Tp = L.T + (J.parent * 7) * nt;
                tq.w = Tp[0]; tq.x = Tp[nt]; tq.y = Tp[2 * nt]; tq.z = Tp[3 * nt];
                ttx = Tp[4 * nt]; tty = Tp[5 * nt]; ttz = Tp[6 * nt];
            }
            aq = qmul(tq, P);
            const float3 t = mat_apply(quat_mat(tq), ttx, tty, ttz, J.ptx, J.pty, J.ptz);
            atx = t.x; aty = t.y; atz = t.z;
        }
        Quat oq = aq;
        float otx = atx, oty = aty, otz = atz;
        if (J.type == kRot) {
            float s, c;
            sincos_pr(0.5f * (J.wn * theta), s, c);
            const Quat E = {c, s * J.ax, s * J.ay, s * J.az};
            oq = qmul(aq, E);
        } else if (J.type == kPrism) {
            const float3 t = mat_apply(quat_mat(aq), atx, aty, atz, J.vx * theta, J.vy * theta, J.vz * theta);
            otx = t.x; oty = t.y; otz = t.z;
        } else if (J.type == kScrew) {
            const float th = J.wn * theta;
            float sh, ch, s, c;
            sincos_pr(0.5f * th, sh, ch);
            sincos_pr(th, s, c);
            float ca, cb;
            if (fabsf(th) < 1e-3f) {
                ca = 0.5f * th;
                cb = (th * th) * 0x1.555556p-3f;
            } else {
                ca = (1.0f - c) / th;
                cb = (th - s) / th;
            }
            const float rx = J.vx * theta, ry = J.vy * theta, rz = J.vz * theta;
            const float cx = __builtin_fmaf(J.ay, rz, -(J.az * ry));
            const float cy = __builtin_fmaf(J.az, rx, -(J.ax * rz));
            const float cz = __builtin_fmaf(J.ax, ry, -(J.ay * rx));
            const float dx = __builtin_fmaf(J.ay, cz, -(J.az * cy));
            const float dy = __builtin_fmaf(J.az, cx, -(J.ax * cz));
            const float dz = __builtin_fmaf(J.ax, cy, -(J.ay * cx));
            const float ex = __builtin_fmaf(cb, dx, __builtin_fmaf(ca, cx, rx));
            const float ey = __builtin_fmaf(cb, dy, __builtin_fmaf(ca, cy, ry));
            const float ez = __builtin_fmaf(cb, dz, __builtin_fmaf(ca, cz, rz));
            const Quat E = {ch, sh * J.ax, sh * J.ay, sh * J.az};
            oq = qmul(aq, E);
            const float3 t = mat_apply(quat_mat(aq), atx, aty, atz, ex, ey, ez);
            otx = t.x; oty = t.y; otz = t.z;
        }
        float* To = L.T + (j * 7) * nt;
        To[0] = oq.w; To[nt] = oq.x; To[2 * nt] = oq.y; To[3 * nt] = oq.z;
        To[4 * nt] = otx; To[5 * nt] = oty; To[6 * nt] = otz;
        pq = oq; ptx = otx; pty = oty; ptz = otz;
    }
}

// ---------------------------------------------------------------------------------- wave checks
// One wave checks one configuration: lane l takes robot spheres l, l+64, ...; every lane walks the
// same obstacle list (LDS broadcast), and a verdict leaves as soon as any lane reports a contact
// (ballot), so the exit is wave-uniform.  The verdicts are an OR of pair tests and the distance
// a min over pairs, so the order of the tests does not change the result.
//
// The configuration's joint frames come from a per-wave FK area T (element c of slot k at
// T[k * ts]); fk_lane fills it for a batch of configurations, one per lane.
//
// flops accumulates algorithmic flops of the tests performed (FMA = 2): 33 per sphere transform,
// 10 / 27 / 36 / 9 per sphere-sphere / capsule / box / halfspace test, 10 per self pair.
constexpr int kFB = 16;        // configurations per FK batch
constexpr int kFBP = kFB + 1;  // padded stride of the per-wave FK area (spreads LDS banks)
constexpr unsigned kFlopsXf = 33, kFlopsSS = 10, kFlopsSC = 27, kFlopsSB = 36, kFlopsSH = 9, kFlopsPair = 10;
constexpr unsigned kFlopsJoint = 128;

__device__ __forceinline__ void frame_from(const float* T, int ts, int js, Mat3& R, float& tx, float& ty, float& tz) {
    const float* F = T + (js * 7) * ts;
    const Quat q = {F[0], F[ts], F[2 * ts], F[3 * ts]};
    R = quat_mat(q);
    tx = F[4 * ts]; ty = F[5 * ts]; tz = F[6 * ts];
}

__device__ __forceinline__ bool wave_any(bool p) { return __ballot(p) != 0ull; }

// Orders this wave's LDS writes before its later LDS reads from other lanes.
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "wavefront");
    __builtin_amdgcn_wave_barrier();
}

// Minimum signed distance of one configuration over every robot-obstacle sphere pair and self
// pair: the flat full sweep (no culling, no early exit).  Lane l takes robot spheres l, l + 64,
// ...; per sphere it keeps the minimum of its centre's distances to the obstacle surfaces and
// subtracts its radius once (the value differs from a per-pair "distance - radii" only by the
// rounding of that subtraction).  World primitives go in unrolled blocks (8 spheres, 4 capsules,
// 4 boxes) with no per-element guard, so each block's LDS broadcasts are issued together, and
// the minimum of a block is a tree of three-operand minima.
__device__ __forceinline__ float wave_min_distance(const Model& M, const float* T, int ts, unsigned long long& flops) {
    const int lane = threadIdx.x & (kWave - 1);
    const int nws = M.n_ws, nwc = M.n_wc, nwb = M.n_wb, nwh = M.n_wh;
    const unsigned long long per_sphere = static_cast<unsigned long long>(nws) * kFlopsSS +
                                          static_cast<unsigned long long>(nwc) * kFlopsSC +
                                          static_cast<unsigned long long>(nwb) * kFlopsSB +
                                          static_cast<unsigned long long>(nwh) * kFlopsSH;
    float best = FLT_MAX;
    for (int s0 = 0; s0 < M.n_spheres; s0 += kWave) {
        const int s = s0 + lane;
        int js = -1;
        float4 sp = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
        if (s < M.n_spheres) { js = M.sphere_js[s]; sp = M.spheres[s]; }
        const bool act = js >= 0;
        const unsigned na = __popcll(__ballot(act));
        if (na == 0) continue;
        float px = 0.0f, py = 0.0f, pz = 0.0f;
        if (act) {
            Mat3 R;
            float tx, ty, tz;
            frame_from(T, ts, js, R, tx, ty, tz);
            const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
            px = p.x; py = p.y; pz = p.z;
        }
        flops += na * (kFlopsXf + per_sphere);
        float cb = FLT_MAX;
        int i = 0;
        for (; i + 8 <= nws; i += 8) {
            float g[8];
#pragma unroll
            for (int k = 0; k < 8; ++k) g[k] = gap_sphere(px, py, pz, M.wsph[i + k]);
            cb = fminf(cb, fminf(fminf(fminf(g[0], g[1]), fminf(g[2], g[3])), fminf(fminf(g[4], g[5]), fminf(g[6], g[7]))));
        }
        for (; i < nws; ++i) cb = fminf(cb, gap_sphere(px, py, pz, M.wsph[i]));
        i = 0;
        for (; i + 4 <= nwc; i += 4) {
            float g[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) g[k] = fast_sqrt(capsule_d2(px, py, pz, M.wcap[i + k])) - M.wcap[i + k].r;
            cb = fminf(cb, fminf(fminf(g[0], g[1]), fminf(g[2], g[3])));
        }
        for (; i < nwc; ++i) cb = fminf(cb, fast_sqrt(capsule_d2(px, py, pz, M.wcap[i])) - M.wcap[i].r);
        i = 0;
        for (; i + 4 <= nwb; i += 4) {
            float g[4];
#pragma unroll
            for (int k = 0; k < 4; ++k) g[k] = gap_box(px, py, pz, M.wbox[i + k]);
            cb = fminf(cb, fminf(fminf(g[0], g[1]), fminf(g[2], g[3])));
        }
        for (; i < nwb; ++i) cb = fminf(cb, gap_box(px, py, pz, M.wbox[i]));
        for (int k = 0; k < nwh; ++k) {
            const float4 h = M.whs[k];
            cb = fminf(cb, __builtin_fmaf(h.z, pz, __builtin_fmaf(h.y, py, h.x * px)) - h.w);
        }
        if (act) best = fminf(best, cb - sp.w);
    }
    for (int p0 = 0; p0 < M.n_pairs; p0 += kWave) {
        const int k = p0 + lane;
        if (k < M.n_pairs) {
            const int pi = static_cast<int>(M.pairs[k] & 0xffffu), pj = static_cast<int>(M.pairs[k] >> 16);
            const float4 a = M.spheres[pi];
            const float4 b = M.spheres[pj];
            Mat3 Ri, Rj;
            float ix, iy, iz, jx, jy, jz;
            frame_from(T, ts, M.sphere_js[pi], Ri, ix, iy, iz);
            frame_from(T, ts, M.sphere_js[pj], Rj, jx, jy, jz);
            const float3 pa = mat_apply(Ri, ix, iy, iz, a.x, a.y, a.z);
            const float3 pb = mat_apply(Rj, jx, jy, jz, b.x, b.y, b.z);
            best = fminf(best, dist_sphere(pa.x, pa.y, pa.z, a.w, make_float4(pb.x, pb.y, pb.z, b.w)));
        }
        flops += static_cast<unsigned long long>(min(kWave, M.n_pairs - p0)) * kFlopsPair;
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) best = fminf(best, __shfl_xor(best, o, kWave));
    return best;
}

// ---------------------------------------------------------------------------------- culled checks
// The verdict of testing every robot sphere against every obstacle and every self pair (the
// reference's exhaustive loop), with less work: each frame's robot spheres are enclosed in a
// bounding sphere (centre c, radius R = max |p - c| + |r| over its spheres), and a fine test of
// obstacle o against the spheres of group g runs only if the bounding sphere does not prove them
// all clear.  Proof: every fine test is a 1-Lipschitz distance of the sphere centre (point,
// segment, orthonormal box, unit-normal halfspace), so for every member p,
//   dist(p, o) >= dist(c, o) - |p - c| >= dist(c, o) - R + |r|,
// and a coarse distance above R + |r_o| + cut + kCullPad rules out the squared-form hit
// (d <= |r + r_o|) and, in margin mode (cut = margin), any distance below the margin.  kCullPad
// (1 mm) covers the FP32 rounding of both transforms; boxes whose axes are not orthonormal and
// halfspaces whose normal is not unit are never culled.  Every fine test that does run is the
// exhaustive loop's own test on the same FP32 values, so the verdict is identical bit for bit.
constexpr float kCullPad = 1e-3f;
constexpr int kCullQ = 2 * kWave;     // refinement queue (obstacle * 64 + group) per wave
constexpr int kUnitCap = 2 * kWave;   // candidate pairs of a sub-unit (wave_unit_check)
constexpr int kUnitMax = 16;          // configurations per sub-unit (one FK batch, kFB)
static_assert(kUnitMax <= 16, "wave_unit_check packs the configuration into 4 bits of a queue entry");

// Per-wave cull area: robot spheres, group spheres, refinement queue, and ucap candidate slots
// (kUnitCap where wave_unit_check runs, else 0).
__host__ __device__ inline int cull_stride_bytes(int nrs, int nj, int ucap) {
    return static_cast<int>(sizeof(float4)) * (nrs + nj) + static_cast<int>(sizeof(int)) * (kCullQ + ucap);
}

// Obstacles of the culled check are numbered halfspaces, boxes, capsules, spheres: the few large
// primitives (a floor, walls) come first, where a colliding configuration is found soonest.
// Distances are compared squared (d > L  <=>  d^2 > L^2 for L >= 0): no square roots, and the
// rounding difference is far inside kCullPad.
__device__ __forceinline__ bool beyond(float d2, float L) { return L < 0.0f || d2 > L * L; }

__device__ __forceinline__ bool coarse_clear(const Model& M, int o, const float4 c, float lim) {
    if (o < M.n_wh) {
        const float4 h = M.whs[o];
        const float n2 = __builtin_fmaf(h.z, h.z, __builtin_fmaf(h.y, h.y, h.x * h.x));
        if (fabsf(n2 - 1.0f) > 2e-4f) return false;
        return __builtin_fmaf(h.z, c.z, __builtin_fmaf(h.y, c.y, h.x * c.x)) - h.w > lim;
    }
    o -= M.n_wh;
    if (o < M.n_wb) {
        const BoxL& b = M.wbox[o];
        if (b.ortho == 0.0f) return false;
        float q0, q1, q2;
        box_q(c.x, c.y, c.z, b, q0, q1, q2);
        const float m0 = fmaxf(q0, 0.0f), m1 = fmaxf(q1, 0.0f), m2 = fmaxf(q2, 0.0f);
        return beyond(__builtin_fmaf(m2, m2, __builtin_fmaf(m1, m1, m0 * m0)), lim);
    }
    o -= M.n_wb;
    if (o < M.n_wc) {
        const CapsuleL& k = M.wcap[o];
        return beyond(capsule_d2(c.x, c.y, c.z, k), lim + fabsf(k.r));
    }
    o -= M.n_wc;
    const float4 w = M.wsph[o];
    const float dx = c.x - w.x, dy = c.y - w.y, dz = c.z - w.z;
    return beyond(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), lim + fabsf(w.w));
}

// Fine test of robot sphere q (world x, y, z, r) against obstacle o: hit (binary) or distance
// below the margin (MARGIN), as the exhaustive loop computes it.  The checks below take the
// mode at run time (M.mode == 1: margin; it is wave-uniform) and branch to one of the two
// instances at each fine-test site, so each check exists once per kernel.
template <bool MARGIN>
__device__ __forceinline__ bool fine_test(const Model& M, int o, const float4 q, unsigned& fl) {
    if (o < M.n_wh) {
        fl += kFlopsSH;
        return MARGIN ? near_halfspace(q.x, q.y, q.z, q.w, M.whs[o], M.margin)
                      : hit_halfspace(q.x, q.y, q.z, q.w, M.whs[o]);
    }
    o -= M.n_wh;
    if (o < M.n_wb) {
        fl += kFlopsSB;
        return MARGIN ? near_box(q.x, q.y, q.z, q.w, M.wbox[o], M.margin) : hit_box(q.x, q.y, q.z, q.w, M.wbox[o]);
    }
    o -= M.n_wb;
    if (o < M.n_wc) {
        fl += kFlopsSC;
        return MARGIN ? near_capsule(q.x, q.y, q.z, q.w, M.wcap[o], M.margin)
                      : hit_capsule(q.x, q.y, q.z, q.w, M.wcap[o]);
    }
    o -= M.n_wc;
    fl += kFlopsSS;
    return MARGIN ? near_sphere(q.x, q.y, q.z, q.w, M.wsph[o], M.margin) : hit_sphere(q.x, q.y, q.z, q.w, M.wsph[o]);
}

__device__ __forceinline__ bool fine_test_mode(const Model& M, bool mg, int o, const float4 q, unsigned& fl) {
    return mg ? fine_test<true>(M, o, q, fl) : fine_test<false>(M, o, q, fl);
}

__device__ __forceinline__ unsigned wave_sum_u32(unsigned v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o, kWave);
    return v;
}

// Number of coarse pairs p in [p0, p1) whose obstacle index o = p / ng lies in [lo, hi).
__device__ __forceinline__ unsigned long long pairs_in(long long p0, long long p1, int ng, int lo, int hi) {
    const long long a = max(p0, static_cast<long long>(lo) * ng), b = min(p1, static_cast<long long>(hi) * ng);
    return b > a ? static_cast<unsigned long long>(b - a) : 0ull;
}

// Flops of the first p1 pairs of the pair steps (halfspaces, boxes, capsules; ng groups).
__device__ __forceinline__ unsigned long long coarse_flops_of(const Model& M, int p1, int ng) {
    const int nwh = M.n_wh, nwb = M.n_wb, no = M.n_wh + M.n_wb + M.n_wc;
    return pairs_in(0, p1, ng, 0, nwh) * kFlopsSH + pairs_in(0, p1, ng, nwh, nwh + nwb) * kFlopsSB +
           pairs_in(0, p1, ng, nwh + nwb, no) * kFlopsSC;
}

// Self pairs k0 + lane and k0 + kWave + lane against the transformed spheres FP (two chunks per
// wave step: their loads are issued together, one ballot for both).
template <bool MARGIN>
__device__ __forceinline__ bool self_pairs_2(const Model& M, const float4* FP, int k0, int lane) {
    const int k = k0 + lane, k2 = k0 + kWave + lane;
    bool h = false;
    if (k < M.n_pairs) {
        const float4 a = FP[M.pairs[k] & 0xffffu], b = FP[M.pairs[k] >> 16];
        h = MARGIN ? near_sphere(a.x, a.y, a.z, a.w, b, M.margin) : hit_sphere(a.x, a.y, a.z, a.w, b);
    }
    if (k2 < M.n_pairs) {
        const float4 a = FP[M.pairs[k2] & 0xffffu], b = FP[M.pairs[k2] >> 16];
        h |= MARGIN ? near_sphere(a.x, a.y, a.z, a.w, b, M.margin) : hit_sphere(a.x, a.y, a.z, a.w, b);
    }
    return h;
}

__device__ __forceinline__ bool self_pairs_mode(const Model& M, bool mg, const float4* FP, int k0, int lane) {
    return mg ? self_pairs_2<true>(M, FP, k0, lane) : self_pairs_2<false>(M, FP, k0, lane);
}

// Lanes per queued entry of a refinement of qn entries (wave-uniform): the largest power of two
// lpe <= 64 with qn * lpe <= 64, 1 when qn > 32.
__device__ __forceinline__ int lanes_per_entry(int qn) {
    int lpe = 1;
    while (lpe < kWave && qn * lpe * 2 <= kWave) lpe *= 2;
    return lpe;
}

__device__ bool wave_culled(const Model& M, const float* T, int ts, unsigned long long& flops) {
    const int lane = threadIdx.x & (kWave - 1);
    char* area = M.cull_base + static_cast<size_t>(threadIdx.x / kWave) * M.cull_stride;
    float4* FP = reinterpret_cast<float4*>(area);          // robot spheres, world frame
    float4* CW = FP + M.n_spheres;                          // group bounding spheres, world frame
    int* Q = reinterpret_cast<int*>(CW + M.n_joints);       // refinement queue
    const int ng = M.n_groups, nrs = M.n_spheres;
    const bool mg = M.mode == 1;
    const float cut = mg ? M.margin : 0.0f;
    PRRTC_SIMT_ASSERT(ng <= M.n_joints && M.n_joints <= kMaxJoints);
    PRRTC_SIMT_ASSERT(M.goff[ng] <= nrs);
    PRRTC_SIMT_ASSERT(M.n_wh + M.n_wb + M.n_wc + M.n_ws < (1 << 25));  // entries o << 6 | g
    PRRTC_SIMT_ASSERT(cull_stride_bytes(nrs, M.n_joints, 0) <= M.cull_stride);
    unsigned lf = 0;  // this lane's flops of fine tests
    auto done = [&](bool v) {
        flops += wave_sum_u32(lf);
        return v;
    };
    for (int s0 = 0; s0 < nrs; s0 += kWave) {
        const int s = s0 + lane;
        bool act = false;
        if (s < nrs) {
            const int js = M.sphere_js[s];
            float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
            if (js >= 0) {
                Mat3 R;
                float tx, ty, tz;
                frame_from(T, ts, js, R, tx, ty, tz);
                const float4 sp = M.spheres[s];
                const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
                o = make_float4(p.x, p.y, p.z, sp.w);
                act = true;
            }
            FP[s] = o;
        }
        flops += static_cast<unsigned long long>(__popcll(__ballot(act))) * kFlopsXf;
    }
    for (int g = lane; g < ng; g += kWave) {
        const float4 c = M.gcoarse[g];
        Mat3 R;
        float tx, ty, tz;
        frame_from(T, ts, M.gframe[g], R, tx, ty, tz);
        const float3 p = mat_apply(R, tx, ty, tz, c.x, c.y, c.z);
        CW[g] = make_float4(p.x, p.y, p.z, c.w + cut + kCullPad);  // the coarse limit, less |r_obstacle|
    }
    flops += static_cast<unsigned long long>(ng) * kFlopsXf;
    wave_sync();
    const int nws = M.n_ws, nwc = M.n_wc, nwb = M.n_wb, nwh = M.n_wh;
    const int no = nwh + nwb + nwc;          // halfspaces, boxes, capsules; spheres follow
    const int total = no * ng;               // world primitives are staged in LDS: far below 2^31 / 64
    int qn = 0;
    // Refine once two pairs are queued (or the scan ends), then wait for twice as many after each
    // clear refinement: a colliding configuration exits early, a free one refines in full batches.
    // (First flush at 2, 4, 8, 16 pairs: 7.6, 7.3, 5.7, 5.1 * 10^8 binary verdicts/s on the large
    // scene; the planner's step is the same for all four.)
    int flush_at = 2;
    // A refinement gives each queued pair lpe lanes (the largest power of two with qn * lpe <= 64),
    // which split the group's members: small early batches keep the wave busy too.
    auto refine = [&]() -> bool {
        wave_sync();
        bool h = false;
        const int lpe = lanes_per_entry(qn);
        const int sub = lane & (lpe - 1);
        for (int e = lane / lpe; e < qn && !h; e += kWave / lpe) {
            const int o = Q[e] >> 6, g = Q[e] & 63;
            PRRTC_SIMT_ASSERT(e < kCullQ && o < M.n_wh + M.n_wb + M.n_wc + M.n_ws && g < ng);
            for (int m = M.goff[g] + sub; m < M.goff[g + 1]; m += lpe) {
                if (fine_test_mode(M, mg, o, FP[M.gmem[m]], lf)) {
                    h = true;
                    break;
                }
            }
        }
        qn = 0;
        const bool any = wave_any(h);
        wave_sync();
        return any;
    };
    auto push = [&](bool dirty, int entry) {
        const unsigned long long mask = __ballot(dirty);
        PRRTC_SIMT_ASSERT(qn + static_cast<int>(__popcll(mask)) <= kCullQ);
        if (dirty) Q[qn + static_cast<int>(__popcll(mask & ((1ull << lane) - 1ull)))] = entry;
        qn += static_cast<int>(__popcll(mask));
        PRRTC_SIMT_OBSERVE(0, qn);
    };
    // Pair steps (halfspaces, boxes, capsules): pair p = o * ng + g (obstacle-major), lane l takes
    // p0 + l, and (o, g) advance by kWave pairs per step without a division (ng <= 64).  Sphere
    // chunks (the many): lanes over a chunk of 64 world spheres (the last one partial), each lane
    // testing its sphere against every group sphere (coarse_clear's
    // sphere test; independent LDS broadcasts, no per-group ballot) into a mask of groups, whose
    // set bits are then queued one per lane per step.  Steps fill the queue up to flush_at; one
    // refinement site serves both.
    const int psteps = (total + kWave - 1) / kWave;
    int o = 0, g = 0, dq = 0, dr = 0;
    if (total > 0) {
        dq = kWave / ng;
        dr = kWave - dq * ng;
        o = lane / ng;
        g = lane - (lane / ng) * ng;
    }
    int pstep = 0, k0 = 0, stests = 0;  // stests: sphere-group tests done
    unsigned long long gm = 0;         // this lane's groups still to queue (chunk k0)
    bool loaded = false;
    for (;;) {
        for (; pstep < psteps && qn < flush_at; ++pstep) {
            bool dirty = false;
            int entry = 0;
            if (o < no) {
                const float4 c = CW[g];
                dirty = !coarse_clear(M, o, c, c.w);
                entry = (o << 6) | g;
            }
            o += dq;
            g += dr;
            if (g >= ng) { g -= ng; ++o; }
            push(dirty, entry);
        }
        if (pstep >= psteps) {
            while (k0 < nws && qn < flush_at) {
                if (!loaded) {
                    gm = 0;
                    if (k0 + lane < nws) {
                        const float4 w = M.wsph[k0 + lane];
                        const float rw = fabsf(w.w);
                        for (int sg = 0; sg < ng; ++sg) {
                            const float4 c = CW[sg];
                            const float dx = c.x - w.x, dy = c.y - w.y, dz = c.z - w.z;
                            if (!beyond(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), c.w + rw))
                                gm |= 1ull << sg;
                        }
                    }
                    loaded = true;
                    stests += min(kWave, nws - k0) * ng;
                }
                if (!wave_any(gm != 0)) {
                    k0 += kWave;
                    loaded = false;
                    continue;
                }
                const bool has = gm != 0;
                push(has, ((no + k0 + lane) << 6) | (has ? static_cast<int>(__builtin_ctzll(gm)) : 0));
                gm &= gm - 1ull;
            }
        }
        const bool all = pstep >= psteps && k0 >= nws;
        if (qn > 0 && refine()) {
            flops += coarse_flops_of(M, min(pstep * kWave, total), ng) + static_cast<unsigned long long>(stests) * kFlopsSS;
            return done(true);
        }
        if (all) break;
        flush_at = min(2 * flush_at, kWave);
    }
    flops += coarse_flops_of(M, total, ng) + static_cast<unsigned long long>(nws) * ng * kFlopsSS;
    for (int p0 = 0; p0 < M.n_pairs; p0 += 2 * kWave) {
        const bool h = self_pairs_mode(M, mg, FP, p0, lane);
        flops += static_cast<unsigned long long>(min(2 * kWave, M.n_pairs - p0)) * kFlopsPair;
        if (wave_any(h)) return done(true);
    }
    return done(false);
}

// Verdict of a sub-unit of nu <= kUnitMax configurations whose FK frames are at T + c (c < nu):
// k + 1 if configuration k collides (the first of those found colliding: the scan stops at the
// first wave step with a world hit, so an earlier configuration may collide too), 0 if all are
// free, -1 if the candidate list overflowed (the caller then checks them one by one).  Needs a
// cull area with kUnitCap candidate slots.  Consecutive waypoints of one edge are close, so one
// coarse pass serves all of them: group g's unit sphere has centre C_g (midpoint of its nu transformed centres) and
// radius R_g + cut + pad + the half-diagonal of the box around them (wave_culled's coarse limit
// plus a bound on max_c |p_c - C_g|).  The coarse tests are 1-Lipschitz in the centre, so a pair
// the unit sphere clears is cleared by every configuration's own group sphere (triangle
// inequality); the pairs it cannot clear form a candidate list.  The candidates of all
// configurations are then tested together (wave_culled's coarse and fine tests, 64 pairs per wave
// step, stopping at the first step with a hit), and, if none hits the world, the self pairs of
// each configuration in order.  Whether any configuration collides is wave_culled's verdict.
__device__ int wave_unit_check(const Model& M, const float* T, int ts, int nu, unsigned long long& flops) {
    const int lane = threadIdx.x & (kWave - 1);
    char* area = M.cull_base + static_cast<size_t>(threadIdx.x / kWave) * M.cull_stride;
    float4* FP = reinterpret_cast<float4*>(area);          // robot spheres (per configuration)
    float4* CW = FP + M.n_spheres;                          // (wave_culled's group spheres)
    int* Q = reinterpret_cast<int*>(CW + M.n_joints);       // refinement queue
    int* LU = Q + kCullQ;                                   // candidate pairs of the sub-unit
    float4* CU = FP;  // unit spheres (n_groups <= n_spheres); free until the first transform below
    const int ng = M.n_groups, nrs = M.n_spheres;
    const bool mg = M.mode == 1;
    const float cut = mg ? M.margin : 0.0f;
    PRRTC_SIMT_ASSERT(nu >= 1 && nu <= kUnitMax && nu <= ts);
    PRRTC_SIMT_ASSERT(ng <= M.n_joints && ng <= nrs && M.goff[ng] <= nrs);
    PRRTC_SIMT_ASSERT(M.n_wh + M.n_wb + M.n_wc + M.n_ws < (1 << 21));  // entries o << 10 | u << 6 | g
    PRRTC_SIMT_ASSERT(cull_stride_bytes(nrs, M.n_joints, kUnitCap) <= M.cull_stride);
    unsigned lf = 0;
    auto done = [&](int v) {
        flops += wave_sum_u32(lf);
        return v;
    };
    // ---- unit spheres: the box around each group's nu transformed centres.  With few groups,
    // lanes take (group, configuration) items, kk lanes per group (kk: nu rounded up to a power of
    // two), and the box is a butterfly of minima and maxima over those lanes; otherwise lane g
    // walks the configurations of group g.  Minima and maxima are exact, so both give the same box.
    auto unit_sphere = [&](int g, float lx, float ly, float lz, float hx, float hy, float hz) {
        // every centre lies in the box [l, h]: within its half-diagonal of the midpoint
        const float cx = 0.5f * (lx + hx), cy = 0.5f * (ly + hy), cz = 0.5f * (lz + hz);
        const float ex = hx - lx, ey = hy - ly, ez = hz - lz;
        const float spread = 0.5f * __builtin_sqrtf(__builtin_fmaf(ez, ez, __builtin_fmaf(ey, ey, ex * ex)));
        // the spread rounded up (relative 1e-6); kCullPad covers the midpoint's rounding
        CU[g] = make_float4(cx, cy, cz, (M.gcoarse[g].w + cut + kCullPad) + spread * 1.000001f);
    };
    int kk = 1;
    while (kk < nu) kk *= 2;
    if (ng * kk <= 4 * kWave) {
        const int gpw = kWave / kk;  // groups per wave step
        for (int g0 = 0; g0 < ng; g0 += gpw) {
            const int g = g0 + lane / kk, k = lane & (kk - 1);
            float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
            if (g < ng && k < nu) {
                const float4 c = M.gcoarse[g];
                Mat3 R;
                float tx, ty, tz;
                frame_from(T + k, ts, M.gframe[g], R, tx, ty, tz);
                const float3 p = mat_apply(R, tx, ty, tz, c.x, c.y, c.z);
                lx = hx = p.x; ly = hy = p.y; lz = hz = p.z;
            }
            for (int o = 1; o < kk; o <<= 1) {
                lx = fminf(lx, __shfl_xor(lx, o, kWave)); ly = fminf(ly, __shfl_xor(ly, o, kWave));
                lz = fminf(lz, __shfl_xor(lz, o, kWave)); hx = fmaxf(hx, __shfl_xor(hx, o, kWave));
                hy = fmaxf(hy, __shfl_xor(hy, o, kWave)); hz = fmaxf(hz, __shfl_xor(hz, o, kWave));
            }
            if (g < ng && k == 0) unit_sphere(g, lx, ly, lz, hx, hy, hz);
        }
    } else if (lane < ng) {
        const int g = lane;
        const float4 c = M.gcoarse[g];
        float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
        for (int k = 0; k < nu; ++k) {
            Mat3 R;
            float tx, ty, tz;
            frame_from(T + k, ts, M.gframe[g], R, tx, ty, tz);
            const float3 p = mat_apply(R, tx, ty, tz, c.x, c.y, c.z);
            lx = fminf(lx, p.x); ly = fminf(ly, p.y); lz = fminf(lz, p.z);
            hx = fmaxf(hx, p.x); hy = fmaxf(hy, p.y); hz = fmaxf(hz, p.z);
        }
        unit_sphere(g, lx, ly, lz, hx, hy, hz);
    }
    flops += static_cast<unsigned long long>(ng) * nu * kFlopsXf;
    wave_sync();
    // ---- unit coarse stage: wave_culled's steps against the unit spheres, into LU
    const int nws = M.n_ws, nwc = M.n_wc, nwb = M.n_wb, nwh = M.n_wh;
    const int no = nwh + nwb + nwc;
    const int total = no * ng;
    int nl = 0;
    auto push = [&](bool dirty, int entry) -> bool {
        const unsigned long long mask = __ballot(dirty);
        const int cnt = static_cast<int>(__popcll(mask));
        if (nl + cnt > kUnitCap) return false;
        if (dirty) LU[nl + static_cast<int>(__popcll(mask & ((1ull << lane) - 1ull)))] = entry;
        nl += cnt;
        PRRTC_SIMT_OBSERVE(1, nl);
        return true;
    };
    if (total > 0) {
        const int dq = kWave / ng, dr = kWave - dq * ng;
        int o = lane / ng, g = lane - (lane / ng) * ng;
        for (int p0 = 0; p0 < total; p0 += kWave) {
            bool dirty = false;
            int entry = 0;
            if (o < no) {
                const float4 c = CU[g];
                dirty = !coarse_clear(M, o, c, c.w);
                entry = (o << 6) | g;
            }
            o += dq;
            g += dr;
            if (g >= ng) { g -= ng; ++o; }
            if (!push(dirty, entry)) {
                flops += coarse_flops_of(M, min(p0 + kWave, total), ng);
                return done(-1);
            }
        }
        flops += coarse_flops_of(M, total, ng);
    }
    // Sphere chunks (the last one partial): each lane tests its world sphere against every unit
    // sphere (independent LDS broadcasts and distance tests, no per-group ballot) into a mask of
    // groups, then the few set bits are pushed, one per lane per pass.
    for (int k0 = 0; k0 < nws; k0 += kWave) {
        unsigned long long gm = 0;
        if (k0 + lane < nws) {
            const float4 w = M.wsph[k0 + lane];
            const float rw = fabsf(w.w);
            for (int sg = 0; sg < ng; ++sg) {
                const float4 c = CU[sg];
                const float dx = c.x - w.x, dy = c.y - w.y, dz = c.z - w.z;
                if (!beyond(__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)), c.w + rw)) gm |= 1ull << sg;
            }
        }
        flops += static_cast<unsigned long long>(min(kWave, nws - k0)) * ng * kFlopsSS;
        while (wave_any(gm != 0)) {
            const bool has = gm != 0;
            const int sg = has ? static_cast<int>(__builtin_ctzll(gm)) : 0;
            if (!push(has, ((no + k0 + lane) << 6) | sg)) return done(-1);
            gm &= gm - 1ull;
        }
    }
    wave_sync();
    // ---- world: every (configuration, candidate) pair at once, lanes over pairs.  The coarse
    // test uses the configuration's own group sphere (wave_culled's); the fine tests of a queued
    // (configuration, obstacle, group) triple transform the group's member spheres with their
    // frame as wave_culled does.  hm: bit u set if configuration u hits the world.
    unsigned hm = 0;
    const int nc = nu * nl;
    for (int k0 = 0; k0 < nc; k0 += kWave) {
        const int k = k0 + lane;
        bool dirty = false;
        int entry = 0;
        if (k < nc) {
            const int u = k / nl;
            PRRTC_SIMT_ASSERT(u < nu && k - u * nl < nl);
            const int en = LU[k - u * nl];
            const int o = en >> 6, g = en & 63;
            PRRTC_SIMT_ASSERT(o < M.n_wh + M.n_wb + M.n_wc + M.n_ws && g < ng);
            const float4 c = M.gcoarse[g];
            Mat3 R;
            float tx, ty, tz;
            frame_from(T + u, ts, M.gframe[g], R, tx, ty, tz);
            const float3 p = mat_apply(R, tx, ty, tz, c.x, c.y, c.z);
            const float lim = c.w + cut + kCullPad;
            dirty = !coarse_clear(M, o, make_float4(p.x, p.y, p.z, lim), lim);
            lf += kFlopsXf + ((o < nwh) ? kFlopsSH : (o < nwh + nwb) ? kFlopsSB : (o < no) ? kFlopsSC : kFlopsSS);
            entry = (o << 10) | (u << 6) | g;
        }
        const unsigned long long mask = __ballot(dirty);
        if (dirty) Q[static_cast<int>(__popcll(mask & ((1ull << lane) - 1ull)))] = entry;
        const int qn = static_cast<int>(__popcll(mask));  // <= kWave <= kCullQ
        PRRTC_SIMT_OBSERVE(2, qn);
        wave_sync();
        const int lpe = lanes_per_entry(qn);
        const int sub = lane & (lpe - 1);
        for (int e = lane / lpe; e < qn; e += kWave / lpe) {
            const int o = Q[e] >> 10, u = (Q[e] >> 6) & 15, g = Q[e] & 63;
            PRRTC_SIMT_ASSERT(e < kCullQ && u < nu && g < ng && o < M.n_wh + M.n_wb + M.n_wc + M.n_ws);
            Mat3 R;
            float tx, ty, tz;
            frame_from(T + u, ts, M.gframe[g], R, tx, ty, tz);
            for (int m = M.goff[g] + sub; m < M.goff[g + 1]; m += lpe) {
                const float4 sp = M.spheres[M.gmem[m]];
                const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
                lf += kFlopsXf;
                if (fine_test_mode(M, mg, o, make_float4(p.x, p.y, p.z, sp.w), lf)) {
                    hm |= 1u << u;
                    break;
                }
            }
        }
        wave_sync();
        if (wave_any(hm != 0u)) break;  // a configuration collides: the sub-unit's verdict is known
    }
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) hm |= static_cast<unsigned>(__shfl_xor(static_cast<int>(hm), o, kWave));
    if (hm) {
        const int first = __builtin_ctz(hm);  // the first of the colliding configurations found
        flops += static_cast<unsigned long long>(kFlopsJoint) * M.n_joints * (first + 1);
        return done(first + 1);
    }
    flops += static_cast<unsigned long long>(kFlopsJoint) * M.n_joints * nu;
    // ---- no world hit: the self pairs of each configuration, in order.  Group pairs whose unit
    // spheres are apart clear all their self pairs for every configuration: with the unit radius
    // R + cut + pad + spread of each group, |p_a - p_b| > |r_a| + |r_b| + 2 (cut + pad) for any
    // member spheres a, b, which rules out the contact and (2 cut >= margin) the margin test; a
    // negative margin adds -cut to the separation.  The surviving pairs (their 16-bit indices, at
    // most 2 kCullQ in the queue's space, else all pairs) are then tested for each configuration,
    // their spheres transformed on the spot.
    if (M.n_pairs > 0) {
        unsigned long long* clear = reinterpret_cast<unsigned long long*>(LU);  // bit gj of row gi
        PRRTC_SIMT_ASSERT(ng <= 64 && kUnitCap * static_cast<int>(sizeof(int)) >= 64 * 8);
        const float extra = fmaxf(-cut, 0.0f);
        if (lane < ng) {
            const float4 ci = CU[lane];
            unsigned long long row = 0ull;
            for (int gj = 0; gj < ng; ++gj) {
                const float4 cj = CU[gj];
                const float dx = ci.x - cj.x, dy = ci.y - cj.y, dz = ci.z - cj.z;
                const float lim = ci.w + cj.w + extra;
                if (__builtin_fmaf(dz, dz, __builtin_fmaf(dy, dy, dx * dx)) > lim * lim) row |= 1ull << gj;
            }
            clear[lane] = row;
        }
        flops += static_cast<unsigned long long>(ng) * ng * kFlopsSS;
        wave_sync();
        unsigned short* QS = reinterpret_cast<unsigned short*>(Q);  // surviving pairs (n_pairs < 2^16)
        constexpr int kSurvivors = 2 * kCullQ;
        int ns = 0;
        bool all = false;
        for (int p0 = 0; p0 < M.n_pairs; p0 += kWave) {
            const int k = p0 + lane;
            bool keep = false;
            if (k < M.n_pairs) {
                const unsigned pr = M.pairs[k];
                const int gi = M.fgroup[M.sphere_js[pr & 0xffffu]], gj = M.fgroup[M.sphere_js[pr >> 16]];
                PRRTC_SIMT_ASSERT(gi >= 0 && gi < ng && gj >= 0 && gj < ng);
                keep = ((clear[gi] >> gj) & 1ull) == 0ull;
            }
            const unsigned long long mask = __ballot(keep);
            const int cnt = static_cast<int>(__popcll(mask));
            PRRTC_SIMT_OBSERVE(4, ns + cnt);
            if (ns + cnt > kSurvivors) {
                all = true;
                if (lane == 0) PRRTC_SIMT_OBSERVE(5, 1);
                break;
            }
            if (keep) QS[ns + static_cast<int>(__popcll(mask & ((1ull << lane) - 1ull)))] = static_cast<unsigned short>(k);
            ns += cnt;
        }
        wave_sync();
        if (all) {
            // too many survivors: every pair of every configuration, as the exhaustive loop
            for (int u = 0; u < nu; ++u) {
                const float* Tu = T + u;
                for (int s0 = 0; s0 < nrs; s0 += kWave) {
                    const int s = s0 + lane;
                    bool act = false;
                    if (s < nrs) {
                        const int js = M.sphere_js[s];
                        float4 o = make_float4(0.0f, 0.0f, 0.0f, 0.0f);
                        if (js >= 0) {
                            Mat3 R;
                            float tx, ty, tz;
                            frame_from(Tu, ts, js, R, tx, ty, tz);
                            const float4 sp = M.spheres[s];
                            const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
                            o = make_float4(p.x, p.y, p.z, sp.w);
                            act = true;
                        }
                        FP[s] = o;
                    }
                    flops += static_cast<unsigned long long>(__popcll(__ballot(act))) * kFlopsXf;
                }
                wave_sync();
                for (int p0 = 0; p0 < M.n_pairs; p0 += 2 * kWave) {
                    const bool h = self_pairs_mode(M, mg, FP, p0, lane);
                    flops += static_cast<unsigned long long>(min(2 * kWave, M.n_pairs - p0)) * kFlopsPair;
                    if (wave_any(h)) return done(u + 1);
                }
                wave_sync();
            }
        } else if (ns > 0) {
            PRRTC_SIMT_ASSERT(ns <= kSurvivors && M.n_pairs <= 65536);
            // each survivor transforms its two spheres itself (faster on MI355X than transforming
            // every robot sphere first: C3 296 against 301 ms per step, dual arm 4.9k against
            // 4.7k plans/s)
            for (int u = 0; u < nu; ++u) {
                for (int e0 = 0; e0 < ns; e0 += kWave) {
                    const int e = e0 + lane;
                    bool h = false;
                    if (e < ns) {
                        const unsigned pr = M.pairs[QS[e]];
                        const int ia = static_cast<int>(pr & 0xffffu), ib = static_cast<int>(pr >> 16);
                        const float4 a = M.spheres[ia], b = M.spheres[ib];
                        Mat3 Ra, Rb;
                        float ax, ay, az, bx, by, bz;
                        frame_from(T + u, ts, M.sphere_js[ia], Ra, ax, ay, az);
                        frame_from(T + u, ts, M.sphere_js[ib], Rb, bx, by, bz);
                        const float3 pa = mat_apply(Ra, ax, ay, az, a.x, a.y, a.z);
                        const float3 pb = mat_apply(Rb, bx, by, bz, b.x, b.y, b.z);
                        const float4 ob = make_float4(pb.x, pb.y, pb.z, b.w);
                        h = mg ? near_sphere(pa.x, pa.y, pa.z, a.w, ob, M.margin) : hit_sphere(pa.x, pa.y, pa.z, a.w, ob);
                    }
                    flops += static_cast<unsigned long long>(min(kWave, ns - e0)) * (2 * kFlopsXf + kFlopsPair);
                    if (wave_any(h)) return done(u + 1);
                }
            }
        }
    }
    return done(0);
}

// The planner's verdict for one configuration (prrtc_planner.cu:501-531), wave-uniform; the
// culled check gives the exhaustive loop's verdict (see wave_culled).
__device__ __forceinline__ bool wave_in_collision(const Model& M, const float* T, int ts, unsigned long long& flops) {
    if (!M.enabled) return false;
    flops += static_cast<unsigned long long>(kFlopsJoint) * M.n_joints;
    return wave_culled(M, T, ts, flops);
}

// wave_unit_check with the mode dispatch of wave_in_collision (nu >= 1 configurations at T + c).
__device__ __forceinline__ int wave_unit_in_collision(const Model& M, const float* T, int ts, int nu,
                                                      unsigned long long& flops) {
    if (!M.enabled) return 0;
    return wave_unit_check(M, T, ts, nu, flops);
}

}  // namespace prrtc

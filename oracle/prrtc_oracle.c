/*
 * prrtc_oracle.c -- CPU oracle for the pRRTC planner path.  TEST INFRASTRUCTURE ONLY.
 *
 * This file is the checker, never the product: only tests/, __graft_entry__.smoke() and the
 * cpu_baseline leg of bench.py may load it.  The product path (prrax_amd/) is the HIP library
 * and fails loudly when that library is missing.
 *
 * It restates, in scalar C, the algorithm of the reference planner kernel
 *   cuda-rrtc/CUDA/prrtc_planner.cu:675-1066  (prrtc_planner_kernel)
 * under ONE fixed schedule of that racy kernel -- the lockstep schedule:
 *   every explorer (= reference block) runs phase k of round r before any explorer starts
 *   phase k+1, and shared-state updates are applied in explorer order; an explorer's connect
 *   nearest-neighbour search sees the other tree as it stood at the start of the round.
 * Every lockstep execution is one legal interleaving of the reference's blocks (they have no
 * ordering between them), so the lockstep planner is "the reference algorithm with the race
 * resolved".  The GPU library implements the same schedule; its trees must equal these
 * bit for bit.
 *
 * Pieces and the reference lines they follow:
 *   halton / prime shuffle ........ prrtc_planner.cu:144-157, 745-765, 783-788
 *   tree-side balance .............. prrtc_planner.cu:790-824
 *   nearest neighbour .............. prrtc_planner.cu:841-862, 949-973 (tie -> lowest index,
 *                                    a deliberate change: the reference's reduction is not
 *                                    lowest-index on ties)
 *   steer + dynamic domain ......... prrtc_planner.cu:864-881, 911-923, 925-937, 46-83
 *   edge discretisation ............ prrtc_planner.cu:883-909 (t_k = (k+1)/G)
 *   append / connect / termination . prrtc_planner.cu:925-1062
 *   binary collision predicate ..... prrtc_planner.cu:232-307, 501-531
 *   margin ("differentiable") ...... prrtc_planner.cu:314-387 (collision iff any pair
 *                                    distance < margin, the reference's own definition at
 *                                    prrtc_planner.cu:203-206)
 *   FK ............................. pyroffi fk_single (absent; called at
 *                                    prrtc_planner.cu:509-522): T_j = T_parent o P_j o Exp(tw_j q_j),
 *                                    output permuted by topo_inv.  Written from the
 *                                    mathematics; pinned against an FP64 numpy FK.
 *   primitive distances ............ pyroffi sphere_*_dist (absent; layouts
 *                                    prrtc_planner.cu:171-175, utils.py:58-116).  Written from
 *                                    the geometry; pinned against FP64 closed forms.
 *
 * Floating-point contract (shared with the HIP library, written twice, independently):
 * compiled with -ffp-contract=off; every fused multiply-add is an explicit fmaf(); sin/cos
 * come from the polynomial below (not libm), so CPU and GPU round identically.
 *
 * What pins each function (tests/test_oracle_*.py):
 *   oc_halton, oc_prime_order ...... known answers, closed-form radical inverse, canonical order
 *   oc_sincos ...................... libm accuracy over +-4 pi, symmetries
 *   oc_fk .......................... FP64 SE3 FK (1e-5), q = 0, single revolute = Rodrigues,
 *                                    mimic = driving joint, a transposed rotation is caught
 *   distances / predicates ......... FP64 closed forms and FP64 predicates away from contact
 *   oc_nearest ..................... brute-force argmin with injected exact ties
 *   oc_plan ........................ properties (collision-free nodes/paths under FP64, forest
 *                                    rooted at the roots, dynamic-domain rules, caps, Halton
 *                                    first round, determinism)
 * Parity unpinned: pyroffi's own FP32 operation order in fk_single / apply_se3_point /
 * sphere_*_dist / colldist_from_sdf (the reference's dependency is absent, so its exact FP32
 * results cannot be reproduced; these functions are pinned to the mathematics instead).
 */
#include <float.h>
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "prrtc_oracle.h"

/* OC_PLANT selects one deliberately wrong rule (a swapped comparison, a dropped term, ...) for
 * the tests that show the oracle's rule tests catch such a mistake (tests/test_oracle_rules.py
 * builds each variant into a temporary directory).  Every real build has OC_PLANT == 0. */
#ifndef OC_PLANT
#define OC_PLANT 0
#endif

/* ---------------------------------------------------------------- sin/cos -------------- */
/* Cody-Waite reduction by pi/2 (three-part split of pi/2) and Taylor polynomials on
 * [-pi/4, pi/4]; coefficients are (-1)^k/n! rounded to float. */
#define OC_TWO_OVER_PI 0x1.45f306p-1f
#define OC_PIO2_HI 0x1.921fb6p+0f
#define OC_PIO2_MID -0x1.777a5cp-25f
#define OC_PIO2_LO -0x1.ee59dap-50f
#define OC_S3 -0x1.555556p-3f
#define OC_S5 0x1.111112p-7f
#define OC_S7 -0x1.a01a02p-13f
#define OC_S9 0x1.71de3ap-19f
#define OC_S11 -0x1.ae6456p-26f
#define OC_C4 0x1.555556p-5f
#define OC_C6 -0x1.6c16c2p-10f
#define OC_C8 0x1.a01a02p-16f
#define OC_C10 -0x1.27e4fcp-22f
#define OC_C12 0x1.1eed8ep-29f

void oc_sincos(float x, float* s_out, float* c_out) {
    const float k = rintf(x * OC_TWO_OVER_PI);
    float r = fmaf(-k, OC_PIO2_HI, x);
    r = fmaf(-k, OC_PIO2_MID, r);
    r = fmaf(-k, OC_PIO2_LO, r);
    const float r2 = r * r;
    float ps = OC_S11;
    ps = fmaf(ps, r2, OC_S9);
    ps = fmaf(ps, r2, OC_S7);
    ps = fmaf(ps, r2, OC_S5);
    ps = fmaf(ps, r2, OC_S3);
    const float sn = fmaf(r * r2, ps, r);
    float pc = OC_C12;
    pc = fmaf(pc, r2, OC_C10);
    pc = fmaf(pc, r2, OC_C8);
    pc = fmaf(pc, r2, OC_C6);
    pc = fmaf(pc, r2, OC_C4);
    const float cs = fmaf(r2, fmaf(r2, pc, -0.5f), 1.0f);
    const int q = ((int)k) & 3;
    float s, c;
    if (q == 0) { s = sn; c = cs; }
    else if (q == 1) { s = cs; c = -sn; }
    else if (q == 2) { s = -sn; c = -cs; }
    else { s = -cs; c = sn; }
    *s_out = s;
    *c_out = c;
}

/* ---------------------------------------------------------------- SE3 algebra ---------- */
typedef struct { float w, x, y, z, tx, ty, tz; } oc_tf;
typedef struct { float m[9]; } oc_mat;

static void qmul(float aw, float ax, float ay, float az, float bw, float bx, float by, float bz,
                 float* w, float* x, float* y, float* z) {
    float t;
    t = aw * bw; t = fmaf(-ax, bx, t); t = fmaf(-ay, by, t); t = fmaf(-az, bz, t); *w = t;
    t = aw * bx; t = fmaf(ax, bw, t); t = fmaf(ay, bz, t); t = fmaf(-az, by, t); *x = t;
    t = aw * by; t = fmaf(-ax, bz, t); t = fmaf(ay, bw, t); t = fmaf(az, bx, t); *y = t;
    t = aw * bz; t = fmaf(ax, by, t); t = fmaf(-ay, bx, t); t = fmaf(az, bw, t); *z = t;
}

static oc_mat quat_to_mat(float w, float x, float y, float z) {
    oc_mat R;
    const float x2 = x + x, y2 = y + y, z2 = z + z;
    const float xx = x * x2, yy = y * y2, zz = z * z2;
    const float xy = x * y2, xz = x * z2, yz = y * z2;
    const float wx = w * x2, wy = w * y2, wz = w * z2;
    R.m[0] = 1.0f - (yy + zz); R.m[1] = xy - wz;          R.m[2] = xz + wy;
    R.m[3] = xy + wz;          R.m[4] = 1.0f - (xx + zz); R.m[5] = yz - wx;
    R.m[6] = xz - wy;          R.m[7] = yz + wx;          R.m[8] = 1.0f - (xx + yy);
    return R;
}

/* out = R p + t, each row fma-chained from t. */
static void mat_apply(const oc_mat* R, float tx, float ty, float tz, float px, float py, float pz,
                      float* ox, float* oy, float* oz) {
    *ox = fmaf(R->m[2], pz, fmaf(R->m[1], py, fmaf(R->m[0], px, tx)));
    *oy = fmaf(R->m[5], pz, fmaf(R->m[4], py, fmaf(R->m[3], px, ty)));
    *oz = fmaf(R->m[8], pz, fmaf(R->m[7], py, fmaf(R->m[6], px, tz)));
}

/* ---------------------------------------------------------------- joint records -------- */
enum { JT_FIXED = 0, JT_ROT = 1, JT_PRISM = 2, JT_SCREW = 3 };

typedef struct {
    int type;
    float ax, ay, az; /* unit rotation axis (ROT/SCREW) */
    float wn;         /* |omega| */
    float vx, vy, vz; /* translation part of the twist */
} oc_joint;

static oc_joint joint_prep(const float* tw) {
    oc_joint j;
    const float vx = tw[0], vy = tw[1], vz = tw[2], wx = tw[3], wy = tw[4], wz = tw[5];
    const float wn2 = fmaf(wz, wz, fmaf(wy, wy, wx * wx));
    const float vn2 = fmaf(vz, vz, fmaf(vy, vy, vx * vx));
    j.vx = vx; j.vy = vy; j.vz = vz;
    j.ax = j.ay = j.az = 0.0f;
    j.wn = 0.0f;
    if (wn2 == 0.0f && vn2 == 0.0f) {
        j.type = JT_FIXED;
    } else if (wn2 == 0.0f) {
        j.type = JT_PRISM;
    } else {
        const float wn = sqrtf(wn2);
        j.wn = wn;
        j.ax = wx / wn; j.ay = wy / wn; j.az = wz / wn;
        j.type = (vn2 == 0.0f) ? JT_ROT : JT_SCREW;
    }
    return j;
}

/* Forward kinematics; T_sorted is in the joint (topological) order of the context arrays. */
static void fk_sorted(const oc_context* c, const float* q, oc_tf* T) {
    for (int j = 0; j < c->n_joints; ++j) {
        const oc_joint jr = joint_prep(&c->twists[6 * j]);
        float theta;
        if (c->mimic_act_idx[j] >= 0) {
            theta = fmaf(q[c->mimic_act_idx[j]], c->mimic_mul[j], c->mimic_off[j]);
        } else if (c->act_idx[j] >= 0) {
            theta = q[c->act_idx[j]];
        } else {
            theta = 0.0f;
        }
        const float* P = &c->parent_tf[7 * j];
        oc_tf A;
        const int p = c->parent_idx[j];
        if (p < 0) {
            A.w = P[0]; A.x = P[1]; A.y = P[2]; A.z = P[3]; A.tx = P[4]; A.ty = P[5]; A.tz = P[6];
        } else {
            const oc_tf* Tp = &T[p];
            qmul(Tp->w, Tp->x, Tp->y, Tp->z, P[0], P[1], P[2], P[3], &A.w, &A.x, &A.y, &A.z);
            const oc_mat Rp = quat_to_mat(Tp->w, Tp->x, Tp->y, Tp->z);
            mat_apply(&Rp, Tp->tx, Tp->ty, Tp->tz, P[4], P[5], P[6], &A.tx, &A.ty, &A.tz);
        }
        oc_tf out = A;
        if (jr.type == JT_ROT) {
            float s, co;
            oc_sincos(0.5f * (jr.wn * theta), &s, &co);
            qmul(A.w, A.x, A.y, A.z, co, s * jr.ax, s * jr.ay, s * jr.az, &out.w, &out.x, &out.y, &out.z);
        } else if (jr.type == JT_PRISM) {
            const oc_mat RA = quat_to_mat(A.w, A.x, A.y, A.z);
            mat_apply(&RA, A.tx, A.ty, A.tz, jr.vx * theta, jr.vy * theta, jr.vz * theta,
                      &out.tx, &out.ty, &out.tz);
        } else if (jr.type == JT_SCREW) {
            const float th = jr.wn * theta;
            float sh, ch, s, co;
            oc_sincos(0.5f * th, &sh, &ch);
            oc_sincos(th, &s, &co);
            float ca, cb;
            if (fabsf(th) < 1e-3f) {
                ca = 0.5f * th;
                cb = (th * th) * 0x1.555556p-3f;
            } else {
                ca = (1.0f - co) / th;
                cb = (th - s) / th;
            }
            const float rx = jr.vx * theta, ry = jr.vy * theta, rz = jr.vz * theta;
            const float cx = fmaf(jr.ay, rz, -(jr.az * ry));
            const float cy = fmaf(jr.az, rx, -(jr.ax * rz));
            const float cz = fmaf(jr.ax, ry, -(jr.ay * rx));
            const float dx = fmaf(jr.ay, cz, -(jr.az * cy));
            const float dy = fmaf(jr.az, cx, -(jr.ax * cz));
            const float dz = fmaf(jr.ax, cy, -(jr.ay * cx));
            const float ex = fmaf(cb, dx, fmaf(ca, cx, rx));
            const float ey = fmaf(cb, dy, fmaf(ca, cy, ry));
            const float ez = fmaf(cb, dz, fmaf(ca, cz, rz));
            qmul(A.w, A.x, A.y, A.z, ch, sh * jr.ax, sh * jr.ay, sh * jr.az, &out.w, &out.x, &out.y, &out.z);
            const oc_mat RA = quat_to_mat(A.w, A.x, A.y, A.z);
            mat_apply(&RA, A.tx, A.ty, A.tz, ex, ey, ez, &out.tx, &out.ty, &out.tz);
        }
        T[j] = out;
    }
}

void oc_fk(const oc_context* c, const float* q, float* T_out) {
    oc_tf T[OC_MAX_JOINTS];
    fk_sorted(c, q, T);
    for (int j = 0; j < c->n_joints; ++j) {
        const oc_tf* t = &T[c->topo_inv[j]];
        float* o = &T_out[7 * j];
        o[0] = t->w; o[1] = t->x; o[2] = t->y; o[3] = t->z; o[4] = t->tx; o[5] = t->ty; o[6] = t->tz;
    }
}

/* ---------------------------------------------------------------- collision ------------ */
/* Binary tests: the reference's "dist <= 0" (prrtc_planner.cu:250-304) evaluated without square
 * roots.  A sphere or capsule whose radius sum rr = r + R is >= 0 is hit iff d^2 <= rr^2; with
 * rr < 0 never (sqrt(d^2) - rr > 0): both at once as d^2 <= rr |rr|.  A box with sr >= 0 is hit iff the squared outside distance
 * is <= sr^2 (it is 0 when the centre is inside); with sr < 0 iff the centre lies deeper than
 * |sr| inside (max_i q_i <= sr, q_i = |l_i| - h_i).  A halfspace iff n.(c - p) <= sr.  These
 * agree with the signed-distance form except within rounding of contact. */
static int hit_sphere(float sx, float sy, float sz, float sr, const float* o) {
    const float dx = sx - o[0], dy = sy - o[1], dz = sz - o[2];
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float rr = sr + o[3];
#if OC_PLANT == 11
    return d2 <= rr * rr;
#else
    return d2 <= rr * fabsf(rr);
#endif
}

static float capsule_d2(float sx, float sy, float sz, const float* cap) {
    const float ax = cap[3] - cap[0], ay = cap[4] - cap[1], az = cap[5] - cap[2];
    const float l2 = fmaf(az, az, fmaf(ay, ay, ax * ax));
    const float inv = (l2 > 0.0f) ? 1.0f / l2 : 0.0f;
    const float wx = sx - cap[0], wy = sy - cap[1], wz = sz - cap[2];
    float t = fmaf(wz, az, fmaf(wy, ay, wx * ax)) * inv;
    t = fminf(fmaxf(t, 0.0f), 1.0f);
    const float dx = fmaf(-t, ax, wx), dy = fmaf(-t, ay, wy), dz = fmaf(-t, az, wz);
    return fmaf(dz, dz, fmaf(dy, dy, dx * dx));
}

static int hit_capsule(float sx, float sy, float sz, float sr, const float* cap) {
    const float rr = sr + cap[6];
    return capsule_d2(sx, sy, sz, cap) <= rr * fabsf(rr);
}

/* q_i = |l_i| - h_i of the centre in the box frame; returns the squared outside distance. */
static float box_q(float sx, float sy, float sz, const float* b, float* qmax) {
    const float wx = sx - b[0], wy = sy - b[1], wz = sz - b[2];
    const float l0 = fmaf(wz, b[5], fmaf(wy, b[4], wx * b[3]));
    const float l1 = fmaf(wz, b[8], fmaf(wy, b[7], wx * b[6]));
    const float l2 = fmaf(wz, b[11], fmaf(wy, b[10], wx * b[9]));
    const float q0 = fabsf(l0) - b[12], q1 = fabsf(l1) - b[13], q2 = fabsf(l2) - b[14];
    const float m0 = fmaxf(q0, 0.0f), m1 = fmaxf(q1, 0.0f), m2 = fmaxf(q2, 0.0f);
    *qmax = fmaxf(q0, fmaxf(q1, q2));
    return fmaf(m2, m2, fmaf(m1, m1, m0 * m0));
}

static int hit_box(float sx, float sy, float sz, float sr, const float* b) {
    float qmax;
    const float d2 = box_q(sx, sy, sz, b, &qmax);
    return (sr >= 0.0f) ? (d2 <= sr * sr) : (qmax <= sr);
}

static int hit_halfspace(float sx, float sy, float sz, float sr, const float* h) {
    const float off = fmaf(h[2], h[5], fmaf(h[1], h[4], h[0] * h[3]));
    const float s = fmaf(h[2], sz, fmaf(h[1], sy, h[0] * sx)) - off;
    return s <= sr;
}

/* Margin tests: the reference's "dist < margin" (prrtc_planner.cu:203-206, 314-387), also
 * without square roots.  With L = r + R + margin, a sphere or capsule is near iff L > 0 and
 * d^2 < L^2 (d^2 < L |L|); a box (L = sr + margin) iff L > 0 and outside^2 < L^2 when the centre is outside
 * (outside^2 > 0), else iff min(max_i q_i, 0) < L; a halfspace iff n.(c - p) - sr < margin. */
static int near_sphere(float sx, float sy, float sz, float sr, const float* o, float margin) {
    const float dx = sx - o[0], dy = sy - o[1], dz = sz - o[2];
    const float d2 = fmaf(dz, dz, fmaf(dy, dy, dx * dx));
    const float L = (sr + o[3]) + margin;
    return d2 < L * fabsf(L);
}

static int near_capsule(float sx, float sy, float sz, float sr, const float* cap, float margin) {
    const float L = (sr + cap[6]) + margin;
    return capsule_d2(sx, sy, sz, cap) < L * fabsf(L);
}

static int near_box(float sx, float sy, float sz, float sr, const float* b, float margin) {
    float qmax;
    const float d2 = box_q(sx, sy, sz, b, &qmax);
    const float L = sr + margin;
    return (d2 > 0.0f) ? (L > 0.0f && d2 < L * L) : (fminf(qmax, 0.0f) < L);
}

static int near_halfspace(float sx, float sy, float sz, float sr, const float* h, float margin) {
    const float off = fmaf(h[2], h[5], fmaf(h[1], h[4], h[0] * h[3]));
    return (fmaf(h[2], sz, fmaf(h[1], sy, h[0] * sx)) - off) - sr < margin;
}

/* Signed distances (the min-distance value and the differentiable cost). */
static float dist_sphere(float sx, float sy, float sz, float sr, const float* o) {
    const float dx = sx - o[0], dy = sy - o[1], dz = sz - o[2];
    return sqrtf(fmaf(dz, dz, fmaf(dy, dy, dx * dx))) - (sr + o[3]);
}

static float dist_capsule(float sx, float sy, float sz, float sr, const float* cap) {
    return sqrtf(capsule_d2(sx, sy, sz, cap)) - (sr + cap[6]);
}

static float dist_box(float sx, float sy, float sz, float sr, const float* b) {
    float qmax;
    const float outside = sqrtf(box_q(sx, sy, sz, b, &qmax));
    const float inside = fminf(qmax, 0.0f);
    return (outside + inside) - sr;
}

static float dist_halfspace(float sx, float sy, float sz, float sr, const float* h) {
    const float off = fmaf(h[2], h[5], fmaf(h[1], h[4], h[0] * h[3]));
    return (fmaf(h[2], sz, fmaf(h[1], sy, h[0] * sx)) - off) - sr;
}

int oc_enabled(const oc_context* c) {
    return c->n_joints > 0 && c->n_spheres > 0 &&
           (c->n_world_spheres > 0 || c->n_world_capsules > 0 || c->n_world_boxes > 0 ||
            c->n_world_halfspaces > 0 || c->n_self_pairs > 0);
}

static int link_ok(const oc_context* c, int l) { return l >= 0 && l < c->n_joints; }

static void sphere_world(const oc_context* c, const oc_tf* T, int s, float* p) {
    const oc_tf* t = &T[c->topo_inv[c->sphere_link_idx[s]]];
    const oc_mat R = quat_to_mat(t->w, t->x, t->y, t->z);
    const float* l = &c->sphere_local[3 * s];
    mat_apply(&R, t->tx, t->ty, t->tz, l[0], l[1], l[2], &p[0], &p[1], &p[2]);
}

/* Minimum signed distance over every valid (sphere, obstacle) and self pair; FLT_MAX when
 * there is none. */
float oc_min_distance(const oc_context* c, const float* q) {
    oc_tf T[OC_MAX_JOINTS];
    fk_sorted(c, q, T);
    float best = FLT_MAX;
    for (int s = 0; s < c->n_spheres; ++s) {
        if (!link_ok(c, c->sphere_link_idx[s])) continue;
        float p[3];
        sphere_world(c, T, s, p);
        const float r = c->sphere_radius[s];
        for (int i = 0; i < c->n_world_spheres; ++i) best = fminf(best, dist_sphere(p[0], p[1], p[2], r, &c->world_spheres[4 * i]));
        for (int i = 0; i < c->n_world_capsules; ++i) best = fminf(best, dist_capsule(p[0], p[1], p[2], r, &c->world_capsules[7 * i]));
        for (int i = 0; i < c->n_world_boxes; ++i) best = fminf(best, dist_box(p[0], p[1], p[2], r, &c->world_boxes[15 * i]));
        for (int i = 0; i < c->n_world_halfspaces; ++i) best = fminf(best, dist_halfspace(p[0], p[1], p[2], r, &c->world_halfspaces[6 * i]));
    }
    for (int k = 0; k < c->n_self_pairs; ++k) {
        const int i = c->self_pairs[2 * k], j = c->self_pairs[2 * k + 1];
        if (i < 0 || j < 0 || i >= c->n_spheres || j >= c->n_spheres) continue;
        if (!link_ok(c, c->sphere_link_idx[i]) || !link_ok(c, c->sphere_link_idx[j])) continue;
        float pi[3], pj[3];
        sphere_world(c, T, i, pi);
        sphere_world(c, T, j, pj);
        const float o[4] = {pj[0], pj[1], pj[2], c->sphere_radius[j]};
        best = fminf(best, dist_sphere(pi[0], pi[1], pi[2], c->sphere_radius[i], o));
    }
    return best;
}

/* Collision verdict: binary mode "any pair dist <= 0", margin mode "any pair dist < margin"
 * (prrtc_planner.cu:203-206, 232-387), each pair with the square-root-free tests above. */
int oc_in_collision(const oc_context* c, const float* q) {
    if (!oc_enabled(c)) return 0;
    const int mg = c->collision_mode == 1;
    const float m = c->collision_margin;
    oc_tf T[OC_MAX_JOINTS];
    fk_sorted(c, q, T);
    for (int s = 0; s < c->n_spheres; ++s) {
        if (!link_ok(c, c->sphere_link_idx[s])) continue;
        float p[3];
        sphere_world(c, T, s, p);
        const float r = c->sphere_radius[s];
        for (int i = 0; i < c->n_world_spheres; ++i) {
            const float* o = &c->world_spheres[4 * i];
            if (mg ? near_sphere(p[0], p[1], p[2], r, o, m) : hit_sphere(p[0], p[1], p[2], r, o)) return 1;
        }
        for (int i = 0; i < c->n_world_capsules; ++i) {
            const float* o = &c->world_capsules[7 * i];
            if (mg ? near_capsule(p[0], p[1], p[2], r, o, m) : hit_capsule(p[0], p[1], p[2], r, o)) return 1;
        }
        for (int i = 0; i < c->n_world_boxes; ++i) {
            const float* o = &c->world_boxes[15 * i];
            if (mg ? near_box(p[0], p[1], p[2], r, o, m) : hit_box(p[0], p[1], p[2], r, o)) return 1;
        }
        for (int i = 0; i < c->n_world_halfspaces; ++i) {
            const float* o = &c->world_halfspaces[6 * i];
            if (mg ? near_halfspace(p[0], p[1], p[2], r, o, m) : hit_halfspace(p[0], p[1], p[2], r, o)) return 1;
        }
    }
    for (int k = 0; k < c->n_self_pairs; ++k) {
        const int i = c->self_pairs[2 * k], j = c->self_pairs[2 * k + 1];
        if (i < 0 || j < 0 || i >= c->n_spheres || j >= c->n_spheres) continue;
        if (!link_ok(c, c->sphere_link_idx[i]) || !link_ok(c, c->sphere_link_idx[j])) continue;
        float pi[3], pj[3];
        sphere_world(c, T, i, pi);
        sphere_world(c, T, j, pj);
        const float o[4] = {pj[0], pj[1], pj[2], c->sphere_radius[j]};
        const float ri = c->sphere_radius[i];
        if (mg ? near_sphere(pi[0], pi[1], pi[2], ri, o, m) : hit_sphere(pi[0], pi[1], pi[2], ri, o)) return 1;
    }
    return 0;
}

void oc_in_collision_batch(const oc_context* c, const float* qs, int n, uint8_t* out) {
    for (int i = 0; i < n; ++i) out[i] = (uint8_t)oc_in_collision(c, &qs[(size_t)i * c->dim]);
}

void oc_min_distance_batch(const oc_context* c, const float* qs, int n, float* out) {
    for (int i = 0; i < n; ++i) out[i] = oc_min_distance(c, &qs[(size_t)i * c->dim]);
}

/* Point-cloud verdict of one configuration: some valid robot sphere s and cloud point p with
 * |c_s - p| <= r_s + r_env, by hit_sphere with the point as a sphere of radius r_env (the
 * definition the point-cloud edge checker uses, prrtc_robogpu.py:103-298).  Brute force over
 * every point: no grid. */
int oc_cloud_hit(const oc_context* c, const float* q, const float* points, int n_points, float r_env) {
    oc_tf T[OC_MAX_JOINTS];
    fk_sorted(c, q, T);
    for (int s = 0; s < c->n_spheres; ++s) {
        if (!link_ok(c, c->sphere_link_idx[s])) continue;
        float p[3];
        sphere_world(c, T, s, p);
        const float r = c->sphere_radius[s];
        for (int k = 0; k < n_points; ++k) {
            const float o[4] = {points[3 * k], points[3 * k + 1], points[3 * k + 2], r_env};
            if (hit_sphere(p[0], p[1], p[2], r, o)) return 1;
        }
    }
    return 0;
}

void oc_cloud_hit_batch(const oc_context* c, const float* qs, int n, const float* points, int n_points, float r_env,
                        uint8_t* out) {
    for (int i = 0; i < n; ++i) out[i] = (uint8_t)oc_cloud_hit(c, &qs[(size_t)i * c->dim], points, n_points, r_env);
}

void oc_sphere_positions(const oc_context* c, const float* q, float* out /* [n_spheres,3] */) {
    oc_tf T[OC_MAX_JOINTS];
    fk_sorted(c, q, T);
    for (int s = 0; s < c->n_spheres; ++s) {
        if (!link_ok(c, c->sphere_link_idx[s])) { out[3 * s] = out[3 * s + 1] = out[3 * s + 2] = NAN; continue; }
        sphere_world(c, T, s, &out[3 * s]);
    }
}

/* ---------------------------------------------------------------- sampler -------------- */
float oc_halton(int base, int index) {
    float fraction = 1.0f, result = 0.0f;
    int n = index;
    while (n > 0) {
        fraction /= (float)base;
        result += fraction * (float)(n % base);
        n /= base;
    }
    return result;
}

void oc_prime_order(int explorer, unsigned seed, int* out16) {
    int primes[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};
    if (explorer != 0 || seed != 0u) {
        unsigned rng = ((unsigned)explorer + seed * 65536u) * 1664525u + 1013904223u;
        for (int i = 15; i > 0; --i) {
            rng = rng * 1664525u + 1013904223u;
            const int j = (int)(rng % (unsigned)(i + 1));
            const int tmp = primes[i]; primes[i] = primes[j]; primes[j] = tmp;
        }
    }
    for (int d = 0; d < 16; ++d) out16[d] = primes[d];
}

/* ---------------------------------------------------------------- nearest neighbour ---- */
static void nn_scan(const float* tree, int max_nodes, int size, int dim, const float* x,
                    int* best_i, float* best_d) {
    float bd = FLT_MAX;
    int bi = 0;
    for (int i = 0; i < size; ++i) {
        float dsq = 0.0f;
        for (int d = 0; d < dim; ++d) {
            const float diff = tree[(size_t)d * max_nodes + i] - x[d];
            dsq = fmaf(diff, diff, dsq);
        }
        if (dsq < bd) { bd = dsq; bi = i; }
    }
    *best_i = bi;
    *best_d = bd;
}

void oc_nearest(const float* tree, int max_nodes, int size, int dim, const float* qs, int nq,
                int* idx, float* dsq) {
    for (int k = 0; k < nq; ++k) nn_scan(tree, max_nodes, size, dim, &qs[(size_t)k * dim], &idx[k], &dsq[k]);
}

/* ---------------------------------------------------------------- planner -------------- */
typedef struct {
    float* cfg;   /* [dim, N] */
    int* parent;  /* [N] */
    float* radius;/* [N] */
    int size;
} oc_tree;

/* ---------------------------------------------------------------- planner rules -------- */
/* Dynamic-domain radius of a node whose extension collided (prrtc_planner.cu:46-64): an
 * unbounded radius (FLT_MAX) becomes dd_radius, any other shrinks by (1 - alpha), not below
 * dd_min_radius. */
void oc_dd_shrink(float* r, float alpha, float r0, float rmin) {
    const float old = *r;
#if OC_PLANT == 4
    *r = (old == FLT_MAX) ? r0 : fmaxf(old * (1.0f + alpha), rmin);
#elif OC_PLANT == 5
    *r = (old == FLT_MAX) ? r0 : fminf(old * (1.0f - alpha), rmin);
#else
    *r = (old == FLT_MAX) ? r0 : fmaxf(old * (1.0f - alpha), rmin);
#endif
}

/* Radius of a node that was extended successfully (prrtc_planner.cu:66-83): grows by
 * (1 + alpha); an unbounded radius stays unbounded. */
void oc_dd_grow(float* r, float alpha) {
#if OC_PLANT == 6
    *r = *r * (1.0f + alpha);
#else
    if (*r != FLT_MAX) *r = *r * (1.0f + alpha);
#endif
}

/* Dynamic-domain skip (prrtc_planner.cu:868): the sample lies outside the nearest node's
 * radius. */
int oc_dd_skips(float radius, float nd) {
#if OC_PLANT == 7
    return radius <= nd;
#else
    return radius < nd;
#endif
}

/* Tree that explorer e of S extends in round `round` (prrtc_planner.cu:790-824); *vamp is the
 * explorer's balance-mode-2 state (0 or 1, 0 before the first round), updated in place.  Round
 * 1 always splits the explorers in halves (mode 0). */
int oc_balance_side(int mode, int round, int e, int S, int size_a, int size_b, float tree_ratio, int* vamp) {
    const int total = size_a + size_b;
#if OC_PLANT == 8
    if (mode == 0 || round == 1) return (e <= S / 2) ? 0 : 1;
#else
    if (mode == 0 || round == 1) return (e < S / 2) ? 0 : 1;
#endif
    if (mode == 1) {
        const int diff = (size_a >= size_b) ? (size_a - size_b) : (size_b - size_a);
        const float ratio = (total > 0) ? ((float)size_a / (float)total) : 0.5f;
#if OC_PLANT == 1
        if (diff < (int)(1.5f * (float)S)) return (e < (int)((float)S * (1.0f - ratio))) ? 0 : 1;
        return (ratio > tree_ratio) ? 0 : 1;
#elif OC_PLANT == 2
        if (diff < (int)(1.5f * (float)S)) return (e < (int)((float)S * ratio)) ? 0 : 1;
        return (ratio < tree_ratio) ? 0 : 1;
#else
        if (diff < (int)(1.5f * (float)S)) return (e < (int)((float)S * (1.0f - ratio))) ? 0 : 1;
        return (ratio < tree_ratio) ? 0 : 1;
#endif
    }
    if (mode == 2) {
        const int o_id = 1 - *vamp;
        const int t_sz = (*vamp == 0) ? size_a : size_b;
        const int o_sz = (o_id == 0) ? size_a : size_b;
        const float ratio = (t_sz > 0) ? fabsf((float)(t_sz - o_sz) / (float)t_sz) : 0.0f;
#if OC_PLANT == 3
        if (ratio >= tree_ratio) *vamp = 1 - *vamp;
#else
        if (ratio < tree_ratio) *vamp = 1 - *vamp;
#endif
        return *vamp;
    }
#if OC_PLANT == 9
    return (size_a < size_b) ? 0 : 1;
#else
    return (size_a <= size_b) ? 0 : 1;
#endif
}

/* Steer (prrtc_planner.cu:870-876): from base toward sample by at most `step`, where nd is the
 * base-to-sample distance; a sample within 1e-8 is taken as is. */
void oc_steer(const float* base, const float* sample, int dim, float nd, float step, float* out) {
#if OC_PLANT == 10
    const float sc = (nd > 1e-8f) ? fminf(1.0f, nd / step) : 1.0f;
#else
    const float sc = (nd > 1e-8f) ? fminf(1.0f, step / nd) : 1.0f;
#endif
    for (int d = 0; d < dim; ++d) out[d] = base[d] + (sample[d] - base[d]) * sc;
}

/* Connect walk (prrtc_planner.cu:975-993): the number of sub-edges n = ceil(dist / step) (0 when
 * the step or the distance is within 1e-8) from `from` to `to`, and the sub-edge vector
 * (to - from) / n (zero when n == 0).  Returns n. */
int oc_connect_steps(const float* from, const float* to, int dim, float dist, float step, float* vec) {
    int n = 0;
#if OC_PLANT == 12
    if (step > 1e-8f && dist > 1e-8f) n = (int)floorf(dist / step);
#else
    if (step > 1e-8f && dist > 1e-8f) n = (int)ceilf(dist / step);
#endif
    for (int d = 0; d < dim; ++d) vec[d] = (n > 0) ? (to[d] - from[d]) / (float)n : 0.0f;
    return n;
}

static int edge_in_collision(const oc_context* c, int dim, int G, const float* base, const float* delta_or_end,
                             int is_connect, float* tmp) {
    for (int k = 0; k < G; ++k) {
        const float tf = (float)(k + 1) / (float)G;
        for (int d = 0; d < dim; ++d) {
            if (is_connect) tmp[d] = base[d] + tf * delta_or_end[d];
            else tmp[d] = base[d] + tf * (delta_or_end[d] - base[d]);
        }
        if (oc_in_collision(c, tmp)) return 1;
    }
    return 0;
}

int oc_plan(const oc_context* c, const oc_params* p, const float* start, const float* goals, int num_goals,
            const float* min_vals, const float* max_vals, oc_outputs* out) {
    const int dim = p->dim, N = p->max_nodes, S = p->num_new_samples, G = p->granularity;
    if (dim < 1 || dim > OC_MAX_DIM || N < 1 || S < 1 || G < 1 || num_goals < 1 || num_goals > N ||
        c->n_joints > OC_MAX_JOINTS)
        return -1;
    oc_tree tr[2];
    tr[0].cfg = out->tree_a; tr[0].parent = out->parents_a; tr[0].radius = out->radii_a;
    tr[1].cfg = out->tree_b; tr[1].parent = out->parents_b; tr[1].radius = out->radii_b;
    for (int d = 0; d < dim; ++d) tr[0].cfg[(size_t)d * N] = start[d];
    tr[0].parent[0] = 0;
    tr[0].radius[0] = FLT_MAX;
    tr[0].size = 1;
    for (int g = 0; g < num_goals; ++g) {
        for (int d = 0; d < dim; ++d) tr[1].cfg[(size_t)d * N + g] = goals[(size_t)g * dim + d];
        tr[1].parent[g] = g;
        tr[1].radius[g] = FLT_MAX;
    }
    tr[1].size = num_goals;

    int* prime = (int*)malloc(sizeof(int) * 16 * S);
    int* side = (int*)malloc(sizeof(int) * S);
    int* nn = (int*)malloc(sizeof(int) * S);
    int* state = (int*)malloc(sizeof(int) * S); /* 0 skip, 1 collided, 2 free */
    int* newi = (int*)malloc(sizeof(int) * S);
    int* onn = (int*)malloc(sizeof(int) * S);
    int* next = (int*)malloc(sizeof(int) * S);
    int* nok = (int*)malloc(sizeof(int) * S);
    int* vamp = (int*)calloc(S, sizeof(int));
    float* cand = (float*)malloc(sizeof(float) * S * dim);
    float* vec = (float*)malloc(sizeof(float) * S * dim);
    float tmp[OC_MAX_DIM], sample[OC_MAX_DIM], curr[OC_MAX_DIM];
    for (int e = 0; e < S; ++e) oc_prime_order(e, p->seed, &prime[16 * e]);

    int status = 0, r = 0;
    out->conn[0] = out->conn[1] = out->conn[2] = -1;
    long long evals = 0;
    for (r = 1;; ++r) {
        if (r > p->max_iterations) { status = OC_STATUS_ITER_CAP; break; }
        const int size_a = tr[0].size, size_b = tr[1].size;
        /* phase 1: sample, side, nearest neighbour, dynamic-domain skip, steer */
        for (int e = 0; e < S; ++e) {
            for (int d = 0; d < dim; ++d) {
                const float val = oc_halton(prime[16 * e + (d % 16)], r);
                sample[d] = min_vals[d] + val * (max_vals[d] - min_vals[d]);
            }
            const int t = oc_balance_side(p->balance_mode, r, e, S, size_a, size_b, p->tree_ratio, &vamp[e]);
            side[e] = t;
            float bd;
            nn_scan(tr[t].cfg, N, tr[t].size, dim, sample, &nn[e], &bd);
            const float nd = sqrtf(bd);
            if (p->dynamic_domain && oc_dd_skips(tr[t].radius[nn[e]], nd)) { state[e] = 0; continue; }
            float base[OC_MAX_DIM];
            for (int d = 0; d < dim; ++d) base[d] = tr[t].cfg[(size_t)d * N + nn[e]];
            oc_steer(base, sample, dim, nd, p->step_size, &cand[e * dim]);
            state[e] = 1;
        }
        /* phase 2: extend-edge FK+CC */
        for (int e = 0; e < S; ++e) {
            if (state[e] == 0) continue;
            float base[OC_MAX_DIM];
            for (int d = 0; d < dim; ++d) base[d] = tr[side[e]].cfg[(size_t)d * N + nn[e]];
            evals += G;
            state[e] = edge_in_collision(c, dim, G, base, &cand[e * dim], 0, tmp) ? 1 : 2;
        }
        /* phase 3: commit.  All dynamic-domain shrinks, then appends (with grows), in explorer order. */
        if (p->dynamic_domain)
            for (int e = 0; e < S; ++e)
                if (state[e] == 1) oc_dd_shrink(&tr[side[e]].radius[nn[e]], p->dd_alpha, p->dd_radius, p->dd_min_radius);
        for (int e = 0; e < S; ++e) {
            newi[e] = -1;
            if (state[e] != 2) continue;
            oc_tree* T = &tr[side[e]];
            if (T->size >= N) { if (status == 0) status = OC_STATUS_TREE_FULL; continue; }
            const int idx = T->size++;
            for (int d = 0; d < dim; ++d) T->cfg[(size_t)d * N + idx] = cand[e * dim + d];
            T->parent[idx] = nn[e];
            T->radius[idx] = FLT_MAX;
            if (p->dynamic_domain) oc_dd_grow(&T->radius[nn[e]], p->dd_alpha);
            newi[e] = idx;
        }
        if (status != 0) break;
        /* phase 4: connect nearest neighbour in the other tree, as the tree stood at the start of
         * the round (the interleaving in which every explorer's connect search runs before the
         * other explorers' appends of this round; the GPU does it during the extend checks) */
        for (int e = 0; e < S; ++e) {
            if (newi[e] < 0) continue;
            const oc_tree* O = &tr[1 - side[e]];
            float bd;
            nn_scan(O->cfg, N, side[e] ? size_a : size_b, dim, &cand[e * dim], &onn[e], &bd);
            float target[OC_MAX_DIM];
            for (int d = 0; d < dim; ++d) target[d] = O->cfg[(size_t)d * N + onn[e]];
            next[e] = oc_connect_steps(&cand[e * dim], target, dim, sqrtf(bd), p->step_size, &vec[e * dim]);
        }
        /* phase 5: connect sub-edges, first failing one ends the walk */
        for (int e = 0; e < S; ++e) {
            if (newi[e] < 0) continue;
            for (int d = 0; d < dim; ++d) curr[d] = cand[e * dim + d];
            int k = 0;
            for (; k < next[e]; ++k) {
                evals += G;
                if (edge_in_collision(c, dim, G, curr, &vec[e * dim], 1, tmp)) break;
                for (int d = 0; d < dim; ++d) curr[d] += vec[e * dim + d];
            }
            nok[e] = k;
        }
        /* phase 6: connect appends, explorer-major; first success or overflow decides */
        for (int e = 0; e < S; ++e) {
            if (newi[e] < 0) continue;
            oc_tree* T = &tr[side[e]];
            int parent = newi[e], overflow = 0;
            for (int d = 0; d < dim; ++d) curr[d] = cand[e * dim + d];
            for (int k = 0; k < nok[e]; ++k) {
                for (int d = 0; d < dim; ++d) curr[d] += vec[e * dim + d];
                if (T->size >= N) { overflow = 1; break; }
                const int idx = T->size++;
                for (int d = 0; d < dim; ++d) T->cfg[(size_t)d * N + idx] = curr[d];
                T->parent[idx] = parent;
                T->radius[idx] = FLT_MAX;
                parent = idx;
            }
            if (overflow) {
                if (status == 0) status = OC_STATUS_TREE_FULL;
            } else if (nok[e] == next[e] && status == 0) {
                status = OC_STATUS_SOLVED;
                if (side[e] == 0) { out->conn[0] = parent; out->conn[1] = onn[e]; }
                else { out->conn[0] = onn[e]; out->conn[1] = parent; }
                out->conn[2] = side[e];
            }
        }
        if (status != 0) break;
    }
    out->sizes[0] = tr[0].size;
    out->sizes[1] = tr[1].size;
    out->iterations = r;
    out->status = status;
    out->fkcc_evals = evals;
    free(prime); free(side); free(nn); free(state); free(newi); free(onn); free(next); free(nok);
    free(vamp); free(cand); free(vec);
    return 0;
}

/* Path from the start root to a goal root (prrtc.py:588-640): walk tree A from conn[0] back to
 * its root, reverse, then walk tree B from conn[1] to its root.  Returns the node count, or -1
 * when the path would exceed max_len. */
int oc_trace_path(const float* tree_a, const float* tree_b, const int* par_a, const int* par_b, int max_nodes,
                  int dim, const int* conn, float* path, int max_len) {
    int na = 0, cur = conn[0];
    int* ia = (int*)malloc(sizeof(int) * max_nodes);
    for (int it = 0; it < max_nodes; ++it) {
        ia[na++] = cur;
        const int par = par_a[cur];
        if (par == cur) break;
        cur = par;
    }
    int len = 0;
    for (int k = na - 1; k >= 0; --k) {
        if (len >= max_len) { free(ia); return -1; }
        for (int d = 0; d < dim; ++d) path[(size_t)len * dim + d] = tree_a[(size_t)d * max_nodes + ia[k]];
        ++len;
    }
    free(ia);
    cur = conn[1];
    for (int it = 0; it < max_nodes; ++it) {
        if (len >= max_len) return -1;
        for (int d = 0; d < dim; ++d) path[(size_t)len * dim + d] = tree_b[(size_t)d * max_nodes + cur];
        ++len;
        const int par = par_b[cur];
        if (par == cur) break;
        cur = par;
    }
    return len;
}

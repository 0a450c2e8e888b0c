// prrtc_common.h -- LDS staging of robot + world and work-group primitives shared by the
// planner and the standalone kernels.
#pragma once
#include "prrtc_device.h"

namespace prrtc {

__host__ __device__ inline size_t align16(size_t x) { return (x + 15) & ~static_cast<size_t>(15); }

// Device-side views of the C-ABI robot and world (by value in kernel arguments).
struct RobotDev {
    const float* twists;
    const float* parent_tf;
    const int* parent_idx;
    const int* act_idx;
    const float* mimic_mul;
    const float* mimic_off;
    const int* mimic_act_idx;
    const int* topo_inv;
    int n_joints;
    const int* sphere_link_idx;
    const float* sphere_local;
    const float* sphere_radius;
    int n_spheres;
};

struct WorldDev {
    const float* ws;
    const float* wc;
    const float* wb;
    const float* wh;
    const int* sp;
    int n_ws, n_wc, n_wb, n_wh, n_sp;              // capacity (padded) counts
    long long s_ws, s_wc, s_wb, s_wh, s_sp;        // per-query strides in elements (0 = shared)
    const int* cnt[5];                             // per-query counts [B] each (spheres, capsules, boxes,
                                                   // halfspaces, self pairs), null = the capacity
};

// Byte offsets of the model inside dynamic LDS, followed by the per-wave FK areas: wave w owns
// wq + w * kMaxDim * fbp floats (configurations, element c of dim d at [d * fbp + c]) and
// wT + w * nj * 7 * fbp floats (joint poses, element c of slot k at [k * fbp + c]).
struct ModelLayout {
    size_t joints, spheres, sphere_js, pairs, wsph, wcap, wbox, whs, gc, gf, go, gm, fg, wq, wT, wcull, end;
    int fbp, cull_stride;
};

__host__ __device__ inline ModelLayout model_layout(int nj, int nrs, int nsp, int nws, int nwc, int nwb, int nwh,
                                                      int nwaves, int fbp, int ucap = 0) {
    ModelLayout L;
    size_t o = 0;
    L.joints = o; o = align16(o + sizeof(JointRec) * nj);
    L.spheres = o; o = align16(o + sizeof(float4) * nrs);
    L.sphere_js = o; o = align16(o + sizeof(int) * nrs);
    L.pairs = o; o = align16(o + sizeof(unsigned) * nsp);
    L.wsph = o; o = align16(o + sizeof(float4) * nws);
    L.wcap = o; o = align16(o + sizeof(CapsuleL) * nwc);
    L.wbox = o; o = align16(o + sizeof(BoxL) * nwb);
    L.whs = o; o = align16(o + sizeof(float4) * nwh);
    L.gc = o; o = align16(o + sizeof(float4) * nj);
    L.gf = o; o = align16(o + sizeof(int) * nj);
    L.go = o; o = align16(o + sizeof(int) * (nj + 1));
    L.gm = o; o = align16(o + sizeof(int) * nrs);
    L.fg = o; o = align16(o + sizeof(int) * nj);
    L.wq = o; o = align16(o + sizeof(float) * kMaxDim * fbp * nwaves);
    L.wT = o; o = align16(o + sizeof(float) * 7 * nj * fbp * nwaves);
    L.cull_stride = static_cast<int>(align16(static_cast<size_t>(cull_stride_bytes(nrs, nj, ucap))));
    L.wcull = o; o = align16(o + static_cast<size_t>(L.cull_stride) * nwaves);
    L.end = o;
    L.fbp = fbp;
    return L;
}

// ---------------------------------------------------------------------------------- LDS atomics
// Atomics on LDS words through explicit address-space-3 pointers, so they lower to ds_* atomics
// (no generic-pointer address-space test).
#define PRRTC_LDS __attribute__((address_space(3)))
__device__ __forceinline__ PRRTC_LDS int* as_lds(int* p) { return (PRRTC_LDS int*)(p); }
// Flag reads/writes are volatile word accesses (LDS is coherent within the work-group); atomic
// load/store builtins on these pointers hit a backend legalisation error when packed FP32 is
// disabled.
__device__ __forceinline__ int lds_load(int* p) { return *reinterpret_cast<volatile int*>(p); }
__device__ __forceinline__ void lds_store(int* p, int v) { *reinterpret_cast<volatile int*>(p) = v; }
__device__ __forceinline__ void lds_min(int* p, int v) {
    __hip_atomic_fetch_min(as_lds(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ void lds_add(int* p, int v) {
    __hip_atomic_fetch_add(as_lds(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_WORKGROUP);
}
__device__ __forceinline__ int lds_cas(int* p, int expected, int desired) {
    __hip_atomic_compare_exchange_strong(as_lds(p), &expected, desired, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                         __HIP_MEMORY_SCOPE_WORKGROUP);
    return expected;
}

// Device-scope atomics on global memory through address-space-1 pointers (plain global_atomic_*;
// a generic pointer would make the compiler test the address space at run time).
#define PRRTC_GLOBAL __attribute__((address_space(1)))
__device__ __forceinline__ int global_add(int* p, int v) {
    return __hip_atomic_fetch_add((PRRTC_GLOBAL int*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void global_add(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_add((PRRTC_GLOBAL unsigned long long*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Loads/stores of data shared by the work-groups of a team (COH = true): relaxed agent-scope
// atomics, which the hardware performs coherently across the XCDs' separate L2 caches, so a
// team barrier needs no L2 write-back/invalidate.  COH = false: plain accesses (one work-group).
template <bool COH, class T>
__device__ __forceinline__ T ldc(const T* p) {
    if constexpr (COH) return __hip_atomic_load((const PRRTC_GLOBAL T*)(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else return *(const PRRTC_GLOBAL T*)(p);
}
template <bool COH, class T>
__device__ __forceinline__ void stc(T* p, T v) {
    if constexpr (COH) __hip_atomic_store((PRRTC_GLOBAL T*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    else *(PRRTC_GLOBAL T*)(p) = v;
}

// ---------------------------------------------------------------------------------- scans
__device__ __forceinline__ int wave_incl_scan(int v) {
    const int lane = threadIdx.x & (kWave - 1);
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(v, o, kWave);
        if (lane >= o) v += t;
    }
    return v;
}

// Exclusive prefix sum over the work-group; scratch holds >= 17 ints of LDS.  Every thread must
// call it.  Returns the exclusive prefix of v; *total receives the sum.
__device__ __forceinline__ int block_excl_scan(int v, int* scratch, int* total) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    const int incl = wave_incl_scan(v);
    if (lane == kWave - 1 || tid == static_cast<int>(blockDim.x) - 1) scratch[wid] = incl;
    __syncthreads();
    if (tid == 0) {
        int run = 0;
        for (int w = 0; w < nw; ++w) {
            const int t = scratch[w];
            scratch[w] = run;
            run += t;
        }
        scratch[16] = run;
    }
    __syncthreads();
    const int res = scratch[wid] + incl - v;
    *total = scratch[16];
    __syncthreads();
    return res;
}

// Exclusive prefix sums of two per-thread values over the work-group (thread order), with their
// totals, behind two barriers (each wave scans both values; the waves' totals go through LDS).
// scratch needs 2 * (waves) ints (at most 16).
__device__ __forceinline__ void block_excl_scan2(int v0, int v1, int* scratch, int& p0, int& p1, int& tot0,
                                                 int& tot1) {
    const int tid = threadIdx.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave, nw = (blockDim.x + kWave - 1) / kWave;
    const int i0 = wave_incl_scan(v0), i1 = wave_incl_scan(v1);
    if (lane == kWave - 1) {
        scratch[wid] = i0;
        scratch[8 + wid] = i1;
    }
    __syncthreads();
    int b0 = 0, b1 = 0, t0 = 0, t1 = 0;
    for (int w = 0; w < nw; ++w) {
        const int c0 = scratch[w], c1 = scratch[8 + w];
        if (w < wid) { b0 += c0; b1 += c1; }
        t0 += c0;
        t1 += c1;
    }
    p0 = b0 + i0 - v0;
    p1 = b1 + i1 - v1;
    tot0 = t0;
    tot1 = t1;
    __syncthreads();  // the scratch words may be rewritten by the next scan
}

// ---------------------------------------------------------------------------------- staging
// Stages robot + world of query q into LDS and fills M.  Must be called by the whole
// work-group; ends with a barrier.  pair_cap is the LDS capacity for self pairs.
// Indices that would reach outside the staged arrays (a parent that is not an earlier joint, an
// actuated/mimic index outside [0, dim), a topo_inv entry outside [0, n_joints)) are treated as
// absent; the host-side context check rejects such robots before they get here.
__device__ inline void stage_model(Model& M, char* lds, const ModelLayout& Lo, const RobotDev& r, const WorldDev& w,
                                   int q, int dim, int mode, float margin, int* scan_scratch) {
    const int tid = threadIdx.x, nt = blockDim.x;
    JointRec* J = reinterpret_cast<JointRec*>(lds + Lo.joints);
    float4* SP = reinterpret_cast<float4*>(lds + Lo.spheres);
    int* SJ = reinterpret_cast<int*>(lds + Lo.sphere_js);
    unsigned* PR = reinterpret_cast<unsigned*>(lds + Lo.pairs);
    float4* WS = reinterpret_cast<float4*>(lds + Lo.wsph);
    CapsuleL* WC = reinterpret_cast<CapsuleL*>(lds + Lo.wcap);
    BoxL* WB = reinterpret_cast<BoxL*>(lds + Lo.wbox);
    float4* WH = reinterpret_cast<float4*>(lds + Lo.whs);
    const int nj = r.n_joints, nrs = r.n_spheres;
    int nws = w.n_ws, nwc = w.n_wc, nwb = w.n_wb, nwh = w.n_wh, nsp = w.n_sp;
    // per-query counts, clamped to [0, capacity]
    if (w.cnt[0]) nws = max(0, min(w.cnt[0][q], w.n_ws));
    if (w.cnt[1]) nwc = max(0, min(w.cnt[1][q], w.n_wc));
    if (w.cnt[2]) nwb = max(0, min(w.cnt[2][q], w.n_wb));
    if (w.cnt[3]) nwh = max(0, min(w.cnt[3][q], w.n_wh));
    if (w.cnt[4]) nsp = max(0, min(w.cnt[4][q], w.n_sp));
    const float* ws = w.ws + w.s_ws * q;
    const float* wc = w.wc + w.s_wc * q;
    const float* wb = w.wb + w.s_wb * q;
    const float* wh = w.wh + w.s_wh * q;
    const int* sp = w.sp + w.s_sp * q;
    for (int j = tid; j < nj; j += nt) {
        JointRec rec;
        joint_prep(r.twists + 6 * j, rec);
        const int p = r.parent_idx[j], ac = r.act_idx[j], mi = r.mimic_act_idx[j];
        rec.parent = (p >= 0 && p < j) ? p : -1;
        rec.act = (ac >= 0 && ac < dim) ? ac : -1;
        rec.mimic = (mi >= 0 && mi < dim) ? mi : -1;
        rec.mul = r.mimic_mul[j];
        rec.off = r.mimic_off[j];
        const float* P = r.parent_tf + 7 * j;
        rec.pw = P[0]; rec.px = P[1]; rec.py = P[2]; rec.pz = P[3];
        rec.ptx = P[4]; rec.pty = P[5]; rec.ptz = P[6];
        J[j] = rec;
    }
    for (int s = tid; s < nrs; s += nt) {
        const int l = r.sphere_link_idx[s];
        const int ti = (l >= 0 && l < nj) ? r.topo_inv[l] : -1;
        SJ[s] = (ti >= 0 && ti < nj) ? ti : -1;
        SP[s] = make_float4(r.sphere_local[3 * s], r.sphere_local[3 * s + 1], r.sphere_local[3 * s + 2],
                            r.sphere_radius[s]);
    }
    for (int k = tid; k < nws; k += nt)
        WS[k] = make_float4(ws[4 * k], ws[4 * k + 1], ws[4 * k + 2], ws[4 * k + 3]);
    for (int k = tid; k < nwc; k += nt) WC[k] = make_capsule(wc + 7 * k);
    for (int k = tid; k < nwb; k += nt) {
        const float* b = wb + 15 * k;
        BoxL o;
        o.cx = b[0]; o.cy = b[1]; o.cz = b[2];
        o.a0x = b[3]; o.a0y = b[4]; o.a0z = b[5];
        o.a1x = b[6]; o.a1y = b[7]; o.a1z = b[8];
        o.a2x = b[9]; o.a2y = b[10]; o.a2z = b[11];
        o.h0 = b[12]; o.h1 = b[13]; o.h2 = b[14];
        // orthonormal axes (Gram matrix within 1e-4 of the identity) make the box test a
        // 1-Lipschitz distance, which the culled check relies on
        const float g00 = o.a0x * o.a0x + o.a0y * o.a0y + o.a0z * o.a0z;
        const float g11 = o.a1x * o.a1x + o.a1y * o.a1y + o.a1z * o.a1z;
        const float g22 = o.a2x * o.a2x + o.a2y * o.a2y + o.a2z * o.a2z;
        const float g01 = o.a0x * o.a1x + o.a0y * o.a1y + o.a0z * o.a1z;
        const float g02 = o.a0x * o.a2x + o.a0y * o.a2y + o.a0z * o.a2z;
        const float g12 = o.a1x * o.a2x + o.a1y * o.a2y + o.a1z * o.a2z;
        const bool ortho = fabsf(g00 - 1.0f) < 1e-4f && fabsf(g11 - 1.0f) < 1e-4f && fabsf(g22 - 1.0f) < 1e-4f &&
                           fabsf(g01) < 1e-4f && fabsf(g02) < 1e-4f && fabsf(g12) < 1e-4f;
        o.ortho = ortho ? 1.0f : 0.0f;
        WB[k] = o;
    }
    for (int k = tid; k < nwh; k += nt) WH[k] = make_halfspace(wh + 6 * k);
    __syncthreads();
    // a serial chain of actuated revolute joints takes fk_lane_chain (same arithmetic, no branches)
    int chain = nj > 0 ? 1 : 0;
    for (int j = 0; j < nj; ++j) {
        const JointRec& R = J[j];
        if (R.type != kRot || R.mimic >= 0 || R.act < 0 || R.parent != j - 1) chain = 0;
    }
    M.chain = chain;
    // Self pairs: keep the valid ones (both indices in range, both spheres on a valid frame),
    // in their original order.
    int base = 0;
    for (int k0 = 0; k0 < nsp; k0 += nt) {
        const int k = k0 + tid;
        int i = -1, j = -1, ok = 0;
        if (k < nsp) {
            i = sp[2 * k]; j = sp[2 * k + 1];
            ok = (i >= 0 && j >= 0 && i < nrs && j < nrs && SJ[i] >= 0 && SJ[j] >= 0) ? 1 : 0;
        }
        int tot;
        const int pre = block_excl_scan(ok, scan_scratch, &tot);
        if (ok) PR[base + pre] = static_cast<unsigned>(i) | (static_cast<unsigned>(j) << 16);
        base += tot;
    }
    // Culling groups: the robot spheres of each frame (frames in order), with a bounding sphere
    // centred on their local AABB, radius max |p - c| + |r|.  Thread j < nj handles frame j.
    float4* GC = reinterpret_cast<float4*>(lds + Lo.gc);
    int* GF = reinterpret_cast<int*>(lds + Lo.gf);
    int* GO = reinterpret_cast<int*>(lds + Lo.go);
    int* GM = reinterpret_cast<int*>(lds + Lo.gm);
    int* FG = reinterpret_cast<int*>(lds + Lo.fg);
    int cnt = 0;
    if (tid < nj)
        for (int s2 = 0; s2 < nrs; ++s2) cnt += (SJ[s2] == tid) ? 1 : 0;
    int n_groups, n_members;
    const int gid = block_excl_scan((tid < nj && cnt > 0) ? 1 : 0, scan_scratch, &n_groups);
    const int goff = block_excl_scan(tid < nj ? cnt : 0, scan_scratch, &n_members);
    if (tid < nj && cnt > 0) {
        float lx = FLT_MAX, ly = FLT_MAX, lz = FLT_MAX, hx = -FLT_MAX, hy = -FLT_MAX, hz = -FLT_MAX;
        int w = goff;
        for (int s2 = 0; s2 < nrs; ++s2) {
            if (SJ[s2] != tid) continue;
            GM[w++] = s2;
            const float4 q = SP[s2];
            lx = fminf(lx, q.x); ly = fminf(ly, q.y); lz = fminf(lz, q.z);
            hx = fmaxf(hx, q.x); hy = fmaxf(hy, q.y); hz = fmaxf(hz, q.z);
        }
        const float cx = 0.5f * (lx + hx), cy = 0.5f * (ly + hy), cz = 0.5f * (lz + hz);
        float rad = 0.0f;
        for (int m = goff; m < w; ++m) {
            const float4 q = SP[GM[m]];
            const float dx = q.x - cx, dy = q.y - cy, dz = q.z - cz;
            rad = fmaxf(rad, __builtin_sqrtf(dx * dx + dy * dy + dz * dz) + fabsf(q.w));
        }
        GC[gid] = make_float4(cx, cy, cz, rad);
        GF[gid] = tid;
        GO[gid] = goff;
    }
    if (tid == 0) GO[n_groups] = n_members;
    if (tid < nj) FG[tid] = (cnt > 0) ? gid : -1;
    M.gcoarse = GC; M.gframe = GF; M.goff = GO; M.gmem = GM; M.fgroup = FG; M.n_groups = n_groups;
    M.cull_base = lds + Lo.wcull;
    M.cull_stride = Lo.cull_stride;
    M.joints = J; M.spheres = SP; M.sphere_js = SJ; M.pairs = PR;
    M.wsph = WS; M.wcap = WC; M.wbox = WB; M.whs = WH;
    M.n_joints = nj; M.n_spheres = nrs; M.n_pairs = base;
    M.n_ws = nws; M.n_wc = nwc; M.n_wb = nwb; M.n_wh = nwh;
    M.enabled = (nj > 0 && nrs > 0 && (nws > 0 || nwc > 0 || nwb > 0 || nwh > 0 || nsp > 0)) ? 1 : 0;
    M.mode = mode;
    M.margin = margin;
    __syncthreads();
}

// Per-wave FK area of the calling wave.
struct WaveArea {
    float* q;
    float* T;
};

__device__ __forceinline__ WaveArea wave_area(char* lds, const ModelLayout& Lo, int nj) {
    const int w = threadIdx.x / kWave;
    WaveArea a;
    a.q = reinterpret_cast<float*>(lds + Lo.wq) + w * kMaxDim * Lo.fbp;
    a.T = reinterpret_cast<float*>(lds + Lo.wT) + w * nj * 7 * Lo.fbp;
    return a;
}


__device__ __forceinline__ int uniform(int v) { return __builtin_amdgcn_readfirstlane(v); }

// ---------------------------------------------------------------------------------- nearest neighbour
constexpr int kEC = 16;  // queries per tree pass

__host__ __device__ inline size_t nn_bytes() { return align16(sizeof(float) * kEC * kMaxDim + 2 * sizeof(int) * kEC * 16); }

__device__ __forceinline__ bool lex_less(float da, int ia, float db, int ib) {
    return da < db || (da == db && ia < ib);
}

// Nearest node (lowest index on ties) of kEC queries (qsh[c * kMaxDim + d], unused rows zero)
// among tree[begin, end).  Every thread of the work-group calls it; on return red_d[c * 16] and
// red_i[c * 16] hold query c's squared distance and index ((FLT_MAX, 0) for an empty range).
// The buffers are passed as LDS pointers (cast by the caller, see lds_f / lds_i): this function
// is not inlined, and generic pointers would make every query read a flat load that waits for
// all outstanding memory traffic.
__device__ __forceinline__ PRRTC_LDS float* lds_f(float* p) { return (PRRTC_LDS float*)(p); }
__device__ __forceinline__ PRRTC_LDS int* lds_i(int* p) { return (PRRTC_LDS int*)(p); }
// COH: nodes at indices >= coh_from may have been written by other work-groups since the last
// grid-wide fence, so they are read with coherent loads; older nodes never change and are read
// through the caches (a wave reads coherently if any of its nodes needs it).
// Rows: the range is cut into rows of nn_row(MAXD) nodes (one scan step of the work-group), and
// only rows row0, row0 + rows, ... are scanned, so that `rows` parts of one tree split it
// evenly, its newest nodes (read coherently) included.
__host__ __device__ constexpr int nn_row_nodes_per_thread(int maxd) { return maxd <= 8 ? 2 : 1; }
template <int MAXD, bool COH = false>
__device__ void nn_tile(const float* __restrict__ tree, int ncap, int begin, int end, int dim,
                        const PRRTC_LDS float* qsh, PRRTC_LDS float* red_d, PRRTC_LDS int* red_i,
                        int coh_from = 0, int row0 = 0, int rows = 1) {
    const int tid = threadIdx.x, nt = blockDim.x, lane = tid & (kWave - 1), wid = tid / kWave;
    const int nw = (nt + kWave - 1) / kWave;
    float bd[kEC];
    int bi[kEC];
#pragma unroll
    for (int c = 0; c < kEC; ++c) { bd[c] = FLT_MAX; bi[c] = 0; }
    // Dimensions d >= dim are zero in the node and in every query row, and fma(0, 0, s) == s, so
    // summing all MAXD terms gives the dim-term sum exactly, without a branch per term.
    // The query rows are re-read (LDS broadcast) for every node: qoff is an opaque zero that
    // keeps the compiler from hoisting all kEC * MAXD query values into registers.  Load indices
    // are clamped into [begin, end).
    int qoff = 0;
    auto load_node = [&](int i, float* xo) {
        const int il = min(i, end - 1);
        const bool coh = COH && __ballot(il >= coh_from) != 0ull;
#pragma unroll
        for (int d = 0; d < MAXD; ++d) {
            const float* p = tree + static_cast<size_t>(min(d, dim - 1)) * ncap + il;
            const float v = coh ? ldc<true>(p) : ldc<false>(p);
            xo[d] = (d < dim) ? v : 0.0f;
        }
    };
    // Raw coordinates (dimensions d >= dim load row dim - 1; the caller zeroes them before use).
    auto load_raw = [&](int i, float* xo) {
        const int il = min(i, end - 1);
        if (COH && __ballot(il >= coh_from) != 0ull) {
#pragma unroll
            for (int d = 0; d < MAXD; ++d) xo[d] = ldc<true>(tree + static_cast<size_t>(min(d, dim - 1)) * ncap + il);
        } else {
#pragma unroll
            for (int d = 0; d < MAXD; ++d) xo[d] = ldc<false>(tree + static_cast<size_t>(min(d, dim - 1)) * ncap + il);
        }
    };
    if constexpr (MAXD <= 8) {
        // Two nodes per step, i and i + nt, scored against each query row read once; the next
        // step's nodes are loaded before this step is scored (indices clamped, so the loads past
        // the end are harmless), and kept raw until then, so that no wait for them comes before
        // the scoring.  A thread visits its nodes in increasing index order (i before i + nt), so
        // on equal distances the lower index still wins.
        const int stride = rows * 2 * nt, first = begin + row0 * 2 * nt + tid;
        float ra[MAXD], rb[MAXD];
        if (first < end) {
            load_raw(first, ra);
            load_raw(first + nt, rb);
        }
        for (int i = first; i < end; i += stride) {
            asm volatile("" : "+v"(qoff));
            // (this copy also frees ra / rb for the next loads: without it -- dim folded to 7 in
            // the 7-term instantiation -- the compiler waited for the prefetch at the top of the
            // loop, and the sample NN of a 64-member team took 76 instead of 53 us per round)
            float xa[MAXD], xb[MAXD];
#pragma unroll
            for (int d = 0; d < MAXD; ++d) {
                xa[d] = (d < dim) ? ra[d] : 0.0f;
                xb[d] = (d < dim) ? rb[d] : 0.0f;
            }
            load_raw(i + stride, ra);
            load_raw(i + stride + nt, rb);
            const bool hb = i + nt < end;
#pragma unroll
            for (int c = 0; c < kEC; ++c) {
                const PRRTC_LDS float* qr = qsh + c * kMaxDim + qoff;
                float da = 0.0f, db = 0.0f;
#pragma unroll
                for (int d = 0; d < MAXD; ++d) {
                    const float q = qr[d];
                    const float ea = xa[d] - q, eb = xb[d] - q;
                    da = __builtin_fmaf(ea, ea, da);
                    db = __builtin_fmaf(eb, eb, db);
                }
                if (da < bd[c]) { bd[c] = da; bi[c] = i; }
                if (hb && db < bd[c]) { bd[c] = db; bi[c] = i + nt; }
            }
        }
    } else {
        for (int i = begin + row0 * nt + tid; i < end; i += rows * nt) {
            asm volatile("" : "+v"(qoff));
            float x[MAXD];
            load_node(i, x);
#pragma unroll
            for (int c = 0; c < kEC; ++c) {
                const PRRTC_LDS float* qr = qsh + c * kMaxDim + qoff;
                float dsq = 0.0f;
#pragma unroll
                for (int d = 0; d < MAXD; ++d) {
                    const float df = x[d] - qr[d];
                    dsq = __builtin_fmaf(df, df, dsq);
                }
                if (dsq < bd[c]) { bd[c] = dsq; bi[c] = i; }
            }
        }
    }
    // Wave minimum of every query in lex_less order (smallest distance, then lowest index), by
    // halving: at each step a lane keeps half of its queries and merges the partner lane's values
    // of the same queries (lane ^ 32, ^ 16, ^ 8, ^ 4); lanes 4q .. 4q + 3 then hold query q and
    // finish with two more exchanges.  The minimum is associative, so the order does not matter.
    {
        auto merge = [&](float kd, int ki, float od, int oi, float& d, int& i) {
            const bool take = lex_less(od, oi, kd, ki);
            d = take ? od : kd;
            i = take ? oi : ki;
        };
        float d8[8], d4[4], d2[2], d1;
        int i8[8], i4[4], i2[2], i1;
        const bool u32 = (lane & 32) != 0, u16 = (lane & 16) != 0, u8 = (lane & 8) != 0, u4 = (lane & 4) != 0;
#pragma unroll
        for (int j = 0; j < 8; ++j) {
            const float od = __shfl_xor(u32 ? bd[j] : bd[8 + j], 32, kWave);
            const int oi = __shfl_xor(u32 ? bi[j] : bi[8 + j], 32, kWave);
            merge(u32 ? bd[8 + j] : bd[j], u32 ? bi[8 + j] : bi[j], od, oi, d8[j], i8[j]);
        }
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            const float od = __shfl_xor(u16 ? d8[j] : d8[4 + j], 16, kWave);
            const int oi = __shfl_xor(u16 ? i8[j] : i8[4 + j], 16, kWave);
            merge(u16 ? d8[4 + j] : d8[j], u16 ? i8[4 + j] : i8[j], od, oi, d4[j], i4[j]);
        }
#pragma unroll
        for (int j = 0; j < 2; ++j) {
            const float od = __shfl_xor(u8 ? d4[j] : d4[2 + j], 8, kWave);
            const int oi = __shfl_xor(u8 ? i4[j] : i4[2 + j], 8, kWave);
            merge(u8 ? d4[2 + j] : d4[j], u8 ? i4[2 + j] : i4[j], od, oi, d2[j], i2[j]);
        }
        {
            const float od = __shfl_xor(u4 ? d2[0] : d2[1], 4, kWave);
            const int oi = __shfl_xor(u4 ? i2[0] : i2[1], 4, kWave);
            merge(u4 ? d2[1] : d2[0], u4 ? i2[1] : i2[0], od, oi, d1, i1);
        }
#pragma unroll
        for (int o = 2; o > 0; o >>= 1) {
            const float od = __shfl_xor(d1, o, kWave);
            const int oi = __shfl_xor(i1, o, kWave);
            merge(d1, i1, od, oi, d1, i1);
        }
        if ((lane & 3) == 0) {
            red_d[(lane >> 2) * 16 + wid] = d1;
            red_i[(lane >> 2) * 16 + wid] = i1;
        }
    }
    __syncthreads();
    if (tid < kEC) {
        float d = red_d[tid * 16];
        int i = red_i[tid * 16];
        for (int w = 1; w < nw; ++w) {
            const float od = red_d[tid * 16 + w];
            const int oi = red_i[tid * 16 + w];
            if (lex_less(od, oi, d, i)) { d = od; i = oi; }
        }
        red_d[tid * 16] = d;
        red_i[tid * 16] = i;
    }
    __syncthreads();
}

// Nearest node of each listed query (query list[k] at qbuf[list[k] * dim]); results to
// out_i/out_d indexed by the list entry.
template <int MAXD>
__device__ void nn_tiles(const float* __restrict__ tree, int ncap, int size, int dim, const int* list, int nl,
                         const float* __restrict__ qbuf, int* out_i, float* out_d, char* nnl) {
    const int tid = threadIdx.x, nt = blockDim.x;
    float* qsh = reinterpret_cast<float*>(nnl);
    float* red_d = qsh + kEC * kMaxDim;
    int* red_i = reinterpret_cast<int*>(red_d + kEC * 16);
    for (int c0 = 0; c0 < nl; c0 += kEC) {
        const int nc = min(kEC, nl - c0);
        for (int k = tid; k < kEC * MAXD; k += nt) {
            const int c = k / MAXD, d = k - (k / MAXD) * MAXD;
            qsh[c * kMaxDim + d] = (c < nc && d < dim) ? qbuf[list[c0 + c] * dim + d] : 0.0f;
        }
        __syncthreads();
        nn_tile<MAXD>(tree, ncap, 0, size, dim, lds_f(qsh), lds_f(red_d), lds_i(red_i));
        if (tid < nc) {
            const int e = list[c0 + tid];
            out_i[e] = red_i[tid * 16];
            out_d[e] = red_d[tid * 16];
        }
        __syncthreads();
    }
}

}  // namespace prrtc

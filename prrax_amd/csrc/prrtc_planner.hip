// prrtc_planner.hip -- persistent lockstep pRRTC planner for MI355X (gfx950) + its C ABI.
//
// One work-group plans one query at a time and pulls the next query from a device-resident
// queue (an atomic counter) when it finishes, so a batch of uneven queries runs in a single
// launch.  Inside a work-group the reference's explorers (its blocks, prrtc_planner.cu:657-674)
// advance in lockstep rounds:
//
//   1  sample (Halton, prime shuffle), pick a tree, nearest neighbour, dynamic-domain skip, steer
//                                                         prrtc_planner.cu:783-881
//   2  FK + collision of every extend edge (G waypoints each)  prrtc_planner.cu:883-909
//   3  commit: all radius shrinks, then appends in explorer order with radius grows
//                                                         prrtc_planner.cu:911-947, 46-83
//   4  connect nearest neighbour in the other tree, step count and step vector
//                                                         prrtc_planner.cu:949-994
//   5  connect sub-edges, evaluated in passes until each explorer's first failing one is known
//                                                         prrtc_planner.cu:996-1018
//   6  connect appends in explorer order; the first success or overflow decides the query
//                                                         prrtc_planner.cu:1020-1062
//
// Each round is one legal interleaving of the reference kernel, chosen so that results do not
// depend on timing: the trees equal the CPU oracle's bit for bit.  Shared state changes only in
// the commit phases, through prefix sums over explorers (no per-node atomics), which also removes
// the reference's unordered-append race (a node is counted only after it is written).
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdio>
#include <cstring>
#include <string>

#include "prrtc_amd.h"
#include "prrtc_common.h"
#include "prrtc_host.h"

namespace prrtc {

constexpr int kPassMult = 2;   // connect items per pass ~= kPassMult * threads

struct PlanArgs {
    int max_iterations;
    float step_size;
    int S;
    int balance_mode;
    float tree_ratio;
    int dd;
    float dd_alpha, dd_radius, dd_min_radius;
    int dim, ncap, G, mode;
    float margin;
    int B;
    RobotDev robot;
    WorldDev world;
    const float* starts;
    const float* goals;
    int gmax;
    const int* goal_counts;
    const float* min_vals;
    const float* max_vals;
    const int* max_nodes_q;
    const unsigned* seeds;
    const int* order;
    float* tree_a;
    float* tree_b;
    int* par_a;
    int* par_b;
    int* sizes;
    int* completed;
    int* iters;
    int* conn;
    int* solved;
    float* time_ms;
    int* status;
    float* path;
    int* path_len;
    int path_cap;
    float* radii;   // [B][2][ncap]
    float* xbuf;    // [B][3][S][dim]: sample, candidate, connect step
    int* queue;     // [1]
    unsigned long long* stats;  // [4]
    ModelLayout ml;
    size_t off_expl, off_primes, off_nn, off_hash, off_misc;
    int hash_size;
    float wall_khz;
};

// Explorer arrays in LDS (S ints each).
enum { X_SIDE, X_NN, X_STATE, X_NEWI, X_ONN, X_NEXT, X_FAIL, X_KD, X_VAMP, X_LIST, X_OFFS, X_AUX0, X_AUX1, X_DSQ,
       X_COUNT };

__host__ __device__ inline size_t expl_bytes(int S) { return align16(sizeof(int) * (size_t)(S + 1) * X_COUNT); }
__host__ __device__ inline int hash_size_for(int S) {
    int h = 64;
    while (h < 2 * S) h <<= 1;
    return h;
}

enum { SC_QUERY = 0, SC_SIZE_A, SC_SIZE_B, SC_STATUS, SC_EVENT, SC_OVERFLOW, SC_CONN0, SC_CONN1, SC_CONN2, SC_COUNT = 16 };

__device__ __forceinline__ float halton(int base, int index) {
    float fraction = 1.0f, result = 0.0f;
    int n = index;
    while (n > 0) {
        fraction /= static_cast<float>(base);
        result += fraction * static_cast<float>(n % base);
        n /= base;
    }
    return result;
}

__device__ __forceinline__ void prime_order(int e, unsigned seed, unsigned char* p) {
    const unsigned char P[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = P[k];
    if (e != 0 || seed != 0u) {
        unsigned rng = (static_cast<unsigned>(e) + seed * 65536u) * 1664525u + 1013904223u;
        for (int i = 15; i > 0; --i) {
            rng = rng * 1664525u + 1013904223u;
            const int j = static_cast<int>(rng % static_cast<unsigned>(i + 1));
            const unsigned char t = p[i];
            p[i] = p[j];
            p[j] = t;
        }
    }
}

template <class F>
__device__ __forceinline__ int compact(int S, F pred, int* list, int* scan) {
    const int tid = threadIdx.x, nt = blockDim.x;
    int base = 0;
    for (int e0 = 0; e0 < S; e0 += nt) {
        const int e = e0 + tid;
        const int ok = (e < S && pred(e)) ? 1 : 0;
        int tot;
        const int pre = block_excl_scan(ok, scan, &tot);
        if (ok) list[base + pre] = e;
        base += tot;
    }
    __syncthreads();
    return base;
}

__device__ __forceinline__ void radius_shrink(float& r, float alpha, float r0, float rmin) {
    r = (r == FLT_MAX) ? r0 : fmaxf(r * (1.0f - alpha), rmin);
}
__device__ __forceinline__ void radius_grow(float& r, float alpha) {
    if (r != FLT_MAX) r = r * (1.0f + alpha);
}

__device__ __forceinline__ void hash_add(int* hk, int* hs, int* hg, int hsize, int key, bool shrink) {
    unsigned h = (static_cast<unsigned>(key) * 2654435761u) & static_cast<unsigned>(hsize - 1);
    for (;;) {
        const int prev = lds_cas(&hk[h], -1, key);
        if (prev == -1 || prev == key) {
            lds_add(shrink ? &hs[h] : &hg[h], 1);
            return;
        }
        h = (h + 1) & static_cast<unsigned>(hsize - 1);
    }
}

template <int MAXD>
__device__ void plan_query(const PlanArgs& a, int q, char* lds) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int dim = a.dim, S = a.S, G = a.G, ncap = a.ncap;
    const int ng = a.goal_counts ? max(1, min(a.goal_counts[q], a.gmax)) : a.gmax;
    const int N = a.max_nodes_q ? min(max(a.max_nodes_q[q], ng), ncap) : ncap;
    const unsigned seed = a.seeds ? a.seeds[q] : 0u;
    int* X = reinterpret_cast<int*>(lds + a.off_expl);
    int* side = X + X_SIDE * (S + 1);
    int* nn = X + X_NN * (S + 1);
    int* state = X + X_STATE * (S + 1);
    int* newi = X + X_NEWI * (S + 1);
    int* onn = X + X_ONN * (S + 1);
    int* next = X + X_NEXT * (S + 1);
    int* fail = X + X_FAIL * (S + 1);
    int* kd = X + X_KD * (S + 1);
    int* vamp = X + X_VAMP * (S + 1);
    int* list = X + X_LIST * (S + 1);
    int* offs = X + X_OFFS * (S + 1);
    int* aux0 = X + X_AUX0 * (S + 1);
    int* aux1 = X + X_AUX1 * (S + 1);
    float* dsq = reinterpret_cast<float*>(X + X_DSQ * (S + 1));
    unsigned char* primes = reinterpret_cast<unsigned char*>(lds + a.off_primes);
    char* nnl = lds + a.off_nn;
    int* hk = reinterpret_cast<int*>(lds + a.off_hash);
    int* hs = hk + a.hash_size;
    int* hg = hs + a.hash_size;
    int* scan = reinterpret_cast<int*>(lds + a.off_misc);
    int* sc = scan + 20;

    const unsigned long long t0 = wall_clock64();
    Model M;
    stage_model(M, lds, a.ml, a.robot, a.world, q, a.dim, a.mode, a.margin, scan);
    const int lane = tid & (kWave - 1), wid = tid / kWave, nwaves = nt / kWave;
    const WaveArea W = wave_area(lds, a.ml, M.n_joints);
    LaneScratch L;  // FK slot of this lane (lanes < kFB) in the wave's FK area
    L.q = W.q + lane;
    L.T = W.T + lane;
    L.nt = kFBP;

    float* TA = a.tree_a + static_cast<size_t>(q) * dim * ncap;
    float* TB = a.tree_b + static_cast<size_t>(q) * dim * ncap;
    int* PA = a.par_a + static_cast<size_t>(q) * ncap;
    int* PB = a.par_b + static_cast<size_t>(q) * ncap;
    float* RA = a.radii + static_cast<size_t>(q) * 2 * ncap;
    float* RB = RA + ncap;
    float* samp = a.xbuf + static_cast<size_t>(q) * 3 * S * dim;
    float* cand = samp + S * dim;
    float* vec = cand + S * dim;
    const float* start = a.starts + static_cast<size_t>(q) * dim;
    const float* goals = a.goals + static_cast<size_t>(q) * a.gmax * dim;

    for (int d = tid; d < dim; d += nt) TA[static_cast<size_t>(d) * ncap] = start[d];
    for (int i = tid; i < ng * dim; i += nt) {
        const int g = i / dim, d = i - (i / dim) * dim;
        TB[static_cast<size_t>(d) * ncap + g] = goals[i];
    }
    if (tid == 0) { PA[0] = 0; RA[0] = FLT_MAX; }
    for (int g = tid; g < ng; g += nt) { PB[g] = g; RB[g] = FLT_MAX; }
    for (int e = tid; e < S; e += nt) { prime_order(e, seed, primes + 16 * e); vamp[e] = 0; }
    for (int h = tid; h < a.hash_size; h += nt) { hk[h] = -1; hs[h] = 0; hg[h] = 0; }
    if (tid == 0) {
        sc[SC_SIZE_A] = 1; sc[SC_SIZE_B] = ng; sc[SC_STATUS] = 0;
        sc[SC_CONN0] = -1; sc[SC_CONN1] = -1; sc[SC_CONN2] = -1;
    }
    __syncthreads();

    unsigned long long evals = 0, flops = 0;  // lane 0 of each wave counts
    int r = 0;
    for (r = 1;; ++r) {
        if (r > a.max_iterations) {
            if (tid == 0) sc[SC_STATUS] = PRRTC_STATUS_ITER_CAP;
            break;
        }
        const int size_a = sc[SC_SIZE_A], size_b = sc[SC_SIZE_B];
        // ---- phase 1a: samples and tree side
        for (int e = tid; e < S; e += nt) {
            const unsigned char* pe = primes + 16 * e;
            for (int d = 0; d < dim; ++d) {
                const float val = halton(pe[d & 15], r);
                samp[e * dim + d] = a.min_vals[d] + val * (a.max_vals[d] - a.min_vals[d]);
            }
            int t;
            const int total = size_a + size_b;
            if (a.balance_mode == 0 || r == 1) {
                t = (e < S / 2) ? 0 : 1;
            } else if (a.balance_mode == 1) {
                const int diff = (size_a >= size_b) ? (size_a - size_b) : (size_b - size_a);
                const float ratio = (total > 0) ? (static_cast<float>(size_a) / static_cast<float>(total)) : 0.5f;
                if (diff < static_cast<int>(1.5f * static_cast<float>(S)))
                    t = (e < static_cast<int>(static_cast<float>(S) * (1.0f - ratio))) ? 0 : 1;
                else
                    t = (ratio < a.tree_ratio) ? 0 : 1;
            } else if (a.balance_mode == 2) {
                const int v = vamp[e];
                const int t_sz = (v == 0) ? size_a : size_b;
                const int o_sz = (v == 0) ? size_b : size_a;
                const float ratio = (t_sz > 0) ? fabsf(static_cast<float>(t_sz - o_sz) / static_cast<float>(t_sz)) : 0.0f;
                const int nv = (ratio < a.tree_ratio) ? 1 - v : v;
                vamp[e] = nv;
                t = nv;
            } else {
                t = (size_a <= size_b) ? 0 : 1;
            }
            side[e] = t;
        }
        __syncthreads();
        // ---- phase 1b: nearest neighbour in the chosen tree
        {
            const int n0 = compact(S, [&](int e) { return side[e] == 0; }, list, scan);
            nn_tiles<MAXD>(TA, ncap, size_a, dim, list, n0, samp, nn, dsq, nnl);
            const int n1 = compact(S, [&](int e) { return side[e] == 1; }, list, scan);
            nn_tiles<MAXD>(TB, ncap, size_b, dim, list, n1, samp, nn, dsq, nnl);
        }
        // ---- phase 1c: dynamic-domain skip and steer
        for (int e = tid; e < S; e += nt) {
            const int t = side[e];
            const float* T = t ? TB : TA;
            const float* R = t ? RB : RA;
            const float nd = __builtin_sqrtf(dsq[e]);
            if (a.dd && R[nn[e]] < nd) {
                state[e] = 0;
            } else {
                const float s = (nd > 1e-8f) ? fminf(1.0f, a.step_size / nd) : 1.0f;
                for (int d = 0; d < dim; ++d) {
                    const float base = T[static_cast<size_t>(d) * ncap + nn[e]];
                    cand[e * dim + d] = base + (samp[e * dim + d] - base) * s;
                }
                state[e] = 1;
            }
            fail[e] = 0;
        }
        __syncthreads();
        // ---- phase 2: extend edges, waypoint t_k = (k+1)/G, k = 0..G-1.  Each wave takes kFB
        // (explorer, waypoint) items, computes their FK one per lane, then checks them one by one.
        {
            const int nv = compact(S, [&](int e) { return state[e] == 1; }, list, scan);
            const int items = nv * G;
            const int nb = (items + kFB - 1) / kFB;
            for (int b = wid; b < nb; b += nwaves) {
                const int it = b * kFB + lane;
                int e = -1;
                if (lane < kFB && it < items) {
                    const int slot = it / G, k = it - slot * G;
                    e = list[slot];
                    if (lds_load(&fail[e])) {
                        e = -1;
                    } else {
                        const float* T = side[e] ? TB : TA;
                        const float tf = static_cast<float>(k + 1) / static_cast<float>(G);
                        for (int d = 0; d < dim; ++d) {
                            const float base = T[static_cast<size_t>(d) * ncap + nn[e]];
                            L.q[d * kFBP] = base + tf * (cand[e * dim + d] - base);
                        }
                        if (M.enabled) fk_lane(M, L);
                    }
                }
                wave_sync();
                for (int c = 0; c < kFB; ++c) {
                    const int ec = uniform(__shfl(e, c, kWave));
                    if (ec < 0 || uniform(lds_load(&fail[ec]))) continue;
                    if (lane == 0) ++evals;
                    if (wave_in_collision(M, W.T + c, kFBP, flops) && lane == 0) lds_store(&fail[ec], 1);
                }
                wave_sync();
            }
            __syncthreads();
        }
        // ---- phase 3: commit -- appends in explorer order, then radius updates
        {
            for (int e = tid; e < S; e += nt)
                if (state[e] == 1 && fail[e]) state[e] = 3;  // 3: collided
            if (tid == 0) sc[SC_OVERFLOW] = 0;
            __syncthreads();
            int base0 = size_a, base1 = size_b;
            for (int e0 = 0; e0 < S; e0 += nt) {
                const int e = e0 + tid;
                const bool fr = e < S && state[e] == 1;
                const int t = fr ? side[e] : 0;
                int tot0, tot1;
                const int p0 = block_excl_scan((fr && t == 0) ? 1 : 0, scan, &tot0);
                const int p1 = block_excl_scan((fr && t == 1) ? 1 : 0, scan, &tot1);
                if (e < S) {
                    int ni = -1;
                    if (fr) {
                        const int idx = t ? base1 + p1 : base0 + p0;
                        if (idx < N) {
                            float* T = t ? TB : TA;
                            for (int d = 0; d < dim; ++d) T[static_cast<size_t>(d) * ncap + idx] = cand[e * dim + d];
                            (t ? PB : PA)[idx] = nn[e];
                            (t ? RB : RA)[idx] = FLT_MAX;
                            ni = idx;
                        } else {
                            sc[SC_OVERFLOW] = 1;
                        }
                    }
                    newi[e] = ni;
                }
                base0 += tot0;
                base1 += tot1;
            }
            if (a.dd) {
                for (int e = tid; e < S; e += nt) {
                    if (state[e] == 3) hash_add(hk, hs, hg, a.hash_size, nn[e] * 2 + side[e], true);
                    else if (newi[e] >= 0) hash_add(hk, hs, hg, a.hash_size, nn[e] * 2 + side[e], false);
                }
                __syncthreads();
                for (int h = tid; h < a.hash_size; h += nt) {
                    const int key = hk[h];
                    if (key < 0) continue;
                    float* R = (key & 1) ? RB : RA;
                    float rad = R[key >> 1];
                    for (int i = 0; i < hs[h]; ++i) radius_shrink(rad, a.dd_alpha, a.dd_radius, a.dd_min_radius);
                    for (int i = 0; i < hg[h]; ++i) radius_grow(rad, a.dd_alpha);
                    R[key >> 1] = rad;
                    hk[h] = -1; hs[h] = 0; hg[h] = 0;
                }
            }
            __syncthreads();
            if (tid == 0) {
                sc[SC_SIZE_A] = min(base0, N);
                sc[SC_SIZE_B] = min(base1, N);
                if (sc[SC_OVERFLOW]) sc[SC_STATUS] = PRRTC_STATUS_TREE_FULL;
            }
            __syncthreads();
            if (sc[SC_STATUS] != 0) break;
        }
        // ---- phase 4: connect nearest neighbour in the other tree
        {
            const int sa = sc[SC_SIZE_A], sb = sc[SC_SIZE_B];
            const int n0 = compact(S, [&](int e) { return newi[e] >= 0 && side[e] == 1; }, list, scan);
            nn_tiles<MAXD>(TA, ncap, sa, dim, list, n0, cand, onn, dsq, nnl);
            const int n1 = compact(S, [&](int e) { return newi[e] >= 0 && side[e] == 0; }, list, scan);
            nn_tiles<MAXD>(TB, ncap, sb, dim, list, n1, cand, onn, dsq, nnl);
            for (int e = tid; e < S; e += nt) {
                if (newi[e] < 0) continue;
                const float* O = side[e] ? TA : TB;
                const float cd = __builtin_sqrtf(dsq[e]);
                int n_ext = 0;
                if (a.step_size > 1e-8f && cd > 1e-8f) n_ext = static_cast<int>(ceilf(cd / a.step_size));
                next[e] = n_ext;
                fail[e] = n_ext;
                kd[e] = 0;
                for (int d = 0; d < dim; ++d)
                    vec[e * dim + d] = (n_ext > 0)
                        ? (O[static_cast<size_t>(d) * ncap + onn[e]] - cand[e * dim + d]) / static_cast<float>(n_ext)
                        : 0.0f;
            }
            __syncthreads();
        }
        // ---- phase 5: connect sub-edges in passes; fail[e] = first failing sub-edge (or n_ext)
        for (;;) {
            const int na = compact(S, [&](int e) { return newi[e] >= 0 && fail[e] == next[e] && kd[e] < next[e]; },
                                   list, scan);
            if (na == 0) break;
            const int mper = max(1, (nwaves * kFB * kPassMult) / (na * G));
            int run = 0;
            for (int s0 = 0; s0 < na; s0 += nt) {
                const int s = s0 + tid;
                int v = 0;
                if (s < na) {
                    const int e = list[s];
                    v = min(mper, next[e] - kd[e]) * G;
                }
                int tot;
                const int pre = block_excl_scan(v, scan, &tot);
                if (s < na) offs[s] = run + pre;
                run += tot;
            }
            if (tid == 0) offs[na] = run;
            __syncthreads();
            const int nb = (run + kFB - 1) / kFB;
            for (int b = wid; b < nb; b += nwaves) {
                const int it = b * kFB + lane;
                int e = -1, j = 0;
                if (lane < kFB && it < run) {
                    int lo = 0, hi = na;  // offs[lo] <= it < offs[hi]
                    while (hi - lo > 1) {
                        const int mid = (lo + hi) >> 1;
                        if (offs[mid] <= it) lo = mid; else hi = mid;
                    }
                    e = list[lo];
                    const int local = it - offs[lo];
                    j = kd[e] + local / G;
                    const int k = local - (local / G) * G;
                    if (j >= lds_load(&fail[e])) {
                        e = -1;
                    } else {
                        const float tf = static_cast<float>(k + 1) / static_cast<float>(G);
                        for (int d = 0; d < dim; ++d) {
                            float c = cand[e * dim + d];
                            const float v = vec[e * dim + d];
                            for (int jj = 0; jj < j; ++jj) c += v;
                            L.q[d * kFBP] = c + tf * v;
                        }
                        if (M.enabled) fk_lane(M, L);
                    }
                }
                wave_sync();
                for (int c = 0; c < kFB; ++c) {
                    const int ec = uniform(__shfl(e, c, kWave));
                    const int jc = uniform(__shfl(j, c, kWave));
                    if (ec < 0 || jc >= uniform(lds_load(&fail[ec]))) continue;
                    if (lane == 0) ++evals;
                    if (wave_in_collision(M, W.T + c, kFBP, flops) && lane == 0) lds_min(&fail[ec], jc);
                }
                wave_sync();
            }
            __syncthreads();
            for (int s = tid; s < na; s += nt) {
                const int e = list[s];
                kd[e] += min(mper, next[e] - kd[e]);
            }
            __syncthreads();
        }
        // ---- phase 6: connect appends (explorer-major); first success or overflow decides
        {
            if (tid == 0) sc[SC_EVENT] = S;
            __syncthreads();
            int base0 = sc[SC_SIZE_A], base1 = sc[SC_SIZE_B];
            for (int e0 = 0; e0 < S; e0 += nt) {
                const int e = e0 + tid;
                const bool act = e < S && newi[e] >= 0;
                const int t = act ? side[e] : 0;
                const int nok = act ? fail[e] : 0;
                int tot0, tot1;
                const int p0 = block_excl_scan(t == 0 ? nok : 0, scan, &tot0);
                const int p1 = block_excl_scan(t == 1 ? nok : 0, scan, &tot1);
                if (act) {
                    const int b = t ? base1 + p1 : base0 + p0;
                    float* T = t ? TB : TA;
                    int* P = t ? PB : PA;
                    float* R = t ? RB : RA;
                    int parent = newi[e];
                    bool over = false;
                    for (int k = 0; k < nok; ++k) {
                        const int idx = b + k;
                        if (idx >= N) { over = true; break; }
                        for (int d = 0; d < dim; ++d) {
                            float c = cand[e * dim + d];
                            const float v = vec[e * dim + d];
                            for (int jj = 0; jj <= k; ++jj) c += v;
                            T[static_cast<size_t>(d) * ncap + idx] = c;
                        }
                        P[idx] = parent;
                        R[idx] = FLT_MAX;
                        parent = idx;
                    }
                    aux0[e] = over ? 1 : 0;
                    aux1[e] = parent;
                    if (over || nok == next[e]) lds_min(&sc[SC_EVENT], e);
                }
                base0 += tot0;
                base1 += tot1;
            }
            __syncthreads();
            if (tid == 0) {
                sc[SC_SIZE_A] = min(base0, N);
                sc[SC_SIZE_B] = min(base1, N);
                const int e = sc[SC_EVENT];
                if (e < S) {
                    if (aux0[e]) {
                        sc[SC_STATUS] = PRRTC_STATUS_TREE_FULL;
                    } else {
                        sc[SC_STATUS] = PRRTC_STATUS_SOLVED;
                        if (side[e] == 0) { sc[SC_CONN0] = aux1[e]; sc[SC_CONN1] = onn[e]; }
                        else { sc[SC_CONN0] = onn[e]; sc[SC_CONN1] = aux1[e]; }
                        sc[SC_CONN2] = side[e];
                    }
                }
            }
            __syncthreads();
            if (sc[SC_STATUS] != 0) break;
        }
    }
    // ---- outputs
    if (lane == 0 && evals) {
        global_add(&a.stats[0], evals);
        global_add(&a.stats[2], flops);
    }
    const int st = sc[SC_STATUS];
    if (tid == 0) {
        const unsigned long long t1 = wall_clock64();
        a.sizes[2 * q] = sc[SC_SIZE_A];
        a.sizes[2 * q + 1] = sc[SC_SIZE_B];
        a.completed[2 * q] = sc[SC_SIZE_A];
        a.completed[2 * q + 1] = sc[SC_SIZE_B];
        a.iters[q] = r;
        a.conn[3 * q] = sc[SC_CONN0];
        a.conn[3 * q + 1] = sc[SC_CONN1];
        a.conn[3 * q + 2] = sc[SC_CONN2];
        a.solved[q] = (st == PRRTC_STATUS_SOLVED) ? 1 : -1;
        a.time_ms[q] = static_cast<float>(static_cast<double>(t1 - t0) / a.wall_khz);
        if (a.status) a.status[q] = st;
        global_add(&a.stats[1], static_cast<unsigned long long>(r));
        if (a.path) {
            float* out = a.path + static_cast<size_t>(q) * a.path_cap * dim;
            int len = 0;
            if (st == PRRTC_STATUS_SOLVED) {
                int na = 0, cur = sc[SC_CONN0];
                for (int it = 0; it < ncap; ++it) {
                    ++na;
                    const int p = PA[cur];
                    if (p == cur) break;
                    cur = p;
                }
                int nb = 0;
                cur = sc[SC_CONN1];
                for (int it = 0; it < ncap; ++it) {
                    ++nb;
                    const int p = PB[cur];
                    if (p == cur) break;
                    cur = p;
                }
                if (na + nb > a.path_cap) {
                    len = -1;
                } else {
                    cur = sc[SC_CONN0];
                    for (int k = na - 1; k >= 0; --k) {
                        for (int d = 0; d < dim; ++d) out[k * dim + d] = TA[static_cast<size_t>(d) * ncap + cur];
                        cur = PA[cur];
                    }
                    cur = sc[SC_CONN1];
                    for (int k = 0; k < nb; ++k) {
                        for (int d = 0; d < dim; ++d) out[(na + k) * dim + d] = TB[static_cast<size_t>(d) * ncap + cur];
                        cur = PB[cur];
                    }
                    len = na + nb;
                }
            }
            a.path_len[q] = len;
        }
    }
    __syncthreads();
}

template <int MAXD>
__global__ __launch_bounds__(512) void plan_kernel(PlanArgs a) {
    extern __shared__ __align__(16) char lds[];
    int* sc = reinterpret_cast<int*>(lds + a.off_misc) + 20;
    for (;;) {
        if (threadIdx.x == 0) sc[SC_QUERY] = global_add(a.queue, 1);
        __syncthreads();
        const int qi = sc[SC_QUERY];
        __syncthreads();
        if (qi >= a.B) return;
        const int q = a.order ? a.order[qi] : qi;
        plan_query<MAXD>(a, q, lds);
    }
}

__global__ void reset_kernel(int* queue, unsigned long long* stats) {
    if (threadIdx.x == 0) *queue = 0;
    if (threadIdx.x < 4) stats[threadIdx.x] = 0ull;
}

// ---------------------------------------------------------------------------------- host side
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

struct WsLayout {
    size_t radii, xbuf, queue, stats, total;
};

WsLayout ws_layout(int B, int dim, int ncap, int S) {
    WsLayout w;
    size_t o = 0;
    w.radii = o; o = align16(o + sizeof(float) * 2 * static_cast<size_t>(ncap) * B);
    w.xbuf = o; o = align16(o + sizeof(float) * 3 * static_cast<size_t>(S) * dim * B);
    w.queue = o; o = align16(o + 16);
    w.stats = o; o = align16(o + 4 * sizeof(unsigned long long));
    w.total = o + 256;
    return w;
}

size_t planner_lds(const ModelLayout& ml, int S, size_t* off_expl, size_t* off_primes, size_t* off_nn,
                   size_t* off_hash, size_t* off_misc, int* hsize) {
    size_t o = ml.end;
    *off_expl = o; o = align16(o + expl_bytes(S));
    *off_primes = o; o = align16(o + 16 * static_cast<size_t>(S));
    *off_nn = o; o = align16(o + nn_bytes());
    *hsize = hash_size_for(S);
    *off_hash = o; o = align16(o + 3 * sizeof(int) * static_cast<size_t>(*hsize));
    *off_misc = o; o = align16(o + sizeof(int) * (20 + SC_COUNT));
    return o;
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
    d.counts = per_query ? w.counts : nullptr;
    return d;
}

int check_robot(const prrtc_robot* r) {
    if (!r) return fail_with(PRRTC_E_INVALID, "robot is null");
    if (r->n_joints < 0 || r->n_joints > PRRTC_MAX_JOINTS)
        return fail_with(PRRTC_E_INVALID, "n_joints must be in [0, 64]");
    if (r->n_robot_spheres < 0) return fail_with(PRRTC_E_INVALID, "n_robot_spheres < 0");
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
    return PRRTC_OK;
}

template <int MAXD>
hipError_t launch_plan(const PlanArgs& a, int grid, int nt, size_t lds, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plan_kernel<MAXD>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(plan_kernel<MAXD>, dim3(grid), dim3(nt), lds, s, a);
    return hipGetLastError();
}

template <int MAXD>
int occupancy(int nt, size_t lds) {
    int blocks = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&blocks, plan_kernel<MAXD>, nt, lds) != hipSuccess) return 1;
    return blocks > 0 ? blocks : 1;
}

int planner_impl(hipStream_t stream, int B, const float* starts, const float* goals, int gmax, const int* goal_counts,
                 const float* min_vals, const float* max_vals, const prrtc_robot* robot, const prrtc_world* world,
                 int world_per_query, const int* max_nodes_q, const unsigned* seeds, const int* order,
                 const prrtc_attrs* at, const prrtc_outputs* out, const prrtc_launch* launch, void* workspace,
                 size_t workspace_bytes) {
    if (!at || !out) return fail_with(PRRTC_E_INVALID, "attrs/outputs are null");
    if (B < 1) return fail_with(PRRTC_E_INVALID, "batch must be >= 1");
    if (at->dim < 1 || at->dim > PRRTC_MAX_DIM) return fail_with(PRRTC_E_INVALID, "dim must be in [1, 16]");
    if (at->num_new_samples < 1 || at->num_new_samples > PRRTC_MAX_SAMPLES)
        return fail_with(PRRTC_E_INVALID, "num_new_samples must be in [1, 1024]");
    if (at->granularity < 1 || at->granularity > PRRTC_MAX_GRANULARITY)
        return fail_with(PRRTC_E_INVALID, "granularity must be in [1, 64]");
    if (at->max_nodes < 1 || at->max_nodes > (1 << 29)) return fail_with(PRRTC_E_INVALID, "max_nodes out of range");
    if (!(at->step_size > 0.0f)) return fail_with(PRRTC_E_INVALID, "step_size must be > 0");
    if (gmax < 1 || gmax > at->max_nodes) return fail_with(PRRTC_E_INVALID, "num_goals must be in [1, max_nodes]");
    if (at->collision_mode < 0 || at->collision_mode > 3) return fail_with(PRRTC_E_INVALID, "collision_mode must be 0..3");
    if (!starts || !goals || !min_vals || !max_vals) return fail_with(PRRTC_E_INVALID, "start/goal/limit pointer is null");
    if (!out->tree_a_configs || !out->tree_b_configs || !out->tree_a_parents || !out->tree_b_parents ||
        !out->tree_sizes || !out->completed || !out->iter_count || !out->connection_info || !out->solved_flag ||
        !out->kernel_time_ms)
        return fail_with(PRRTC_E_INVALID, "one of the ten reference outputs is null");
    if (out->path && (!out->path_len || out->path_capacity < 1))
        return fail_with(PRRTC_E_INVALID, "path needs path_len and path_capacity >= 1");
    int rc = check_robot(robot);
    if (rc) return rc;
    rc = check_world(world);
    if (rc) return rc;
    if (world_per_query && !world->counts) return fail_with(PRRTC_E_INVALID, "per-query world needs counts");
    const int dim = at->dim, S = at->num_new_samples, ncap = at->max_nodes;
    const WsLayout wl = ws_layout(B, dim, ncap, S);
    if (!workspace || workspace_bytes < wl.total) return fail_with(PRRTC_E_WORKSPACE, "workspace too small");

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_with(PRRTC_E_HIP, "hipGetDevice failed");
    const DevInfo& di = dev_info(dev);
    int nt = (launch && launch->threads > 0) ? launch->threads : 256;
    if (nt % kWave || nt > 512) return fail_with(PRRTC_E_INVALID, "planner threads must be a multiple of 64, <= 512");

    PlanArgs a;
    std::memset(&a, 0, sizeof(a));
    a.max_iterations = at->max_iterations; a.step_size = at->step_size; a.S = S; a.balance_mode = at->balance_mode;
    a.tree_ratio = at->tree_ratio; a.dd = at->dynamic_domain ? 1 : 0; a.dd_alpha = at->dd_alpha;
    a.dd_radius = at->dd_radius; a.dd_min_radius = at->dd_min_radius; a.dim = dim; a.ncap = ncap;
    a.G = at->granularity; a.mode = (at->collision_mode == 1) ? 1 : 0; a.margin = at->collision_margin; a.B = B;
    a.robot = robot_dev(*robot);
    a.world = world_dev(*world, world_per_query != 0);
    a.starts = starts; a.goals = goals; a.gmax = gmax; a.goal_counts = goal_counts;
    a.min_vals = min_vals; a.max_vals = max_vals; a.max_nodes_q = max_nodes_q; a.seeds = seeds; a.order = order;
    a.tree_a = out->tree_a_configs; a.tree_b = out->tree_b_configs; a.par_a = out->tree_a_parents;
    a.par_b = out->tree_b_parents; a.sizes = out->tree_sizes; a.completed = out->completed; a.iters = out->iter_count;
    a.conn = out->connection_info; a.solved = out->solved_flag; a.time_ms = out->kernel_time_ms; a.status = out->status;
    a.path = out->path; a.path_len = out->path_len; a.path_cap = out->path_capacity;
    char* ws = static_cast<char*>(workspace);
    a.radii = reinterpret_cast<float*>(ws + wl.radii);
    a.xbuf = reinterpret_cast<float*>(ws + wl.xbuf);
    a.queue = reinterpret_cast<int*>(ws + wl.queue);
    a.stats = reinterpret_cast<unsigned long long*>(ws + wl.stats);
    a.wall_khz = di.wall_khz;

    size_t lds = 0;
    for (;;) {
        a.ml = model_layout(robot->n_joints, robot->n_robot_spheres, world->n_self_pairs, world->n_world_spheres,
                            world->n_world_capsules, world->n_world_boxes, world->n_world_halfspaces, nt / kWave, kFBP);
        lds = planner_lds(a.ml, S, &a.off_expl, &a.off_primes, &a.off_nn, &a.off_hash, &a.off_misc, &a.hash_size);
        if (lds <= static_cast<size_t>(di.max_lds) || nt <= kWave) break;
        nt -= kWave;
    }
    if (lds > static_cast<size_t>(di.max_lds))
        return fail_with(PRRTC_E_LDS, "robot + world + explorer state exceed the work-group LDS (" +
                                           std::to_string(lds) + " > " + std::to_string(di.max_lds) + " bytes)");
    const bool small = dim <= 8;
    int slots = (launch && launch->slots > 0) ? launch->slots
                                               : (small ? occupancy<8>(nt, lds) : occupancy<16>(nt, lds)) * di.cus;
    slots = std::max(1, std::min(slots, B));
    hipLaunchKernelGGL(reset_kernel, dim3(1), dim3(64), 0, stream, a.queue, a.stats);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    e = small ? launch_plan<8>(a, slots, nt, lds, stream) : launch_plan<16>(a, slots, nt, lds, stream);
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    return PRRTC_OK;
}

}  // namespace prrtc

using namespace prrtc;

extern "C" {

size_t prrtc_planner_workspace_bytes(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples) {
    if (batch < 1 || dim < 1 || max_nodes < 1 || num_new_samples < 1) return 0;
    return ws_layout(batch, dim, max_nodes, num_new_samples).total;
}

int prrtc_planner(prrtc_stream_t stream, const float* start, const float* goals, int32_t num_goals,
                  const float* min_vals, const float* max_vals, const prrtc_robot* robot, const prrtc_world* world,
                  const prrtc_attrs* attrs, const prrtc_outputs* out, const prrtc_launch* launch, void* workspace,
                  size_t workspace_bytes) {
    return planner_impl(reinterpret_cast<hipStream_t>(stream), 1, start, goals, num_goals, nullptr, min_vals, max_vals,
                        robot, world, 0, nullptr, nullptr, nullptr, attrs, out, launch, workspace, workspace_bytes);
}

int prrtc_planner_batch(prrtc_stream_t stream, int32_t batch, const float* starts, const float* goals,
                        int32_t max_goals, const int32_t* goal_counts, const float* min_vals, const float* max_vals,
                        const prrtc_robot* robot, const prrtc_world* world, int32_t world_per_query,
                        const int32_t* max_nodes_q, const uint32_t* seeds, const int32_t* order,
                        const prrtc_attrs* attrs, const prrtc_outputs* out, const prrtc_launch* launch,
                        void* workspace, size_t workspace_bytes) {
    return planner_impl(reinterpret_cast<hipStream_t>(stream), batch, starts, goals, max_goals, goal_counts, min_vals,
                        max_vals, robot, world, world_per_query, max_nodes_q, seeds, order, attrs, out, launch,
                        workspace, workspace_bytes);
}

size_t prrtc_planner_stats_offset(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples) {
    if (batch < 1 || dim < 1 || max_nodes < 1 || num_new_samples < 1) return 0;
    return ws_layout(batch, dim, max_nodes, num_new_samples).stats;
}

const char* prrtc_last_error(void) { return g_err.c_str(); }

const char* prrtc_version(void) { return "prrtc_amd 0.1.0 (gfx950)"; }

}  // extern "C"

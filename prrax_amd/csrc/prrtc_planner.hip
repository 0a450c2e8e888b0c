// prrtc_planner.hip -- lockstep pRRTC planner for MI355X (gfx950) + its C ABI.
//
// Each query is planned by a team of W work-groups.  The reference's explorers (its blocks,
// prrtc_planner.cu:657-674) advance in lockstep rounds:
//
//   1  sample (Halton, prime shuffle), pick a tree, nearest neighbour, dynamic-domain skip, steer
//                                                         prrtc_planner.cu:783-881
//   2  FK + collision of every extend edge (G waypoints each)  prrtc_planner.cu:883-909
//   3  commit: all radius shrinks, then appends in explorer order with radius grows
//                                                         prrtc_planner.cu:911-947, 46-83
//   4  connect nearest neighbour of each successful extend in the other tree as it stood at the
//      round's start, step count and step vector          prrtc_planner.cu:949-994
//   5  connect sub-edges, checked in parallel work units; each explorer's first failing sub-edge
//      is an order-independent minimum
//                                                         prrtc_planner.cu:996-1018
//   6  connect appends in explorer order; the first success or overflow decides the query
//                                                         prrtc_planner.cu:1020-1062
//
// Each round is one legal interleaving of the reference kernel, chosen so that results do not
// depend on timing: the trees equal the CPU oracle's bit for bit, whatever the team size.
// Shared state changes only in the commit phases, through prefix sums over explorers (no
// per-node atomics), which also removes the reference's unordered-append race.
//
// Scheduling (two launches on one stream, no host synchronisation in between):
//   pass 1  one work-group per query, queries pulled from a device queue; a query still open
//           after `pass1_rounds` rounds (or, once every query has started, after
//           kPass1MinRounds) saves its round state and joins the continue list;
//   pass 2  a cooperative launch: the continue list is split over teams of W work-groups
//           (W = slots / open queries, capped), so the few long queries of an uneven batch
//           get the whole GPU.  A single query skips pass 1 (budget 0).
// Inside a team, explorer e belongs to member e % W; work whose result every member needs
// (prefix sums, sizes, status) is computed redundantly from global state after a team barrier.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <mutex>
#include <string>

#include "prrtc_amd.h"
#include "prrtc_common.h"
#include "prrtc_host.h"

namespace prrtc {

// Per-query explorer arrays kept in global memory ([B][XG_COUNT][S] ints).  XG_KEY / XG_OKEY are
// 64-bit nearest-neighbour keys (two int slots each, first so that they stay 8-byte aligned;
// XG_COUNT is even).
enum { XG_KEY = 0, XG_OKEY = 2, XG_FAIL = 4, XG_VAMP, XG_ECLAIM, XG_CLAIM, XG_CFAIL, XG_PAD, XG_COUNT };
static_assert(XG_COUNT % 2 == 0, "keys of every query must stay 8-byte aligned");
// Per-query round state ([B][QS_COUNT] ints) for the pass-1 -> pass-2 hand-over.
enum { QS_ROUND, QS_SIZE_A, QS_SIZE_B, QS_TIME, QS_STOP, QS_COUNT = 8 };
// Control words ([kCtl] ints).  CTL_NC + i counts continue list i (three lists rotate over the
// epochs of pass 2: one read, one filled, one cleared).
enum { CTL_QUEUE, CTL_ERROR, CTL_NC, kCtl = 16 };
// uint64 statistics at the end of the workspace (layout in prrtc_amd.h, prrtc_planner_stats_offset).
constexpr int kStats = 32;

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
    // workspace
    float* radii;                // [B][2][ncap]
    int* xg;                     // [B][XG_COUNT][S]
    int* qstate;                 // [B][QS_COUNT]
    int* ctl;                    // [kCtl]
    int* cont;                   // [3][B] continue lists
    unsigned* tsync;             // team barrier lines (make_team); the last team is the grid
    unsigned long long* stats;   // [kStats]
    // scheduling
    int pass1_rounds;            // round budget of pass 1 (0: pass 1 only initialises)
    int wmax;                    // team size cap
    int epoch_rounds;            // round budget of a query per pass-2 epoch (teams re-form between epochs)
    int slots;
    // LDS layout
    ModelLayout ml;
    size_t off_expl, off_cand, off_primes, off_sampler, off_nn, off_hash, off_misc;
    int hash_size;
    float wall_khz;
};

// Explorer arrays in each member's LDS (S + 1 ints each).
enum { X_SIDE, X_NN, X_STATE, X_NEWI, X_ONN, X_NEXT, X_NOK, X_VAMP, X_LIST, X_AUX0, X_AUX1, X_COUNT };

__host__ __device__ inline size_t expl_bytes(int S) { return align16(sizeof(int) * (size_t)(S + 1) * X_COUNT); }
__host__ __device__ inline int hash_size_for(int S) {
    int h = 64;
    while (h < 2 * S) h <<= 1;
    return h;
}

enum { SC_QUERY = 0, SC_SIZE_A, SC_SIZE_B, SC_STATUS, SC_EVENT, SC_OVERFLOW, SC_CONN0, SC_CONN1, SC_CONN2, SC_ERR,
       SC_DRAIN, SC_STOP, SC_COUNT = 16 };

// Pass 1 gives a query at least this many rounds; after that, once no query is left to start,
// the query moves on to pass 2 (teams) instead of spending the rest of its budget on one
// work-group while the others idle.  (C3 steps: 1 / 2 / 4 / 8 / 12 / 16 / 24 rounds 259.5 /
// 256.1 / 253.3 / 251.3 / 251.4 / 251.4 / 260.0 ms; 256 dual-arm queries at S = 128, all started
// at once: 4 / 6 / 8 / 12 / never 51.0 / 51.0 / 50.7 / 59.0 / 65.4 ms.)
constexpr int kPass1MinRounds = 8;

// ---------------------------------------------------------------------------------- sampler
// Explorer e's order of the 16 Halton bases {3, 5, ..., 59}, as indices into that list.
__device__ __forceinline__ void prime_order(int e, unsigned seed, unsigned char* p) {
#pragma unroll
    for (int k = 0; k < 16; ++k) p[k] = static_cast<unsigned char>(k);
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

// The Halton radical inverse of index n in base b (prrtc_planner.cu:144-157): fraction = 1;
// while n > 0 { fraction /= b; result += fraction * (n % b); n /= b }.  Per-work-group tables
// in LDS make it a short loop (sampler_bytes() at one 16-byte aligned address):
//   frac[16][kHaltonDigits]  frac[k][j] is the fraction that multiplies digit j in base k,
//                            b^-(j+1) by the same repeated correctly rounded division, so the
//                            products and sums of sample_dim are bit-identical to the loop's;
//   magic[16] (u64)          floor(2^64 / b) + 1: q = high 64 bits of n * magic is n / b exactly
//                            for every 32-bit n (the error n * (magic * b - 2^64) / b < 1 / b);
//   base[16], lo[kMaxDim], span[kMaxDim] = hi - lo;
//   hv[16]                   this round's radical inverse in each base (halton_round): the
//                            index r is the round, common to all explorers, so a round needs
//                            16 radical inverses, not one per explorer and dimension.
constexpr int kHaltonDigits = 20;  // 3^20 > 2^31: every digit of a non-negative int index
constexpr int kSmpMagic = 16 * kHaltonDigits * 4, kSmpBase = kSmpMagic + 16 * 8, kSmpLo = kSmpBase + 16 * 4,
              kSmpSpan = kSmpLo + kMaxDim * 4, kSmpHv = kSmpSpan + kMaxDim * 4;
__host__ __device__ inline size_t sampler_bytes() { return align16(kSmpHv + 16 * 4); }

// Fills the tables (every thread calls it; a __syncthreads must follow before use).  The tables
// depend on the batch's limits only: each planner kernel fills them once, in its prologue.
__device__ __forceinline__ void sampler_init(char* tab, const float* lo, const float* hi, int dim) {
    const int tid = threadIdx.x;
    if (tid < 16) {
        const unsigned char P[16] = {3, 5, 7, 11, 13, 17, 19, 23, 29, 31, 37, 41, 43, 47, 53, 59};
        unsigned b = 3;
#pragma unroll
        for (int k = 0; k < 16; ++k)
            if (tid == k) b = P[k];
        float f = 1.0f;
        float* F = reinterpret_cast<float*>(tab) + tid * kHaltonDigits;
        for (int j = 0; j < kHaltonDigits; ++j) {
            f /= static_cast<float>(b);
            F[j] = f;
        }
        reinterpret_cast<unsigned long long*>(tab + kSmpMagic)[tid] = ~0ull / b + 1ull;  // b odd
        reinterpret_cast<unsigned*>(tab + kSmpBase)[tid] = b;
    }
    if (tid < dim) {
        reinterpret_cast<float*>(tab + kSmpLo)[tid] = lo[tid];
        reinterpret_cast<float*>(tab + kSmpSpan)[tid] = hi[tid] - lo[tid];
    }
}

// This round's radical inverses hv[k] of r in base k (threads k < 16; a __syncthreads must
// follow before sample_dim reads them).
__device__ __forceinline__ void halton_round(char* tab, int r) {
    const int k = threadIdx.x;
    if (k >= 16) return;
    const unsigned long long mg = reinterpret_cast<const unsigned long long*>(tab + kSmpMagic)[k];
    const unsigned b = reinterpret_cast<const unsigned*>(tab + kSmpBase)[k];
    const float* F = reinterpret_cast<const float*>(tab) + k * kHaltonDigits;
    const unsigned long long mlo = mg & 0xffffffffull, mhi = mg >> 32;
    float result = 0.0f;
    unsigned n = static_cast<unsigned>(r);
    for (int j = 0; n > 0; ++j) {
        const unsigned long long t = ((static_cast<unsigned long long>(n) * mlo) >> 32) + static_cast<unsigned long long>(n) * mhi;
        const unsigned q = static_cast<unsigned>(t >> 32);
        result += F[j] * static_cast<float>(n - q * b);
        n = q;
    }
    reinterpret_cast<float*>(tab + kSmpHv)[k] = result;
}

// Explorer e's sample of this round in dimension d (prrtc_planner.cu:783-788); pe is the
// explorer's base order from prime_order: lo[d] + radical_inverse(r, base) * (hi[d] - lo[d]).
__device__ __forceinline__ float sample_dim(const char* tab, const unsigned char* pe, int d) {
    const float* f = reinterpret_cast<const float*>(tab);
    return f[kSmpLo / 4 + d] + f[kSmpHv / 4 + pe[d & 15]] * f[kSmpSpan / 4 + d];
}

// ---------------------------------------------------------------------------------- helpers
// Explorers e < S with pred(e), in increasing order, into list; returns their count.  Wave 0
// does it alone, one ballot per 64 explorers (no block-wide scan), and publishes the count in
// scan[slot] behind one work-group barrier.  Two calls in a row must use different slots (a
// wave may still be reading the first count when wave 0 writes the second).
template <class F>
__device__ __forceinline__ int compact(int S, F pred, int* list, int* scan, int slot = 19) {
    const int tid = threadIdx.x;
    if (tid < kWave) {
        int base = 0;
        for (int e0 = 0; e0 < S; e0 += kWave) {
            const int e = e0 + tid;
            const bool ok = e < S && pred(e);
            const unsigned long long m = __ballot(ok);
            if (ok) list[base + static_cast<int>(__popcll(m & ((1ull << tid) - 1ull)))] = e;
            base += static_cast<int>(__popcll(m));
        }
        if (tid == 0) scan[slot] = base;
    }
    __syncthreads();
    return scan[slot];
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

// Global words shared by a team: relaxed agent-scope atomics through address-space-1 pointers.
__device__ __forceinline__ int gload(const int* p) {
    return __hip_atomic_load((const PRRTC_GLOBAL int*)(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gstore(int* p, int v) {
    __hip_atomic_store((PRRTC_GLOBAL int*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gmin(int* p, int v) {
    __hip_atomic_fetch_min((PRRTC_GLOBAL int*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ unsigned long long gload64(const unsigned long long* p) {
    return __hip_atomic_load((const PRRTC_GLOBAL unsigned long long*)(p), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gstore64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_store((PRRTC_GLOBAL unsigned long long*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void gmin64(unsigned long long* p, unsigned long long v) {
    __hip_atomic_fetch_min((PRRTC_GLOBAL unsigned long long*)(p), v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
// Nearest-neighbour key: squared distances are >= 0, so their bit patterns order like the values
// and the key order is the (distance, lowest index) order.
__device__ __forceinline__ unsigned long long nn_key(float dsq, int idx) {
    return (static_cast<unsigned long long>(__float_as_uint(dsq)) << 32) | static_cast<unsigned>(idx);
}
constexpr unsigned long long kNoKey = ~0ull;

// Waypoints per work unit of a check phase over `edges` edges of G waypoints (a unit is ku
// consecutive waypoints of one edge, checked in order by one wave after one FK of all of them;
// a collision skips the rest of the unit): the smallest power of two ku <= fb whose units are no
// more than the cap below; fb when even fb-waypoint units exceed it.  Uniform over the team.
__device__ __forceinline__ int unit_size(int edges, int G, int waves, int fb) {
    // Large teams (the latency-bound tail of a batch): units fit the waves, one each.  Small
    // teams and single work-groups (many queries at once, throughput-bound): up to two units per
    // wave, which balances better than one larger unit.
    const long long cap = (waves <= 256) ? 2ll * waves : static_cast<long long>(waves);
    int ku = 1;
    while (ku < fb && static_cast<long long>(edges) * ((G + ku - 1) / ku) > cap) ku <<= 1;
    return min(ku, fb);
}

// Team barrier words.  Every counter has a 256-byte line of its own, so arrivals on one counter
// do not queue behind another's, and the generation word that waiting work-groups poll is apart
// from the counters.  Team t (t = slots: the grid) owns lines 2t (arrivals) and 2t + 1
// (generation); a team of W >= kGroupMinW work-groups also owns up to kSyncGroups group lines
// in a second area (make_team).
constexpr int kSyncLine = 64;     // words per line
constexpr int kSyncGroups = 16;   // group counters of a large team
constexpr int kGroupMinW = 128;   // smallest team that arrives in groups
__host__ __device__ inline int sync_group_count(int W) {
    return W >= 2 * kGroupMinW ? kSyncGroups : (W >= kGroupMinW ? kSyncGroups / 2 : 1);
}
// Members of group g (members m with m % G == g): W / G, one more for the first W % G groups.
__device__ __forceinline__ unsigned group_size(int W, int G, int g) {
    return static_cast<unsigned>(W / G + (g < W % G ? 1 : 0));
}

// Team of W work-groups working on one query.
struct Team {
    int W, m;
    unsigned* sync;  // line pair: arrivals (top counter), generation (sync + kSyncLine)
    unsigned* grp;   // group counter lines (G > 1)
    int G;           // 1, or the groups members arrive in (member m in group m % G)
    int* error;
    unsigned gen;
    bool released;   // thread 0: its arrival was the last one (team_arrive), no wait needed
};

// Waits until every global access this thread issued has completed; also a compiler barrier.
__device__ __forceinline__ void drain_memory() { asm volatile("s_waitcnt vmcnt(0) lgkmcnt(0)" ::: "memory"); }

// Barrier over the team's work-groups (a plain work-group barrier when W == 1).  Everything the
// members share goes through agent-scope atomics (ldc/stc<true>, gload/gstore), so it suffices
// that each thread's accesses have completed before its work-group counts in: relaxed counter
// and generation words, no L2 write-back or invalidate (a team may span XCDs).  A large team
// arrives in G groups: the last arrival of a group counts in at the top counter, so no counter
// sees more than max(ceil(W / G), G) arrivals (atomics on one address are serialised).  The last
// arrival at the top opens the next generation.  A counter of n arrivals per barrier is not reset
// between barriers: arrival k * n (mod n == 0) is the last one of a barrier, since no member can
// arrive at the next barrier before the generation opens, after every arrival of this one.  The
// team's last barrier (`reset`, before its words go to another team shape) resets each counter
// and waits for that store before counting in above it or opening the generation.  Waiting is
// bounded (seconds), so a broken team raises the error flag instead of hanging.
__device__ __forceinline__ bool count_in(unsigned* cnt, unsigned n, bool reset) {
    const unsigned prev = __hip_atomic_fetch_add((PRRTC_GLOBAL unsigned*)cnt, 1u, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT);
    if ((prev + 1u) % n != 0u) return false;
    if (reset) {
        __hip_atomic_store((PRRTC_GLOBAL unsigned*)cnt, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        drain_memory();
    }
    return true;
}
// A barrier in two halves: team_arrive counts the work-group in (and opens the next generation
// if it is the last arrival); team_wait waits for that generation.  Work between the two must
// not touch memory the other members read or write before the barrier (for instance the next
// round's samples, which live in the work-group's LDS).
__device__ __forceinline__ void team_arrive(Team& t, bool reset = false) {
    if (t.W > 1) {
        drain_memory();
        __syncthreads();
        if (threadIdx.x == 0) {
            unsigned* gen = t.sync + kSyncLine;
            bool last = true;
            if (t.G > 1) last = count_in(t.grp + (t.m % t.G) * kSyncLine, group_size(t.W, t.G, t.m % t.G), reset);
            if (last) last = count_in(t.sync, static_cast<unsigned>(t.G > 1 ? t.G : t.W), reset);
            if (last) __hip_atomic_fetch_add((PRRTC_GLOBAL unsigned*)gen, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            t.released = last;
        }
    } else {
        __syncthreads();
    }
}
__device__ __forceinline__ void team_wait(Team& t) {
    if (t.W > 1) {
        if (threadIdx.x == 0) {
            unsigned* gen = t.sync + kSyncLine;
            if (!t.released) {
                unsigned spins = 0;
                while (__hip_atomic_load((PRRTC_GLOBAL unsigned*)gen, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) ==
                       t.gen) {
                    __builtin_amdgcn_s_sleep(1);
                    if (++spins > (1u << 26)) {
                        gstore(t.error, 1);
                        break;
                    }
                    // Another work-group gave up on a barrier: the run has failed, stop waiting.
                    if ((spins & 1023u) == 0u && gload(t.error)) break;
                }
            }
            drain_memory();
        }
        t.gen += 1u;
        __syncthreads();
    } else {
        __syncthreads();
    }
}

__device__ __forceinline__ void team_barrier(Team& t, bool reset = false) {
    team_arrive(t, reset);
    team_wait(t);
}

// Barrier over the whole grid between epochs: as team_barrier, plus agent-scope release/acquire
// fences (L2 write-back and invalidate), so plain stores of one epoch (continue lists, saved
// round state) are visible to the work-groups of the next.
__device__ __forceinline__ void grid_barrier(Team& t) {
    drain_memory();  // the write-back below must see every wave's stores
    __syncthreads();
    if (threadIdx.x == 0) __threadfence();
    team_barrier(t);
    if (threadIdx.x == 0) __threadfence();
    __syncthreads();
}

// ---------------------------------------------------------------------------------- one query
struct QueryView {
    int q, N, ng;
    unsigned seed;
    float *TA, *TB, *RA, *RB;
    int *PA, *PB;
    unsigned long long *g_key, *g_okey;  // (squared distance bits << 32 | node) of the extend / connect NN
    int *g_fail, *g_vamp;
    int *g_eclaim, *g_claim, *g_cfail;  // extend / connect claim counters, first failing sub-edge
    int* qs;
};

template <int MAXD, bool COH>
__device__ void run_query(const PlanArgs& a, const QueryView& V, char* lds, Team& team, int r_begin, int r_budget,
                          unsigned long long t0, int pass, int* push_cnt, int* push_list,
                          unsigned long long deadline = 0ull) {
    const int tid = threadIdx.x, nt = blockDim.x;
    const int lane = tid & (kWave - 1), wid = tid / kWave, nwaves = nt / kWave;
    const int dim = a.dim, S = a.S, G = a.G, ncap = a.ncap, N = V.N, W = team.W, m = team.m;
    int* X = reinterpret_cast<int*>(lds + a.off_expl);
    int* side = X + X_SIDE * (S + 1);
    int* nn = X + X_NN * (S + 1);
    int* state = X + X_STATE * (S + 1);
    int* newi = X + X_NEWI * (S + 1);
    int* onn = X + X_ONN * (S + 1);
    int* next = X + X_NEXT * (S + 1);
    int* nok = X + X_NOK * (S + 1);
    int* vamp = X + X_VAMP * (S + 1);
    int* list = X + X_LIST * (S + 1);
    int* aux0 = X + X_AUX0 * (S + 1);
    int* aux1 = X + X_AUX1 * (S + 1);
    float* cand = reinterpret_cast<float*>(lds + a.off_cand);
    float* vec = cand + S * dim;
    unsigned char* primes = reinterpret_cast<unsigned char*>(lds + a.off_primes);
    char* nnl = lds + a.off_nn;
    float* qsh = reinterpret_cast<float*>(nnl);
    float* red_d = qsh + kEC * kMaxDim;
    int* red_i = reinterpret_cast<int*>(red_d + kEC * 16);
    int* hk = reinterpret_cast<int*>(lds + a.off_hash);
    int* hs = hk + a.hash_size;
    int* hg = hs + a.hash_size;
    int* scan = reinterpret_cast<int*>(lds + a.off_misc);
    int* sc = scan + 20;

    Model M;
    stage_model(M, lds, a.ml, a.robot, a.world, V.q, a.dim, a.mode, a.margin, scan);
    const WaveArea WA = wave_area(lds, a.ml, M.n_joints);
    const int fbp = a.ml.fbp, fb = fbp - 1;  // FK batch: fb configurations per wave, stride fbp
    LaneScratch L;  // FK slot of this lane (lanes < fb) in the wave's FK area
    L.q = WA.q + lane;
    L.T = WA.T + lane;
    L.nt = fbp;
    float* TA = V.TA;
    float* TB = V.TB;
    float* RA = V.RA;
    float* RB = V.RB;
    int* PA = V.PA;
    int* PB = V.PB;

    for (int e = tid; e < S; e += nt) {
        prime_order(e, V.seed, primes + 16 * e);
        vamp[e] = V.g_vamp[e];
    }
    for (int h = tid; h < a.hash_size; h += nt) { hk[h] = -1; hs[h] = 0; hg[h] = 0; }
    if (tid == 0) {
        sc[SC_SIZE_A] = V.qs[QS_SIZE_A];
        sc[SC_SIZE_B] = V.qs[QS_SIZE_B];
        sc[SC_STATUS] = 0;
        sc[SC_CONN0] = -1; sc[SC_CONN1] = -1; sc[SC_CONN2] = -1;
        sc[SC_ERR] = 0;
    }
    __syncthreads();

    // Nodes that existed when this run began were written before the last grid-wide fence (or
    // kernel boundary) and never change: NN scans read them through the caches.
    const int coh_a = sc[SC_SIZE_A], coh_b = sc[SC_SIZE_B];
    unsigned long long evals = 0, flops = 0;  // lane 0 of each wave counts
    // Per-phase wall time (member 0, thread 0): up to B1, B3 (extend checks), the commit, B4
    // (connect NN), B7 (connect prep and walks), B8.
    unsigned long long ph[6] = {0, 0, 0, 0, 0, 0};
    unsigned long long tph = wall_clock64();
    auto mark = [&](int k) {
        const unsigned long long now = wall_clock64();
        ph[k] += now - tph;
        tph = now;
    };
    // Nearest neighbours of the explorers that tree(e) picks (0 / 1 / -1 = none), query point
    // qpt(e, d), into keys[e].  Work: tiles of kEC explorers (tree-0 tiles, then tree-1 tiles).
    // With W >= tiles, tile k goes to the W / tiles members m = k (mod tiles), each scanning its
    // part of the tree; otherwise members take whole tiles in turn.  Partial results merge
    // through a 64-bit minimum of (squared distance, index), which equals the serial scan's.
    auto nn_phase = [&](auto tree, auto qpt, unsigned long long* keys, int sz0, int sz1) {
        const int n0 = compact(S, [&](int e) { return tree(e) == 0; }, list, scan, 18);
        const int n1 = compact(S, [&](int e) { return tree(e) == 1; }, list + n0, scan, 19);
        const int nt0 = (n0 + kEC - 1) / kEC, ntl = nt0 + (n1 + kEC - 1) / kEC;
        if (ntl == 0) return;
        int first = m, stride = W, parts = 1, part = 0;
        if (W >= ntl) {
            first = m % ntl;
            stride = ntl;
            parts = W / ntl;
            part = m / ntl;
            if (part >= parts) return;
        }
        for (int k = first; k < ntl; k += stride) {
            const int t = (k < nt0) ? 0 : 1;
            const int c0 = t ? n0 + (k - nt0) * kEC : k * kEC;
            const int nc = min(kEC, (t ? n0 + n1 : n0) - c0);
            const int sz = t ? sz1 : sz0;
            // part `part` of `parts` scans rows part, part + parts, ... of `span` nodes (nn_tile),
            // so the newest nodes, which are read coherently, are shared out too
            const int span = nn_row_nodes_per_thread(MAXD) * nt;
            const int full = sz / span, tail = sz - full * span;
            const int mine = ((full > part) ? (full - part + parts - 1) / parts : 0) * span +
                             ((full % parts == part) ? tail : 0);  // nodes of the part's rows
            if (mine == 0) continue;
            const int lo = 0, hi = sz;
            for (int q = tid; q < kEC * MAXD; q += nt) {
                const int c = q / MAXD, d = q - (q / MAXD) * MAXD;
                qsh[c * kMaxDim + d] = (c < nc && d < dim) ? qpt(list[c0 + c], d) : 0.0f;
            }
            __syncthreads();
            // 7 dimensions: the 7-term instantiation (the 8-term one adds fma(0, 0, s) = s, the
            // same sums at 8/7 of the arithmetic)
            bool done7 = false;
            if constexpr (MAXD == 8) {
                if (dim == 7) {
                    nn_tile<7, COH>(t ? TB : TA, ncap, lo, hi, dim, lds_f(qsh), lds_f(red_d), lds_i(red_i),
                                    t ? coh_b : coh_a, part, parts);
                    done7 = true;
                }
            }
            if (!done7)
                nn_tile<MAXD, COH>(t ? TB : TA, ncap, lo, hi, dim, lds_f(qsh), lds_f(red_d), lds_i(red_i),
                                   t ? coh_b : coh_a, part, parts);
            if (tid == 0) {  // node bytes, passes and (query, node) pairs (no return value: the atomics do not wait)
                global_add(&a.stats[16], 4ull * static_cast<unsigned long long>(dim) * static_cast<unsigned>(mine));
                global_add(&a.stats[17], 1ull);
                global_add(&a.stats[19], static_cast<unsigned long long>(nc) * static_cast<unsigned>(mine));
            }
            if (tid < nc) gmin64(&keys[list[c0 + tid]], nn_key(red_d[tid * 16], red_i[tid * 16]));
            __syncthreads();
        }
    };
    // Connect searches of this round's successful extends against the other tree's first sz0 /
    // sz1 nodes.
    auto connect_nn = [&](int sz0, int sz1) {
        nn_phase([&](int e) { return newi[e] >= 0 ? 1 - side[e] : -1; },
                 [&](int e, int d) { return cand[e * dim + d]; }, V.g_okey, sz0, sz1);
    };
    int r_end = (r_budget >= 0) ? r_begin + r_budget : 0x7fffffff;  // r_budget < 0: no budget
    bool sampled = false;  // cand holds this round's samples (drawn during the previous B8)
    int r = r_begin;
    for (;; ++r) {
        if (r > a.max_iterations) {
            if (tid == 0) sc[SC_STATUS] = PRRTC_STATUS_ITER_CAP;
            __syncthreads();
            break;
        }
        if (r >= r_end) break;  // round budget spent: the query goes to the next continue list
        if (pass == 1 && r - r_begin >= kPass1MinRounds) {
            if (tid == 0) sc[SC_DRAIN] = (gload(a.ctl + CTL_QUEUE) >= a.B) ? 1 : 0;
            __syncthreads();
            if (sc[SC_DRAIN]) break;  // every query has started: continue in pass 2
        }
        const int size_a = sc[SC_SIZE_A], size_b = sc[SC_SIZE_B];
        // ---- phase 1a: tree side of every explorer (each member keeps a full copy); owners clear
        // the shared collision flags of their explorers for this round.
        for (int e = tid; e < S; e += nt) {
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
            if (e % W == m) {
                gstore(&V.g_fail[e], 0);
                gstore(&V.g_eclaim[e], 0);
                gstore(&V.g_claim[e], 0);
                gstore(&V.g_cfail[e], 0x7fffffff);
            }
        }
        // This round's samples of every explorer, into cand (free until phase 1c steers them),
        // unless the previous round drew them while waiting at B8.
        if (!sampled) {
            halton_round(lds + a.off_sampler, r);
            __syncthreads();
            for (int e = tid; e < S; e += nt)
                for (int d = 0; d < dim; ++d) cand[e * dim + d] = sample_dim(lds + a.off_sampler, primes + 16 * e, d);
        }
        sampled = false;
        __syncthreads();
        // ---- phase 1b: nearest neighbour of each explorer's sample in its tree
        nn_phase([&](int e) { return side[e]; },
                 [&](int e, int d) { return cand[e * dim + d]; }, V.g_key,
                 size_a, size_b);
        team_barrier(team);  // B1: nearest neighbours of all explorers are in global memory
        mark(0);
        // ---- phase 1c: dynamic-domain skip and steer, every explorer (redundant per member).
        // The nearest node's radius and coordinates are loaded together (one memory latency, not
        // dim + 1); the coordinates stay in vec (free until the connect prep) for phase 2.
        for (int e = tid; e < S; e += nt) {
            const int t = side[e];
            const float* T = t ? TB : TA;
            const float* R = t ? RB : RA;
            const unsigned long long key = gload64(&V.g_key[e]);
            const int ni = static_cast<int>(key & 0xffffffffull);
            nn[e] = ni;
            float base[MAXD];
#pragma unroll
            for (int d = 0; d < MAXD; ++d) base[d] = ldc<COH>(T + static_cast<size_t>(min(d, dim - 1)) * ncap + ni);
            const float rad = a.dd ? ldc<COH>(R + ni) : 0.0f;
            const float nd = __builtin_sqrtf(__uint_as_float(static_cast<unsigned>(key >> 32)));
            if (a.dd && rad < nd) {
                state[e] = 0;
            } else {
                const float s = (nd > 1e-8f) ? fminf(1.0f, a.step_size / nd) : 1.0f;
#pragma unroll
                for (int d = 0; d < MAXD; ++d) {
                    if (d < dim) {
                        const float smp = cand[e * dim + d];  // the sample
                        cand[e * dim + d] = base[d] + (smp - base[d]) * s;
                        vec[e * dim + d] = base[d];
                    }
                }
                state[e] = 1;
            }
        }
        __syncthreads();
        // ---- phases 2 (kind 0) and 5 (kind 1) share one check loop, so that the FK and the
        // collision checks are instantiated once (the kernels' register budget); the phase-3
        // commit and the connect prep run between them, after B3.
        //   phase 2: extend edges, waypoint t_k = (k+1)/G, k = 0..G-1.  Work unit = ku consecutive
        //     waypoints of one edge (FK one per lane, then checks in order; a collision anywhere
        //     on the edge stops the remaining units of that edge).
        //   phase 5: connect walks.  Work unit = ku consecutive waypoints of sub-edge j of one
        //     explorer's walk; a collision at sub-edge j lowers the explorer's first failing
        //     sub-edge g_cfail[e] (an order-independent minimum), and units past it are skipped.
        bool ended = false;  // phase 3 ended the query (tree full)
#pragma clang loop unroll(disable)
        for (int kind = 0; kind < 2; ++kind) {
            if (kind == 1) {
            team_barrier(team);  // B3: every extend verdict is in global memory
            mark(1);
            // ---- phase 3: commit -- appends in explorer order (prefix sums), radius updates
            {
                for (int e = tid; e < S; e += nt) {
                    if (state[e] == 1 && gload(&V.g_fail[e])) state[e] = 3;  // 3: collided
                    if (e % W == m) gstore64(&V.g_key[e], kNoKey);  // read by every member before B3
                }
                if (tid == 0) sc[SC_OVERFLOW] = 0;
                __syncthreads();
                int base0 = size_a, base1 = size_b;
                for (int e0 = 0; e0 < S; e0 += nt) {
                    const int e = e0 + tid;
                    const bool fr = e < S && state[e] == 1;
                    const int t = fr ? side[e] : 0;
                    int p0, p1, tot0, tot1;
                    block_excl_scan2((fr && t == 0) ? 1 : 0, (fr && t == 1) ? 1 : 0, scan, p0, p1, tot0, tot1);
                    if (e < S) {
                        int ni = -1;
                        if (fr) {
                            const int idx = t ? base1 + p1 : base0 + p0;
                            if (idx < N) {
                                ni = idx;
                                if (e % W == m) {
                                    float* T = t ? TB : TA;
                                    for (int d = 0; d < dim; ++d)
                                        stc<COH>(T + static_cast<size_t>(d) * ncap + idx, cand[e * dim + d]);
                                    stc<COH>((t ? PB : PA) + idx, nn[e]);
                                    stc<COH>((t ? RB : RA) + idx, FLT_MAX);
                                }
                            } else {
                                sc[SC_OVERFLOW] = 1;
                            }
                        }
                        newi[e] = ni;
                    }
                    base0 += tot0;
                    base1 += tot1;
                }
                if (a.dd && m == 0) {
                    for (int e = tid; e < S; e += nt) {
                        if (state[e] == 3) hash_add(hk, hs, hg, a.hash_size, nn[e] * 2 + side[e], true);
                        else if (newi[e] >= 0) hash_add(hk, hs, hg, a.hash_size, nn[e] * 2 + side[e], false);
                    }
                    __syncthreads();
                    for (int h = tid; h < a.hash_size; h += nt) {
                        const int key = hk[h];
                        if (key < 0) continue;
                        float* R = (key & 1) ? RB : RA;
                        float rad = ldc<COH>(R + (key >> 1));
                        for (int i = 0; i < hs[h]; ++i) radius_shrink(rad, a.dd_alpha, a.dd_radius, a.dd_min_radius);
                        for (int i = 0; i < hg[h]; ++i) radius_grow(rad, a.dd_alpha);
                        stc<COH>(R + (key >> 1), rad);
                        hk[h] = -1; hs[h] = 0; hg[h] = 0;
                    }
                }
                __syncthreads();
                if (tid == 0) {
                    sc[SC_SIZE_A] = min(base0, N);
                    sc[SC_SIZE_B] = min(base1, N);
                    if (sc[SC_OVERFLOW]) sc[SC_STATUS] = PRRTC_STATUS_TREE_FULL;
                }
            }
            // The connect NN and walk read only nodes that existed at the start of the round
            // (appends and radius updates are seen from the next round, after B7 / B8).
            __syncthreads();
            mark(2);
            if (sc[SC_STATUS] != 0) {
                ended = true;
                break;
            }
            // ---- phase 4: connect nearest neighbour of each successful extend in the other tree
            // as it stood at the start of the round (its first size_a / size_b nodes) -- the
            // interleaving in which each explorer's connect search precedes the other explorers'
            // appends of this round.  Only the extends that passed are searched (most collide);
            // a team merges its members' partial results before the connect prep reads them.
            connect_nn(size_a, size_b);
            team_barrier(team);  // B4: connect NN keys merged
            mark(3);
            for (int e = tid; e < S; e += nt) {
                if (newi[e] < 0) continue;
                const float* O = side[e] ? TA : TB;
                const unsigned long long key = gload64(&V.g_okey[e]);
                const int oi = static_cast<int>(key & 0xffffffffull);
                onn[e] = oi;
                float tgt[MAXD];  // the other tree's nearest node, loaded together
    #pragma unroll
                for (int d = 0; d < MAXD; ++d) tgt[d] = ldc<COH>(O + static_cast<size_t>(min(d, dim - 1)) * ncap + oi);
                const float cd = __builtin_sqrtf(__uint_as_float(static_cast<unsigned>(key >> 32)));
                int n_ext = 0;
                if (a.step_size > 1e-8f && cd > 1e-8f) n_ext = static_cast<int>(ceilf(cd / a.step_size));
                next[e] = n_ext;
    #pragma unroll
                for (int d = 0; d < MAXD; ++d)
                    if (d < dim)
                        vec[e * dim + d] = (n_ext > 0) ? (tgt[d] - cand[e * dim + d]) / static_cast<float>(n_ext) : 0.0f;
            }
            __syncthreads();
            }
            int nl, ku, cps, total = 0;
            bool deal, once;
            if (kind == 0) {
                nl = compact(S, [&](int e) { return state[e] == 1; }, list, scan);
                ku = unit_size(nl, G, W * nwaves, fb);
                cps = (G + ku - 1) / ku;
                // Units no more than waves: a fixed deal, unit us to wave us (no claiming; a unit
                // of an edge already known to collide is skipped).  Otherwise units are claimed
                // dynamically (claim counter per explorer), which balances the load and skips
                // edges known to collide.
                once = static_cast<long long>(nl) * cps <= W * nwaves;
                deal = once || static_cast<long long>(nl) * G < W * nwaves;
            } else {
                nl = compact(S, [&](int e) { return newi[e] >= 0; }, list, scan);
                // sub-edges of all walks: summed by wave 0, published behind one barrier
                if (tid < kWave) {
                    unsigned acc = 0;
                    for (int k = tid; k < nl; k += kWave) acc += static_cast<unsigned>(next[list[k]]);
                    acc = wave_sum_u32(acc);
                    if (tid == 0) scan[17] = static_cast<int>(acc);
                }
                __syncthreads();
                const int nsub = scan[17];
                ku = unit_size(nsub, G, W * nwaves, fb);
                cps = (G + ku - 1) / ku;
                // Units no more than waves: a fixed deal, unit us to wave us (unit offsets of the
                // walkers in aux0; units past a walk's first failure found so far are skipped).
                // Otherwise units are claimed dynamically, and a collision at sub-edge j lowers the
                // explorer's cap to j, so units past the first failure stop being handed out.
                once = static_cast<long long>(nsub) * cps <= W * nwaves;
                deal = once || static_cast<long long>(nsub) * G < W * nwaves;
                if (deal) {
                    for (int k0 = 0; k0 < nl; k0 += nt) {
                        const int k = k0 + tid;
                        int tot;
                        const int pre = block_excl_scan(k < nl ? next[list[k]] * cps : 0, scan, &tot);
                        if (k < nl) aux0[k] = total + pre;
                        total += tot;
                    }
                    __syncthreads();
                }
            }
            // consecutive units (often one explorer's) go to different members' CUs
            int us = wid * W + m;
            int start = ((m * nwaves + wid) * S) / (W * nwaves);
            for (;;) {
                int e = -1, jj = 0, k0 = 0;
                if (kind == 0) {
                    if (deal) {
                        if (us >= nl * cps) break;
                        e = list[us / cps];
                        k0 = (us - (us / cps) * cps) * ku;
                        us += W * nwaves;
                        if (uniform(gload(&V.g_fail[e]))) continue;
                    } else {
                        for (int b0 = 0; b0 < S && e < 0; b0 += kWave) {
                            bool ok = false;
                            int ex = 0;
                            if (b0 + lane < S) {
                                ex = start + b0 + lane;
                                if (ex >= S) ex -= S;
                                ok = state[ex] == 1 && gload(&V.g_fail[ex]) == 0 && gload(&V.g_eclaim[ex]) < cps;
                            }
                            const unsigned long long mask = __ballot(ok);
                            if (mask) e = __shfl(ex, __ffsll(static_cast<long long>(mask)) - 1, kWave);
                        }
                        e = uniform(e);
                        if (e < 0) break;
                        start = e;
                        int u = 0;
                        if (lane == 0)
                            u = __hip_atomic_fetch_add((PRRTC_GLOBAL int*)&V.g_eclaim[e], 1, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                        u = uniform(__shfl(u, 0, kWave));
                        if (u >= cps) continue;  // another wave took the last unit
                        k0 = u * ku;
                    }
                } else {
                    if (deal) {
                        if (us >= total) break;
                        int lo = 0, hi = nl - 1;  // the last walker whose offset is <= us
                        while (lo < hi) {
                            const int mid = (lo + hi + 1) >> 1;
                            if (aux0[mid] <= us) lo = mid;
                            else hi = mid - 1;
                        }
                        e = list[lo];
                        const int rr = us - aux0[lo];
                        jj = rr / cps;
                        k0 = (rr - jj * cps) * ku;
                        us += W * nwaves;
                    } else {
                        for (int b0 = 0; b0 < S && e < 0; b0 += kWave) {
                            bool ok = false;
                            int ex = 0;
                            if (b0 + lane < S) {
                                ex = start + b0 + lane;
                                if (ex >= S) ex -= S;
                                if (newi[ex] >= 0) {
                                    const int cap = min(next[ex], gload(&V.g_cfail[ex])) * cps;
                                    ok = gload(&V.g_claim[ex]) < cap;
                                }
                            }
                            const unsigned long long mask = __ballot(ok);
                            if (mask) e = __shfl(ex, __ffsll(static_cast<long long>(mask)) - 1, kWave);
                        }
                        e = uniform(e);
                        if (e < 0) break;
                        start = e;
                        int u = 0;
                        if (lane == 0)
                            u = __hip_atomic_fetch_add((PRRTC_GLOBAL int*)&V.g_claim[e], 1, __ATOMIC_RELAXED,
                                                       __HIP_MEMORY_SCOPE_AGENT);
                        u = uniform(__shfl(u, 0, kWave));
                        jj = u / cps;
                        k0 = (u - jj * cps) * ku;
                        if (jj >= next[e]) continue;  // past the end of the walk
                    }
                    if (jj >= uniform(gload(&V.g_cfail[e]))) continue;  // past the first failure
                }
                const int nk = min(ku, G - k0);
                if (lane < nk) {
                    const float tf = static_cast<float>(k0 + lane + 1) / static_cast<float>(G);
                    for (int d = 0; d < dim; ++d) {
                        if (kind == 0) {
                            const float base = vec[e * dim + d];  // the nearest node (phase 1c)
                            L.q[d * fbp] = base + tf * (cand[e * dim + d] - base);
                        } else {
                            float c = cand[e * dim + d];
                            const float v = vec[e * dim + d];
                            for (int q = 0; q < jj; ++q) c += v;
                            L.q[d * fbp] = c + tf * v;
                        }
                    }
                    if (M.enabled) fk_lane(M, L);
                }
                wave_sync();
                // Sub-units of up to kUnitMax consecutive waypoints share one coarse pass
                // (wave_unit_check); a sub-unit whose candidate list overflows is checked one
                // waypoint at a time.  Only whether some waypoint collides is recorded.
                for (int c0 = 0; c0 < nk; c0 += kUnitMax) {
                    // (the flag / cap was read when the unit was taken; re-read from the second
                    // sub-unit)
                    if (c0 > 0 && (kind == 0 ? uniform(gload(&V.g_fail[e])) != 0 : jj >= uniform(gload(&V.g_cfail[e]))))
                        break;
                    const int nu = min(kUnitMax, nk - c0);
                    int v = (nu > 1) ? wave_unit_in_collision(M, WA.T + c0, fbp, nu, flops) : -1;
                    if (v < 0) {
                        v = 0;
                        for (int c = 0; c < nu && v == 0; ++c)
                            if (wave_in_collision(M, WA.T + c0 + c, fbp, flops)) v = c + 1;
                    }
                    if (lane == 0) evals += (v > 0) ? v : nu;
                    if (v > 0) {
                        if (lane == 0) {
                            if (kind == 0) gstore(&V.g_fail[e], 1);
                            else gmin(&V.g_cfail[e], jj);
                        }
                        break;
                    }
                }
                wave_sync();
                // Units no more than waves: they are dealt (unit us to wave us), so a wave checks
                // at most one and no claim counters are read or raised.
                if (once) break;
            }
        }
        if (ended) break;
        team_barrier(team);  // B7: every connect walk's length is in global memory
        mark(4);
        // ---- phase 6: connect appends (explorer-major); first success or overflow decides
        {
            for (int e = tid; e < S; e += nt) {
                nok[e] = (newi[e] >= 0) ? min(next[e], gload(&V.g_cfail[e])) : 0;
                if (e % W == m) gstore64(&V.g_okey[e], kNoKey);  // read by every member before B7
            }
            if (tid == 0) sc[SC_EVENT] = S;
            __syncthreads();
            int base0 = sc[SC_SIZE_A], base1 = sc[SC_SIZE_B];
            for (int e0 = 0; e0 < S; e0 += nt) {
                const int e = e0 + tid;
                const bool act = e < S && newi[e] >= 0;
                const int t = act ? side[e] : 0;
                const int n_ok = act ? nok[e] : 0;
                int p0, p1, tot0, tot1;
                block_excl_scan2(t == 0 ? n_ok : 0, t == 1 ? n_ok : 0, scan, p0, p1, tot0, tot1);
                if (act) {
                    const int b = t ? base1 + p1 : base0 + p0;
                    const bool over = b + n_ok > N;
                    if (e % W == m) {
                        float* T = t ? TB : TA;
                        int* P = t ? PB : PA;
                        float* R = t ? RB : RA;
                        int parent = newi[e];
                        for (int k = 0; k < n_ok && b + k < N; ++k) {
                            const int idx = b + k;
                            for (int d = 0; d < dim; ++d) {
                                float c = cand[e * dim + d];
                                const float v = vec[e * dim + d];
                                for (int jj = 0; jj <= k; ++jj) c += v;
                                stc<COH>(T + static_cast<size_t>(d) * ncap + idx, c);
                            }
                            stc<COH>(P + idx, parent);
                            stc<COH>(R + idx, FLT_MAX);
                            parent = idx;
                        }
                    }
                    aux0[e] = over ? 1 : 0;
                    aux1[e] = (n_ok > 0) ? b + n_ok - 1 : newi[e];
                    if (over || n_ok == next[e]) lds_min(&sc[SC_EVENT], e);
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
        }
        // B8: connect nodes are in global memory.  Between arriving and waiting, the work-group
        // draws the next round's samples (LDS only; phase 6 no longer reads cand).  With a
        // deadline, member 0 decides before it arrives whether this is the epoch's last round;
        // every member reads that decision after the barrier.
        if (deadline != 0ull && m == 0 && tid == 0) gstore(&V.qs[QS_STOP], wall_clock64() >= deadline ? 1 : 0);
        team_arrive(team);
        halton_round(lds + a.off_sampler, r + 1);
        __syncthreads();
        for (int e = tid; e < S; e += nt)
            for (int d = 0; d < dim; ++d) cand[e * dim + d] = sample_dim(lds + a.off_sampler, primes + 16 * e, d);
        sampled = true;
        team_wait(team);
        mark(5);
        if (W > 1) {
            if (tid == 0) sc[SC_ERR] = gload(team.error);
            __syncthreads();
        }
        if (deadline != 0ull) {
            if (tid == 0) sc[SC_STOP] = gload(&V.qs[QS_STOP]);
            __syncthreads();
            if (sc[SC_STOP]) r_end = r + 1;  // the epoch's time is up: save the state after this round
        }
        if (sc[SC_STATUS] != 0 || sc[SC_ERR]) break;
    }
    // ---- counters
    if (lane == 0 && evals) {
        global_add(&a.stats[0], evals);
        global_add(&a.stats[2], flops);
    }
    if (m == 0 && tid == 0) {
        const int base = pass == 1 ? 4 : 10;  // pass 1 / pass 2 phase times
        for (int k = 0; k < 6; ++k) global_add(&a.stats[base + k], ph[k]);
    }

    const int st = sc[SC_ERR] ? -1 : sc[SC_STATUS];
    const unsigned long long t1 = wall_clock64();
    const float dt = static_cast<float>(static_cast<double>(t1 - t0) / a.wall_khz);
    if (st == 0) {
        // round budget spent: save the round state (every member holds the same values)
        if (m == 0) {
            for (int e = tid; e < S; e += nt) V.g_vamp[e] = vamp[e];
            if (tid == 0) {
                V.qs[QS_ROUND] = r;
                V.qs[QS_SIZE_A] = sc[SC_SIZE_A];
                V.qs[QS_SIZE_B] = sc[SC_SIZE_B];
                V.qs[QS_TIME] = __float_as_int(__int_as_float(V.qs[QS_TIME]) + dt);
                const int k = global_add(push_cnt, 1);
                push_list[k] = V.q;
            }
        }
        __syncthreads();
        return;
    }
    if (m == 0 && tid == 0) {
        const int q = V.q;
        a.sizes[2 * q] = sc[SC_SIZE_A];
        a.sizes[2 * q + 1] = sc[SC_SIZE_B];
        a.completed[2 * q] = sc[SC_SIZE_A];
        a.completed[2 * q + 1] = sc[SC_SIZE_B];
        a.iters[q] = r;
        a.conn[3 * q] = sc[SC_CONN0];
        a.conn[3 * q + 1] = sc[SC_CONN1];
        a.conn[3 * q + 2] = sc[SC_CONN2];
        a.solved[q] = (st == PRRTC_STATUS_SOLVED) ? 1 : -1;
        a.time_ms[q] = __int_as_float(V.qs[QS_TIME]) + dt;
        if (a.status) a.status[q] = st;
        global_add(&a.stats[1], static_cast<unsigned long long>(r));
        if (a.path) {
            float* out = a.path + static_cast<size_t>(q) * a.path_cap * dim;
            int len = 0;
            if (st == PRRTC_STATUS_SOLVED) {
                int na = 0, cur = sc[SC_CONN0];
                for (int it = 0; it < ncap; ++it) {
                    ++na;
                    const int p = ldc<COH>(PA + cur);
                    if (p == cur) break;
                    cur = p;
                }
                int nb = 0;
                cur = sc[SC_CONN1];
                for (int it = 0; it < ncap; ++it) {
                    ++nb;
                    const int p = ldc<COH>(PB + cur);
                    if (p == cur) break;
                    cur = p;
                }
                if (na + nb > a.path_cap) {
                    len = -1;
                } else {
                    cur = sc[SC_CONN0];
                    for (int k = na - 1; k >= 0; --k) {
                        for (int d = 0; d < dim; ++d) out[k * dim + d] = ldc<COH>(TA + static_cast<size_t>(d) * ncap + cur);
                        cur = ldc<COH>(PA + cur);
                    }
                    cur = sc[SC_CONN1];
                    for (int k = 0; k < nb; ++k) {
                        for (int d = 0; d < dim; ++d)
                            out[(na + k) * dim + d] = ldc<COH>(TB + static_cast<size_t>(d) * ncap + cur);
                        cur = ldc<COH>(PB + cur);
                    }
                    len = na + nb;
                }
            }
            a.path_len[q] = len;
        }
    }
    __syncthreads();
}

__device__ QueryView query_view(const PlanArgs& a, int q) {
    QueryView V;
    V.q = q;
    V.ng = a.goal_counts ? max(1, min(a.goal_counts[q], a.gmax)) : a.gmax;
    V.N = a.max_nodes_q ? min(max(a.max_nodes_q[q], V.ng), a.ncap) : a.ncap;
    V.seed = a.seeds ? a.seeds[q] : 0u;
    V.TA = a.tree_a + static_cast<size_t>(q) * a.dim * a.ncap;
    V.TB = a.tree_b + static_cast<size_t>(q) * a.dim * a.ncap;
    V.PA = a.par_a + static_cast<size_t>(q) * a.ncap;
    V.PB = a.par_b + static_cast<size_t>(q) * a.ncap;
    V.RA = a.radii + static_cast<size_t>(q) * 2 * a.ncap;
    V.RB = V.RA + a.ncap;
    int* xg = a.xg + static_cast<size_t>(q) * XG_COUNT * a.S;
    V.g_key = reinterpret_cast<unsigned long long*>(xg + XG_KEY * a.S);
    V.g_okey = reinterpret_cast<unsigned long long*>(xg + XG_OKEY * a.S);
    V.g_fail = xg + XG_FAIL * a.S;
    V.g_vamp = xg + XG_VAMP * a.S;
    V.g_eclaim = xg + XG_ECLAIM * a.S;
    V.g_claim = xg + XG_CLAIM * a.S;
    V.g_cfail = xg + XG_CFAIL * a.S;
    V.qs = a.qstate + static_cast<size_t>(q) * QS_COUNT;
    return V;
}

// Roots and round state of a new query (prrtc_planner.cu:97-125, 1179-1252).
__device__ void init_query(const PlanArgs& a, const QueryView& V) {
    const int tid = threadIdx.x, nt = blockDim.x, dim = a.dim, ncap = a.ncap;
    const float* start = a.starts + static_cast<size_t>(V.q) * dim;
    const float* goals = a.goals + static_cast<size_t>(V.q) * a.gmax * dim;
    for (int d = tid; d < dim; d += nt) V.TA[static_cast<size_t>(d) * ncap] = start[d];
    for (int i = tid; i < V.ng * dim; i += nt) {
        const int g = i / dim, d = i - (i / dim) * dim;
        V.TB[static_cast<size_t>(d) * ncap + g] = goals[i];
    }
    if (tid == 0) {
        V.PA[0] = 0;
        V.RA[0] = FLT_MAX;
        V.qs[QS_ROUND] = 1;
        V.qs[QS_SIZE_A] = 1;
        V.qs[QS_SIZE_B] = V.ng;
        V.qs[QS_TIME] = __float_as_int(0.0f);
    }
    for (int g = tid; g < V.ng; g += nt) { V.PB[g] = g; V.RB[g] = FLT_MAX; }
    for (int e = tid; e < a.S; e += nt) {
        V.g_vamp[e] = 0;
        V.g_key[e] = kNoKey;
        V.g_okey[e] = kNoKey;
    }
    __syncthreads();
}

// Pass 1: one work-group per query from the device queue.
// Barrier words of team t (t = slots: the grid).  Group areas: teams of W >= kGroupMinW number
// at most kMaxSlots / kGroupMinW; the grid takes the area after them.
constexpr int kMaxSlots = 8192;
constexpr int kGroupTeams = kMaxSlots / kGroupMinW + 1;
__host__ __device__ inline size_t sync_words() {
    return static_cast<size_t>(kSyncLine) * (2 * (kMaxSlots + 1) + static_cast<size_t>(kSyncGroups) * kGroupTeams);
}
__device__ __forceinline__ Team make_team(const PlanArgs& a, int t, int W, int m) {
    Team team{W, m, a.tsync + static_cast<size_t>(2 * kSyncLine) * t, nullptr, sync_group_count(W), a.ctl + CTL_ERROR,
              0u};
    if (team.G > 1) {
        const int gt = (t == a.slots) ? kGroupTeams - 1 : t;  // a grouped team's index is < slots / 128
        team.grp = a.tsync + static_cast<size_t>(2 * kSyncLine) * (kMaxSlots + 1) +
                   static_cast<size_t>(kSyncGroups * kSyncLine) * gt;
    }
    return team;
}

template <int MAXD>
__global__ __launch_bounds__(512) void plan_pass1(PlanArgs a) {
    extern __shared__ __align__(16) char lds[];
    int* sc = reinterpret_cast<int*>(lds + a.off_misc) + 20;
    Team team{1, 0, nullptr, nullptr, 1, a.ctl + CTL_ERROR, 0u};
    sampler_init(lds + a.off_sampler, a.min_vals, a.max_vals, a.dim);  // batch-wide; synced below
    for (;;) {
        if (threadIdx.x == 0) sc[SC_QUERY] = global_add(a.ctl + CTL_QUEUE, 1);
        __syncthreads();
        const int qi = sc[SC_QUERY];
        __syncthreads();
        if (qi >= a.B) return;
        const unsigned long long t0 = wall_clock64();
        const QueryView V = query_view(a, a.order ? a.order[qi] : qi);
        init_query(a, V);
        run_query<MAXD, false>(a, V, lds, team, 1, a.pass1_rounds, t0, 1, a.ctl + CTL_NC, a.cont);
    }
}

// Team sizes of a tail of open queries (2 .. kTailMax of them, at least two work-groups per query):
// wave 0 of every work-group computes the same layout from the same saved round states, so the
// members of a team agree on it without another grid barrier.  A query's remaining rounds are
// estimated from its trees' growth so far (a tree that fills ends the query): a tree of n nodes
// after r rounds reaches max_nodes N after about r (N / n)^kGrowthExp rounds, at most the rounds
// left to max_iterations.  (The C3 queries that fill their trees grow as r^0.75 after their first
// hundred rounds, faster before: q455 holds 1.8k, 6.6k, 19.3k and 56.0k nodes after 25, 100, 400
// and 1600 rounds.  Exponents 1.0, 1.15, 1.33 and 1.5 gave C3 steps of 255.2, 255.3, 258.2 and
// 257.8 ms, at 150 us of epoch per round.)  A team's round time is modelled as proportional to 1 + kTeamLatency / W
// (a latency-bound part and a part that spreads over the members; the longest C3 query takes
// 169, 123 and 116 us per round on teams of 64, 128 and 256).  Work-groups go, a few at a time,
// to the query with the latest estimated finish (remaining rounds x round time), lowest list
// index on ties, up to the team cap.  start[k] .. start[k + 1] are the work-groups of team k (the
// query at position k of the list); work-groups from start[ncont] on have no team this epoch.
// One lane per query, so at most one wave's worth (tails of up to 16 / 32 / 64 queries: C3 245.2
// / 244.5 / 244.7 ms, 256 dual-arm queries 52.1 / 50.6 / 50.2 ms).
constexpr int kTailMax = kWave;
constexpr float kTeamLatency = 43.0f;
constexpr float kGrowthExp = 1.15f;
// Epoch time per round of budget.  C3 steps (with kGrowthExp 1.15): 150 us 255.3 ms, 200 us
// 256.1, 250 us 254.3, 300 us 254.0; 600 us (epochs of up to 38 ms) 290.3.
constexpr float kEpochRoundUs = 250.0f;

__device__ void tail_teams(const PlanArgs& a, int cur, int ncont, int* start) {
    const int lane = threadIdx.x & (kWave - 1);
    float rem = 0.0f;
    if (lane < ncont) {
        const int q = a.cont[cur * a.B + lane];
        const int* qs = a.qstate + static_cast<size_t>(q) * QS_COUNT;
        const int ng = a.goal_counts ? max(1, min(a.goal_counts[q], a.gmax)) : a.gmax;
        const int N = a.max_nodes_q ? min(max(a.max_nodes_q[q], ng), a.ncap) : a.ncap;
        const int done = qs[QS_ROUND] - 1;
        float rm = static_cast<float>(max(a.max_iterations - done, 1));
        if (done > 0) {
            const int size[2] = {qs[QS_SIZE_A], qs[QS_SIZE_B]};
            for (int t = 0; t < 2; ++t) {
                const float fill = static_cast<float>(N) / static_cast<float>(max(size[t], 1));
                rm = fminf(rm, static_cast<float>(done) * (exp2f(kGrowthExp * log2f(fmaxf(fill, 1.0f))) - 1.0f));
            }
        }
        rem = fmaxf(rm, 1.0f);
    }
    int W = lane < ncont ? 1 : 0;
    for (int left = a.slots - ncont; left > 0;) {
        const int give = max(1, left / (2 * ncont));  // a few dozen steps, finer as they run out
        const float fin = (lane < ncont && W < a.wmax) ? rem * (1.0f + kTeamLatency / static_cast<float>(W)) : -1.0f;
        // wave-wide maximum of (estimated finish, lowest lane): finish times are positive, so
        // their bits order as unsigned integers
        unsigned long long key = fin < 0.0f ? 0ull
                                            : (static_cast<unsigned long long>(__float_as_uint(fin)) << 32) |
                                                  static_cast<unsigned>(kWave - 1 - lane);
#pragma unroll
        for (int o = kWave / 2; o > 0; o >>= 1) {
            const unsigned long long other = __shfl_xor(key, o, kWave);
            key = other > key ? other : key;
        }
        if (key == 0ull) break;  // every team at the cap
        if (lane == kWave - 1 - static_cast<int>(key & 0xffffffffull)) W = min(W + give, max(a.wmax, 1));
        left -= give;
    }
    int incl = W;  // inclusive prefix sum over the lanes
#pragma unroll
    for (int o = 1; o < kWave; o <<= 1) {
        const int t = __shfl_up(incl, o, kWave);
        if (lane >= o) incl += t;
    }
    if (lane == 0) start[0] = 0;
    if (lane < ncont) start[lane + 1] = incl;
}

// Pass 2: teams of work-groups resume the queries left open by pass 1, in epochs.  In each
// epoch the open queries are dealt to teams -- a tail of at most kTailMax queries to teams sized
// by tail_teams, more queries to equal teams of W = slots / open queries (capped) -- and run for
// at most 4 << epoch rounds (capped at `epoch_rounds`); the queries still open then form the next
// epoch's list, so the teams grow as queries finish.  A grid barrier separates the epochs.
template <int MAXD>
__global__ __launch_bounds__(512) void plan_pass2(PlanArgs a) {
    extern __shared__ __align__(16) char lds[];
    __shared__ int s_ncont;
    __shared__ int s_start[kTailMax + 1];
    const int b = blockIdx.x;
    Team grid = make_team(a, a.slots, a.slots, b);
    sampler_init(lds + a.off_sampler, a.min_vals, a.max_vals, a.dim);  // batch-wide; synced below
    for (int ep = 0;; ++ep) {
        const int cur = ep % 3, nxt = (ep + 1) % 3;
        if (threadIdx.x == 0) s_ncont = gload(&a.ctl[CTL_NC + cur]);
        __syncthreads();
        const int ncont = s_ncont;
        if (ncont == 0) return;  // every work-group reads the same count after the grid barrier
        // list (ep + 2) % 3 was read in the previous epoch and is filled in the next one
        if (b == 0 && threadIdx.x == 0) gstore(&a.ctl[CTL_NC + (ep + 2) % 3], 0);
        int W, team_id, member, nteams;
        const bool tail = ncont >= 2 && ncont <= kTailMax && a.slots >= 2 * ncont;
        if (tail) {
            if (threadIdx.x < kWave) tail_teams(a, cur, ncont, s_start);
            __syncthreads();
            team_id = ncont;  // no team
            for (int k = 0; k < ncont; ++k)
                if (b >= s_start[k] && b < s_start[k + 1]) team_id = k;
            W = (team_id < ncont) ? s_start[team_id + 1] - s_start[team_id] : 1;
            member = (team_id < ncont) ? b - s_start[team_id] : 0;
            nteams = ncont;
        } else {
            W = max(1, min(a.wmax, a.slots / ncont));  // every resident work-group has a team
            if (a.slots % (8 * W) == 0) {  // keep a team on one XCD (work-groups are dealt round-robin)
                team_id = (b % 8) + 8 * ((b / 8) / W);
                member = (b / 8) % W;
            } else {
                team_id = b / W;
                member = b % W;
            }
            nteams = a.slots / W;
        }
        // Epoch budgets grow 4, 8, 16, ... up to epoch_rounds: early epochs end soon, so the
        // teams re-form (and grow) while most queries of a batch finish.
        const int budget = (a.epoch_rounds > 0) ? min(a.epoch_rounds, 4 << min(ep, 20)) : -1;
        // Tail teams run at different speeds, so their epochs end at a time, not a round count:
        // about `budget` rounds of kEpochRoundUs each, and at most 4 x budget rounds.
        const unsigned long long span = (tail && budget > 0)
            ? static_cast<unsigned long long>(kEpochRoundUs * static_cast<float>(budget) * a.wall_khz * 1e-3f) : 0ull;
        if (team_id < nteams && team_id < ncont) {
            Team team = make_team(a, team_id, W, member);
            if (threadIdx.x == 0)  // the words are at rest: every barrier on them ended before the grid barrier
                team.gen = __hip_atomic_load((PRRTC_GLOBAL unsigned*)(team.sync + kSyncLine), __ATOMIC_ACQUIRE,
                                             __HIP_MEMORY_SCOPE_AGENT);
            for (int k = team_id; k < ncont; k += nteams) {
                const unsigned long long t0 = wall_clock64();
                const QueryView V = query_view(a, a.cont[cur * a.B + k]);
                run_query<MAXD, true>(a, V, lds, team, V.qs[QS_ROUND], span ? 4 * budget : budget, t0, 2,
                                a.ctl + CTL_NC + nxt, a.cont + nxt * a.B, span ? t0 + span : 0ull);
                // members leave the query together before the next one; the team's last barrier
                // resets its counters for the next epoch's team shapes
                team_barrier(team, k + nteams >= ncont);
            }
        }
        grid_barrier(grid);
    }
}

__global__ void reset_kernel(int* ctl, unsigned long long* stats, unsigned* tsync, unsigned long long shape,
                             unsigned long long fk_tile) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t < kCtl) ctl[t] = 0;
    if (t < kStats) stats[t] = (t == 3) ? shape : (t == 18) ? fk_tile : 0ull;
    const int lines = static_cast<int>(sync_words() / kSyncLine);
    for (int i = t; i < lines; i += gridDim.x * blockDim.x) tsync[static_cast<size_t>(i) * kSyncLine] = 0u;
}

// ---------------------------------------------------------------------------------- host side
struct WsLayout {
    size_t radii, xg, qstate, ctl, cont, tsync, stats, total;
};

WsLayout ws_layout(int B, int ncap, int S) {
    WsLayout w;
    size_t o = 0;
    w.radii = o; o = align16(o + sizeof(float) * 2 * static_cast<size_t>(ncap) * B);
    w.xg = o; o = align16(o + sizeof(int) * XG_COUNT * static_cast<size_t>(S) * B);
    w.qstate = o; o = align16(o + sizeof(int) * QS_COUNT * static_cast<size_t>(B));
    w.ctl = o; o = align16(o + sizeof(int) * kCtl);
    w.cont = o; o = align16(o + sizeof(int) * 3 * static_cast<size_t>(B));
    o = (o + 255) & ~static_cast<size_t>(255);  // barrier lines start on a 256-byte boundary
    w.tsync = o; o = align16(o + sizeof(unsigned) * sync_words());
    w.stats = o; o = align16(o + kStats * sizeof(unsigned long long));
    w.total = o + 256;
    return w;
}

size_t planner_lds(const ModelLayout& ml, int S, int dim, PlanArgs& a) {
    size_t o = ml.end;
    a.off_expl = o; o = align16(o + expl_bytes(S));
    a.off_cand = o; o = align16(o + sizeof(float) * 2 * static_cast<size_t>(S) * dim);
    a.off_primes = o; o = align16(o + 16 * static_cast<size_t>(S));
    a.off_sampler = o; o = align16(o + sampler_bytes());
    a.off_nn = o; o = align16(o + nn_bytes());
    a.hash_size = hash_size_for(S);
    a.off_hash = o; o = align16(o + 3 * sizeof(int) * static_cast<size_t>(a.hash_size));
    a.off_misc = o; o = align16(o + sizeof(int) * (20 + SC_COUNT));
    return o;
}

// Private memory (scratch) of the planner kernels per lane.  The team pass is a cooperative launch
// that keeps every resident work-group alive at once, so its scratch is needed for every lane of
// the GPU together: a launch needing more than kMaxPassScratch bytes in total is refused instead
// of being left to the runtime's scratch allocation.
constexpr size_t kMaxPassScratch = size_t(64) << 20;

template <int MAXD>
size_t pass_scratch_per_lane() {
    hipFuncAttributes f1, f2;
    std::memset(&f1, 0, sizeof(f1));
    std::memset(&f2, 0, sizeof(f2));
    if (hipFuncGetAttributes(&f1, reinterpret_cast<const void*>(&plan_pass1<MAXD>)) != hipSuccess) return 0;
    if (hipFuncGetAttributes(&f2, reinterpret_cast<const void*>(&plan_pass2<MAXD>)) != hipSuccess) return 0;
    return std::max(f1.localSizeBytes, f2.localSizeBytes);
}

template <int MAXD>
int occupancy(int nt, size_t lds) {
    int b1 = 0, b2 = 0;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b1, plan_pass1<MAXD>, nt, lds) != hipSuccess) b1 = 1;
    if (hipOccupancyMaxActiveBlocksPerMultiprocessor(&b2, plan_pass2<MAXD>, nt, lds) != hipSuccess) b2 = 1;
    return std::max(1, std::min(b1, b2));
}

template <int MAXD>
hipError_t launch_passes(PlanArgs& a, int nt, size_t lds, hipStream_t s) {
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plan_pass1<MAXD>),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    e = hipFuncSetAttribute(reinterpret_cast<const void*>(&plan_pass2<MAXD>),
                            hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(plan_pass1<MAXD>, dim3(std::min(a.slots, a.B)), dim3(nt), lds, s, a);
    e = hipGetLastError();
    if (e != hipSuccess) return e;
    void* args[] = {&a};
    return hipLaunchCooperativeKernel(reinterpret_cast<const void*>(&plan_pass2<MAXD>), dim3(a.slots), dim3(nt), args,
                                      static_cast<unsigned>(lds), s);
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
    if (world_per_query && !(world->world_spheres_count && world->world_capsules_count && world->world_boxes_count &&
                             world->world_halfspaces_count && world->self_pairs_count))
        return fail_with(PRRTC_E_INVALID, "per-query worlds need the five count vectors");
    const int dim = at->dim, S = at->num_new_samples, ncap = at->max_nodes;
    const WsLayout wl = ws_layout(B, ncap, S);
    if (!workspace || workspace_bytes < wl.total) return fail_with(PRRTC_E_WORKSPACE, "workspace too small");

    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_with(PRRTC_E_HIP, "hipGetDevice failed");
    const DevInfo& di = dev_info(dev);
    // 512 threads (8 waves) per work-group by default: the planner kernels take 256 VGPRs, so one
    // such work-group fills a CU, and a team of them has half the members (half the per-member
    // redundant work and barrier arrivals) of 256-thread work-groups.
    int nt = (launch && launch->threads > 0) ? launch->threads : 512;
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
    a.xg = reinterpret_cast<int*>(ws + wl.xg);
    a.qstate = reinterpret_cast<int*>(ws + wl.qstate);
    a.ctl = reinterpret_cast<int*>(ws + wl.ctl);
    a.cont = reinterpret_cast<int*>(ws + wl.cont);
    a.tsync = reinterpret_cast<unsigned*>(ws + wl.tsync);
    a.stats = reinterpret_cast<unsigned long long*>(ws + wl.stats);
    a.wall_khz = di.wall_khz;

    // LDS: the FK batch (16 configurations per wave, or 8 when that lets more work-groups stay
    // resident, as for large robots), shrinking the work-group if even that does not fit.
    const bool small = dim <= 8;
    const int fk_batch = launch ? launch->fk_batch : 0;
    if (fk_batch != 0 && fk_batch != kFB && fk_batch != kFB / 2)
        return fail_with(PRRTC_E_INVALID, "fk_batch must be 0, 8 or 16");
    size_t lds = 0;
    int per_cu = 0;
    for (;;) {
        for (const int fbp : {kFBP, kFB / 2 + 1}) {
            if (fk_batch != 0 && fbp != fk_batch + 1) continue;
            PlanArgs t = a;
            t.ml = model_layout(robot->n_joints, robot->n_robot_spheres, world->n_self_pairs, world->n_world_spheres,
                                world->n_world_capsules, world->n_world_boxes, world->n_world_halfspaces, nt / kWave,
                                fbp, kUnitCap);
            const size_t l = planner_lds(t.ml, S, dim, t);
            if (l > static_cast<size_t>(di.max_lds)) continue;
            const int occ = small ? occupancy<8>(nt, l) : occupancy<16>(nt, l);
            if (occ > per_cu) {
                per_cu = occ;
                lds = l;
                a = t;
            }
        }
        if (per_cu > 0 || nt <= kWave) break;
        nt -= kWave;
    }
    if (per_cu == 0)
        return fail_with(PRRTC_E_LDS, "robot + world + explorer state exceed the work-group LDS (max " +
                                           std::to_string(di.max_lds) + " bytes)");
    const int resident = per_cu * di.cus;
    const int slots = (launch && launch->slots > 0) ? std::min(launch->slots, resident) : resident;
    a.slots = std::max(1, std::min(slots, kMaxSlots));
    // A lone query goes straight to a team; a batch first gives every query a short solo run.
    // Pass 1 (one work-group per query) pays off only when the batch fills the resident slots;
    // a smaller batch starts directly on teams of >= 2 work-groups.
    a.pass1_rounds = (launch && launch->pass1_rounds >= 0) ? launch->pass1_rounds : (2 * B <= a.slots ? 0 : 24);
    a.wmax = (launch && launch->team_max > 0) ? launch->team_max : 512;
    a.epoch_rounds = (launch && launch->epoch_rounds > 0) ? launch->epoch_rounds : 64;
    const unsigned long long shape = static_cast<unsigned long long>(a.slots) |
                                     (static_cast<unsigned long long>(nt) << 16) |
                                     (static_cast<unsigned long long>(a.pass1_rounds & 0xffff) << 32) |
                                     (static_cast<unsigned long long>(a.wmax & 0xffff) << 48);
    const size_t scratch_lane = small ? pass_scratch_per_lane<8>() : pass_scratch_per_lane<16>();
    if (scratch_lane * static_cast<size_t>(nt) * static_cast<size_t>(a.slots) > kMaxPassScratch)
        return fail_with(PRRTC_E_HIP, "planner kernels use " + std::to_string(scratch_lane) +
                                          " bytes of private memory per lane: a cooperative launch of " +
                                          std::to_string(a.slots) + " work-groups would need more than " +
                                          std::to_string(kMaxPassScratch >> 20) + " MiB of scratch at once");
    hipLaunchKernelGGL(reset_kernel, dim3(8), dim3(256), 0, stream, a.ctl, a.stats, a.tsync, shape,
                       static_cast<unsigned long long>(a.ml.fbp - 1));
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    e = small ? launch_passes<8>(a, nt, lds, stream) : launch_passes<16>(a, nt, lds, stream);
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    return PRRTC_OK;
}

// ---------------------------------------------------------------------------------- no-allocation entry
// prrtc_planner_noalloc takes its workspace from `scratch`, a file-scope device array: every GPU
// the library runs on holds its own copy until the process ends.  In host code the array's name
// is only a stand-in for its device address, so the address is looked up once per device (with
// that device current) and the kernels receive it as their workspace pointer.  Calls share the
// array: calls that overlap on one GPU race.  Everything the planner reads from its workspace
// before writing it (control words, statistics, barrier lines) is cleared by reset_kernel at the
// start of every call, so a call never sees the previous call's contents.
static __device__ __align__(256) unsigned char scratch[PRRTC_NOALLOC_SCRATCH_BYTES];

static std::mutex g_scratch_mu;
static void* g_scratch_ptr[64];

// Device address of this device's `scratch` (looked up on the first call on the device).
static int scratch_of_device(int dev, void** ptr) {
    if (dev < 0 || dev >= 64) return fail_with(PRRTC_E_INVALID, "device index out of range");
    std::lock_guard<std::mutex> lock(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* p = nullptr;
        const hipError_t e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !p)
            return fail_with(PRRTC_E_HIP, std::string("hipGetSymbolAddress(scratch) failed: ") + hipGetErrorString(e));
        g_scratch_ptr[dev] = p;
    }
    *ptr = g_scratch_ptr[dev];
    return PRRTC_OK;
}

}  // namespace prrtc

using namespace prrtc;

extern "C" {

size_t prrtc_planner_workspace_bytes(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples) {
    if (batch < 1 || dim < 1 || max_nodes < 1 || num_new_samples < 1) return 0;
    return ws_layout(batch, max_nodes, num_new_samples).total;
}

size_t prrtc_planner_stats_offset(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples) {
    if (batch < 1 || dim < 1 || max_nodes < 1 || num_new_samples < 1) return 0;
    return ws_layout(batch, max_nodes, num_new_samples).stats;
}

int prrtc_planner(prrtc_stream_t stream, const float* start, const float* goals, int32_t num_goals,
                  const float* min_vals, const float* max_vals, const prrtc_robot* robot, const prrtc_world* world,
                  const prrtc_attrs* attrs, const prrtc_outputs* out, const prrtc_launch* launch, void* workspace,
                  size_t workspace_bytes) {
    return planner_impl(reinterpret_cast<hipStream_t>(stream), 1, start, goals, num_goals, nullptr, min_vals, max_vals,
                        robot, world, 0, nullptr, nullptr, nullptr, attrs, out, launch, workspace, workspace_bytes);
}

int prrtc_planner_noalloc(prrtc_stream_t stream, const float* start, const float* goals, int32_t num_goals,
                          const float* min_vals, const float* max_vals, const prrtc_robot* robot,
                          const prrtc_world* world, const prrtc_attrs* attrs, const prrtc_outputs* out,
                          const prrtc_launch* launch) {
    if (!attrs) return fail_with(PRRTC_E_INVALID, "attrs/outputs are null");
    if (attrs->max_nodes < 1 || attrs->max_nodes > (1 << 29)) return fail_with(PRRTC_E_INVALID, "max_nodes out of range");
    if (attrs->num_new_samples < 1 || attrs->num_new_samples > PRRTC_MAX_SAMPLES)
        return fail_with(PRRTC_E_INVALID, "num_new_samples must be in [1, 1024]");
    const size_t need = ws_layout(1, attrs->max_nodes, attrs->num_new_samples).total;
    if (need > sizeof(scratch))
        return fail_with(PRRTC_E_WORKSPACE, "max_nodes " + std::to_string(attrs->max_nodes) + " with num_new_samples " +
                                                std::to_string(attrs->num_new_samples) + " needs " +
                                                std::to_string(need) + " workspace bytes; prrtc_planner_noalloc holds " +
                                                std::to_string(sizeof(scratch)) + " (use prrtc_planner)");
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_with(PRRTC_E_HIP, "hipGetDevice failed");
    void* ws = nullptr;
    const int rc = scratch_of_device(dev, &ws);
    if (rc) return rc;
    return planner_impl(reinterpret_cast<hipStream_t>(stream), 1, start, goals, num_goals, nullptr, min_vals, max_vals,
                        robot, world, 0, nullptr, nullptr, nullptr, attrs, out, launch, ws, sizeof(scratch));
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

}  // extern "C"

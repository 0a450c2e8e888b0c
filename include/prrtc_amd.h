/*
 * prrtc_amd.h -- C ABI of the MI355X pRRTC planner library (_prrtc_planner_lib.so).
 *
 * Plain pointers and sizes only; every array pointer is DEVICE memory unless stated otherwise,
 * and every call is enqueue-only on the given HIP stream (no host synchronisation, no device
 * allocation: scratch comes from the caller's workspace).  Return value: 0 on success, a
 * negative PRRTC_E_* code otherwise; prrtc_last_error() gives the message (thread-local).
 *
 * Each entry point replaces one XLA FFI handler of the reference; the mapping is given on the
 * declaration.  The reference's argument groups are kept, in order, as structs:
 *   26 buffer arguments  -> start/goals/min/max + prrtc_robot + prrtc_world (world primitives,
 *                           self pairs and the six coarse arrays; batch calls add the reference's
 *                           five per-query count vectors)
 *   10 results           -> prrtc_outputs (first ten fields, same order and dtypes)
 *   14 attributes        -> prrtc_attrs (same order and dtypes)
 */
#ifndef PRRTC_AMD_H
#define PRRTC_AMD_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct ihipStream_t* prrtc_stream_t; /* hipStream_t */

#define PRRTC_OK 0
#define PRRTC_E_INVALID -1  /* bad argument (sizes, limits, null pointers) */
#define PRRTC_E_HIP -2      /* HIP runtime error */
#define PRRTC_E_WORKSPACE -3/* workspace too small */
#define PRRTC_E_LDS -4      /* robot + world do not fit the work-group's LDS */

/* Per-query status (the reference only has solved_flag 1/-1, prrtc_planner.cu:776,929,1058). */
#define PRRTC_STATUS_SOLVED 1
#define PRRTC_STATUS_ITER_CAP 2
#define PRRTC_STATUS_TREE_FULL 3

#define PRRTC_MAX_JOINTS 64
#define PRRTC_MAX_DIM 16
#define PRRTC_MAX_SAMPLES 1024
#define PRRTC_MAX_GRANULARITY 64

/* Robot FK + sphere model: collision_context keys fk_* and sphere_* (README.md:81-93),
 * arguments 5-15 of PrrtcPlannerFfi (prrtc_planner.cu:1075-1085). */
typedef struct {
    const float* fk_twists;        /* [n_joints, 6]  (v, omega) */
    const float* fk_parent_tf;     /* [n_joints, 7]  (w, x, y, z, tx, ty, tz) */
    const int32_t* fk_parent_idx;  /* [n_joints] */
    const int32_t* fk_act_idx;     /* [n_joints] */
    const float* fk_mimic_mul;     /* [n_joints] */
    const float* fk_mimic_off;     /* [n_joints] */
    const int32_t* fk_mimic_act_idx; /* [n_joints] */
    const int32_t* fk_topo_inv;    /* [n_joints] */
    int32_t n_joints;
    const int32_t* sphere_link_idx;/* [n_robot_spheres] */
    const float* sphere_local;     /* [n_robot_spheres, 3] */
    const float* sphere_radius;    /* [n_robot_spheres] */
    int32_t n_robot_spheres;
} prrtc_robot;

/* World primitives, self pairs and the coarse-checker arrays: arguments 16-26 of PrrtcPlannerFfi
 * (prrtc_planner.cu:1086-1096), in the same order.  Single-query calls: arrays hold n_* rows.
 * Batch calls with per-query worlds (PrrtcPlannerBatchCtxFfi, prrtc_planner.cu:1946-1980): arrays
 * are padded [batch, n_*, F] (n_* = the padded capacity) and the five per-query counts are the
 * reference's five device i32 [batch] vectors (world_spheres_count ... self_pairs_count, arguments
 * 21-25 there); a NULL count vector means every query uses n_*.
 * The six coarse arrays (utils.py:160-227; prrtc.py:76-110) are accepted and checked for
 * consistent sizes; they change no result: the culled check derives its own per-frame bounding
 * spheres (collision modes 2/3 return the binary verdict, prrtc_planner.cu:421-425). */
typedef struct {
    const float* world_spheres;    /* [.., 4]  (x, y, z, r) */
    const float* world_capsules;   /* [.., 7]  (p1, p2, r) */
    const float* world_boxes;      /* [.., 15] (c, a1, a2, a3, half lengths) */
    const float* world_halfspaces; /* [.., 6]  (n, p); free side is +n */
    const int32_t* self_pairs;     /* [.., 2] */
    int32_t n_world_spheres, n_world_capsules, n_world_boxes, n_world_halfspaces, n_self_pairs;
    const int32_t* world_spheres_count;    /* batch only: [batch] or NULL */
    const int32_t* world_capsules_count;   /* batch only: [batch] or NULL */
    const int32_t* world_boxes_count;      /* batch only: [batch] or NULL */
    const int32_t* world_halfspaces_count; /* batch only: [batch] or NULL */
    const int32_t* self_pairs_count;       /* batch only: [batch] or NULL */
    const int32_t* coarse_sphere_link_idx; /* [n_coarse_spheres] */
    const float* coarse_sphere_local;      /* [n_coarse_spheres, 3] */
    const float* coarse_sphere_radius;     /* [n_coarse_spheres] */
    const int32_t* fine_link_offsets;      /* [n_fine_link_offsets] */
    const int32_t* coarse_self_pairs;      /* [n_coarse_self_pairs, 2] */
    const int32_t* self_pair_ranges;       /* [n_self_pair_ranges, 2] */
    int32_t n_coarse_spheres, n_fine_link_offsets, n_coarse_self_pairs, n_self_pair_ranges;
} prrtc_world;

/* The 14 FFI attributes, same order (prrtc_planner.cu:2103-2116, prrtc.py:537-550). */
typedef struct {
    int32_t max_iterations;
    float step_size;
    int32_t num_new_samples;
    int32_t balance_mode;
    float tree_ratio;
    int32_t dynamic_domain;
    float dd_alpha;
    float dd_radius;
    float dd_min_radius;
    int32_t dim;
    int32_t max_nodes;
    int32_t granularity;
    int32_t collision_mode;  /* 0 binary, 1 margin ("differentiable"), 2/3 coarse (binary verdict) */
    float collision_margin;
} prrtc_attrs;

/* The 10 FFI results (prrtc.py:384-395), plus optional extras (NULL = not wanted).
 * Batch calls prefix every array with [batch] (prrtc.py:1088-1099). */
typedef struct {
    float* tree_a_configs;     /* [dim, max_nodes] SoA */
    float* tree_b_configs;     /* [dim, max_nodes] */
    int32_t* tree_a_parents;   /* [max_nodes]; a root is its own parent */
    int32_t* tree_b_parents;   /* [max_nodes] */
    int32_t* tree_sizes;       /* [2] */
    int32_t* completed;        /* [2] (equal to tree_sizes: every counted node is written) */
    int32_t* iter_count;       /* [1] lockstep rounds run */
    int32_t* connection_info;  /* [3] a_idx, b_idx, extended tree */
    int32_t* solved_flag;      /* [1] 1 solved, -1 failed */
    float* kernel_time_ms;     /* [1] device time of this query */
    int32_t* status;           /* [1] PRRTC_STATUS_* (optional) */
    float* path;               /* [path_capacity, dim] start -> goal (optional) */
    int32_t* path_len;         /* [1] nodes written, 0 if unsolved, -1 if over capacity */
    int32_t path_capacity;
} prrtc_outputs;

/* Launch shape (NULL = defaults chosen by the library).  Results never depend on it. */
typedef struct {
    int32_t threads;      /* threads per work-group, multiple of 64 (planner <= 512, checks <= 1024); 0 = auto */
    int32_t slots;        /* resident work-groups; 0 = all that fit */
    int32_t pass1_rounds; /* planner: rounds a query runs on one work-group before it is handed to a team of
                             work-groups (once every query of the batch has started, a query moves on
                             after 4 rounds); -1 = auto (24 when the batch fills the resident
                             work-groups, else 0: every query starts on a team) */
    int32_t team_max;     /* planner: largest team (work-groups per query); 0 = auto (512) */
    int32_t epoch_rounds; /* planner: most rounds a query runs between two re-formings of the teams of
                             the team pass (epochs run 4, 8, 16, ... rounds up to this); 0 = auto (64).
                             With 2 to 64 open queries the teams are sized by each query's estimated
                             remaining rounds and an epoch lasts about 250 us per round of its budget
                             (at most 4x its rounds), so teams of different speeds end it together */
    int32_t fk_batch;     /* planner: configurations per wave in one FK + check tile (8 or 16; a check
                             work unit covers at most this many waypoints); 0 = auto (16, or 8 when
                             that lets more work-groups stay resident) */
} prrtc_launch;

/* Workspace bytes a planner call needs (radii + per-query explorer state + queue counters). */
size_t prrtc_planner_workspace_bytes(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples);

/* Byte offset, inside that workspace, of 32 uint64 counters the last planner call left there:
 * [0] configurations FK+collision-checked, [1] lockstep rounds summed over queries,
 * [2] algorithmic FP32 flops of that checking (FMA = 2), [3] launch shape chosen (bits 0-15
 * resident work-groups, 16-31 threads per work-group, 32-47 pass-1 rounds, 48-63 team cap),
 * [4..9] / [10..15] device wall-clock ticks spent in the round phases of the per-query pass /
 * the team pass (sample + NN, extend checks, commit, connect NN, connect walk, connect commit),
 * [16] algorithmic bytes of the nearest-neighbour scans (4 * dim per node per pass of a tile of
 * up to 16 queries), [17] such tile passes, [18] configurations per FK tile of a wave (8 or 16),
 * [19] (query, node) pairs those passes scored (3 * dim flops each), [20..31] zero.
 * No reference counterpart (the reference reports only kernel_time_ms). */
size_t prrtc_planner_stats_offset(int32_t batch, int32_t dim, int32_t max_nodes, int32_t num_new_samples);

/* Single query.  Replaces PrrtcPlannerFfi (prrtc_planner.cu:2063-2117, handler 1069-1396):
 * start [dim], goals [num_goals, dim], min/max [dim]. */
int prrtc_planner(prrtc_stream_t stream, const float* start, const float* goals, int32_t num_goals,
                  const float* min_vals, const float* max_vals, const prrtc_robot* robot, const prrtc_world* world,
                  const prrtc_attrs* attrs, const prrtc_outputs* out, const prrtc_launch* launch, void* workspace,
                  size_t workspace_bytes);

/* Bytes of the device array prrtc_planner_noalloc takes its workspace from (per GPU). */
#define PRRTC_NOALLOC_SCRATCH_BYTES (16u << 20)

/* prrtc_planner without a workspace argument and without any device allocation.  No counterpart
 * in the reference: it is for callers that cannot allocate (code under HIP graph capture, tight
 * serving loops, harnesses that forbid hipMalloc).  The workspace is a static device array of
 * PRRTC_NOALLOC_SCRATCH_BYTES on each GPU, held until the process ends; a call whose
 * prrtc_planner_workspace_bytes(1, dim, max_nodes, num_new_samples) exceeds it returns
 * PRRTC_E_WORKSPACE (16 MiB covers max_nodes up to 1.5 million at any num_new_samples).
 * Calls share the array: calls that overlap on one GPU (two streams, or two threads) race.
 * Results equal prrtc_planner's. */
int prrtc_planner_noalloc(prrtc_stream_t stream, const float* start, const float* goals, int32_t num_goals,
                          const float* min_vals, const float* max_vals, const prrtc_robot* robot,
                          const prrtc_world* world, const prrtc_attrs* attrs, const prrtc_outputs* out,
                          const prrtc_launch* launch);

/* Batch of independent queries in one persistent launch.  Replaces PrrtcPlannerBatchFfi
 * (prrtc_planner.cu:2007-2061; one shared world: world_per_query == 0, count vectors ignored) and
 * PrrtcPlannerBatchCtxFfi (prrtc_planner.cu:1946-2005; padded per-query worlds with the five
 * count vectors: world_per_query == 1).
 * starts [batch, dim], goals [batch, max_goals, dim].  Extensions over the reference (all may be
 * NULL): goal_counts [batch] (uneven goal sets), max_nodes_q [batch] (per-query tree capacity,
 * <= attrs->max_nodes), seeds [batch] (sampler seed; 0 = the reference's sequence), order [batch]
 * (queue order, a permutation of 0..batch-1; results do not depend on it). */
int prrtc_planner_batch(prrtc_stream_t stream, int32_t batch, const float* starts, const float* goals,
                        int32_t max_goals, const int32_t* goal_counts, const float* min_vals, const float* max_vals,
                        const prrtc_robot* robot, const prrtc_world* world, int32_t world_per_query,
                        const int32_t* max_nodes_q, const uint32_t* seeds, const int32_t* order,
                        const prrtc_attrs* attrs, const prrtc_outputs* out, const prrtc_launch* launch,
                        void* workspace, size_t workspace_bytes);

/* Collision verdict of n configs [n, dim] against one robot + world (no reference FFI target;
 * it is the planner's prrtc_config_in_collision, prrtc_planner.cu:501-531, exposed for parity
 * and throughput).  out_hit [n] uint8 (may be NULL): the verdict of collision_mode (binary: any
 * pair distance <= 0; 1: any pair distance < collision_margin).  out_min_dist [n] float (may be
 * NULL): the smallest signed distance over every pair (a full sweep; square roots within 1 ulp).
 * work_flops (may be NULL): a device uint64 that the call atomically increments by the
 * algorithmic flops of the checking it did. */
int prrtc_check_configs(prrtc_stream_t stream, const float* configs, int32_t n, int32_t dim, const prrtc_robot* robot,
                        const prrtc_world* world, int32_t collision_mode, float collision_margin, uint8_t* out_hit,
                        float* out_min_dist, const prrtc_launch* launch, uint64_t* work_flops);

/* FK of n configs: out [n, n_joints, 7] world poses in the original joint order (pyroffi
 * fk_single semantics, prrtc_planner.cu:509-522). */
int prrtc_forward_kinematics(prrtc_stream_t stream, const float* configs, int32_t n, int32_t dim,
                             const prrtc_robot* robot, float* out);

/* Replaces PrrtcNearestNeighborFfi (prrtc_nearest_neighbor.cu:86-141): nearest node of each
 * query [batch, dim] among the first tree_size nodes of tree [dim, max_nodes]; ties go to the
 * lowest index; dists are Euclidean (sqrt), as the reference returns. */
int prrtc_nearest_neighbor(prrtc_stream_t stream, const float* tree_configs, int32_t dim, int32_t max_nodes,
                           int32_t tree_size, const float* queries, int32_t batch, int32_t* out_idx, float* out_dist);

/* Replaces PrrtcExtendFfi (prrtc_extend.cu:84-140): steer from tree node nearest[b] toward
 * samples[b] by at most step_size (prrtc_planner.cu:870-876 arithmetic, with its 1e-8 guard). */
int prrtc_extend(prrtc_stream_t stream, const float* tree_configs, int32_t dim, int32_t max_nodes,
                 const int32_t* nearest, const float* samples, int32_t batch, float step_size, float* out_configs,
                 int32_t* out_parents, int32_t* out_valid);

/* Differentiable sphere-penetration cost (no reference counterpart; the nearest reference code
 * is the margin predicate, prrtc_planner.cu:314-387).  For each of n configs [n, dim]:
 *   cost = sum over robot spheres i, world spheres w of softplus_beta(r_i + r_w - |c_i - o_w|),
 *   softplus_beta(x) = log(1 + exp(beta x)) / beta,
 * and, when out_grad [n, dim] is not NULL, d cost / d config through the FK chain (hand-written
 * adjoint; FK is recomputed, no per-link transforms are stored).  world_spheres [n_ws, 4]. */
int prrtc_sphere_cost(prrtc_stream_t stream, const float* configs, int32_t n, int32_t dim, const prrtc_robot* robot,
                      const float* world_spheres, int32_t n_world_spheres, float beta, float* out_cost,
                      float* out_grad, uint64_t* work_flops);

/* Workspace bytes of prrtc_refine_paths. */
size_t prrtc_refine_workspace_bytes(int32_t n_paths, int32_t n_waypoints, int32_t dim);

/* Path refinement (no reference counterpart): `iterations` gradient steps on the interior
 * waypoints of paths [n_paths, n_waypoints, dim] (updated in place) of
 *   sum of prrtc_sphere_cost over all waypoints (world->world_spheres) + smooth_weight * sum |q_{k+1} - q_k|^2.
 * Each step computes every waypoint's candidate q - step * gradient (clamped to [min_vals,
 * max_vals]; both NULL = no clamp), then moves the even interior waypoints and after them the
 * odd ones: a waypoint takes its candidate only if the candidate and 15 interior points of each
 * edge to its neighbours (t = j/16) pass the exact predicate (collision_mode/margin against the
 * whole world, as prrtc_check_configs), so every edge of a refined path is checked.  End points
 * never move.  out_cost [iterations + 1] (may be NULL): the total cost before each step and
 * after the last. */
int prrtc_refine_paths(prrtc_stream_t stream, float* paths, int32_t n_paths, int32_t n_waypoints, int32_t dim,
                       const float* min_vals, const float* max_vals, const prrtc_robot* robot,
                       const prrtc_world* world, int32_t collision_mode, float collision_margin, float beta,
                       float smooth_weight, float step, int32_t iterations, float* out_cost, void* workspace,
                       size_t workspace_bytes, uint64_t* work_flops);

/* Environment point cloud for edge validation.  The reference's point-cloud planner validates
 * its edges against a cloud on pyroffi's OptiX checker, each point a sphere of radius r_env
 * (prrtc_robogpu.py:103-298, README.md:107-148); that checker is absent from the reference tree,
 * so this one is defined here: a robot sphere (c, r) hits point p iff |c - p| <= r + r_env.
 * The cloud lives in a uniform grid of cubic cells of edge `cell` over [lo, hi] (points outside
 * go to the nearest border cell), points sorted by cell; prrtc_cloud_build fills this struct,
 * whose pointers are DEVICE memory inside the caller's buffer. */
typedef struct {
    const float* points;       /* [n_points, 4] (x, y, z, 0), sorted by cell */
    const int32_t* cell_start; /* [nx * ny * nz + 1]: cell (x, y, z) = (z ny + y) nx + x holds
                                  points[cell_start[c] .. cell_start[c + 1]) */
    int32_t n_points;
    int32_t nx, ny, nz;
    float lo[3];
    float cell;
    float r_env;
} prrtc_cloud;

/* Bytes of device memory prrtc_cloud_build needs (0 for invalid arguments).  lo, hi: HOST
 * float[3], the grid's bounds; the grid may hold at most 2^24 cells. */
size_t prrtc_cloud_bytes(int32_t n_points, const float* lo, const float* hi, float cell);

/* Builds the grid of points [n_points, 3] (device) into `memory` on `stream` (count, scan and
 * scatter kernels, no host synchronisation) and fills *out.  No reference counterpart (it stands
 * where pyroffi's OptiX BVH build stood). */
int prrtc_cloud_build(prrtc_stream_t stream, const float* points, int32_t n_points, const float* lo, const float* hi,
                      float cell, float r_env, void* memory, size_t memory_bytes, prrtc_cloud* out);

/* Edge validation, the call prrtc_robogpu.py:186-188 makes of the OptiX checker
 * (check_edges_collision_free): waypoints [n_edges, granularity, dim]; out_free [n_edges] uint8 is
 * 1 iff every waypoint of the edge is free of the world primitives and self pairs (the verdict of
 * collision_mode, as prrtc_check_configs) and, when cloud is not NULL, of the cloud. */
int prrtc_check_edges(prrtc_stream_t stream, const float* waypoints, int32_t n_edges, int32_t granularity, int32_t dim,
                      const prrtc_robot* robot, const prrtc_world* world, int32_t collision_mode,
                      float collision_margin, const prrtc_cloud* cloud, uint8_t* out_free, const prrtc_launch* launch,
                      uint64_t* work_flops);

/* Message of the last failing call on this thread. */
const char* prrtc_last_error(void);

/* Library version string. */
const char* prrtc_version(void);

#ifdef __cplusplus
}
#endif
#endif

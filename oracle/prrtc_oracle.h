/* prrtc_oracle.h -- CPU oracle for the pRRTC planner path.  TEST INFRASTRUCTURE ONLY:
 * loaded by tests/, __graft_entry__.smoke() and bench.py's cpu_baseline leg, never by the
 * product library. */
#ifndef PRRTC_ORACLE_H
#define PRRTC_ORACLE_H
#include <stdint.h>

#define OC_MAX_JOINTS 64
#define OC_MAX_DIM 16

#define OC_STATUS_SOLVED 1
#define OC_STATUS_ITER_CAP 2
#define OC_STATUS_TREE_FULL 3

typedef struct {
    int n_joints, dim;
    const float* twists;        /* [n_j,6]  (v, omega) */
    const float* parent_tf;     /* [n_j,7]  (w,x,y,z,tx,ty,tz) */
    const int* parent_idx;      /* [n_j] */
    const int* act_idx;         /* [n_j] */
    const float* mimic_mul;     /* [n_j] */
    const float* mimic_off;     /* [n_j] */
    const int* mimic_act_idx;   /* [n_j] */
    const int* topo_inv;        /* [n_j] */
    int n_spheres;
    const int* sphere_link_idx; /* [n_rs] */
    const float* sphere_local;  /* [n_rs,3] */
    const float* sphere_radius; /* [n_rs] */
    int n_world_spheres;
    const float* world_spheres; /* [.,4] */
    int n_world_capsules;
    const float* world_capsules;/* [.,7] */
    int n_world_boxes;
    const float* world_boxes;   /* [.,15] */
    int n_world_halfspaces;
    const float* world_halfspaces; /* [.,6] */
    int n_self_pairs;
    const int* self_pairs;      /* [.,2] */
    int collision_mode;         /* 0 binary, 1 margin */
    float collision_margin;
} oc_context;

typedef struct {
    int max_iterations;
    float step_size;
    int num_new_samples;
    int balance_mode;
    float tree_ratio;
    int dynamic_domain;
    float dd_alpha, dd_radius, dd_min_radius;
    int dim;
    int max_nodes;
    int granularity;
    unsigned seed;
} oc_params;

typedef struct {
    float* tree_a;   /* [dim, max_nodes] */
    float* tree_b;
    int* parents_a;  /* [max_nodes] */
    int* parents_b;
    float* radii_a;  /* [max_nodes] */
    float* radii_b;
    int sizes[2];
    int iterations;
    int conn[3];
    int status;
    long long fkcc_evals;
} oc_outputs;

void oc_sincos(float x, float* s, float* c);
void oc_fk(const oc_context* c, const float* q, float* T_out);
int oc_enabled(const oc_context* c);
int oc_in_collision(const oc_context* c, const float* q);
void oc_in_collision_batch(const oc_context* c, const float* qs, int n, uint8_t* out);
float oc_min_distance(const oc_context* c, const float* q);
void oc_min_distance_batch(const oc_context* c, const float* qs, int n, float* out);
void oc_sphere_positions(const oc_context* c, const float* q, float* out);
int oc_cloud_hit(const oc_context* c, const float* q, const float* points, int n_points, float r_env);
void oc_cloud_hit_batch(const oc_context* c, const float* qs, int n, const float* points, int n_points, float r_env,
                        uint8_t* out);
float oc_halton(int base, int index);
void oc_prime_order(int explorer, unsigned seed, int* out16);
void oc_nearest(const float* tree, int max_nodes, int size, int dim, const float* qs, int nq, int* idx, float* dsq);
void oc_dd_shrink(float* r, float alpha, float r0, float rmin);
void oc_dd_grow(float* r, float alpha);
int oc_dd_skips(float radius, float nd);
int oc_balance_side(int mode, int round, int e, int S, int size_a, int size_b, float tree_ratio, int* vamp);
void oc_steer(const float* base, const float* sample, int dim, float nd, float step, float* out);
int oc_connect_steps(const float* from, const float* to, int dim, float dist, float step, float* vec);
int oc_plan(const oc_context* c, const oc_params* p, const float* start, const float* goals, int num_goals,
            const float* min_vals, const float* max_vals, oc_outputs* out);
int oc_trace_path(const float* tree_a, const float* tree_b, const int* par_a, const int* par_b, int max_nodes,
                  int dim, const int* conn, float* path, int max_len);
#endif

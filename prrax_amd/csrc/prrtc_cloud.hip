// prrtc_cloud.hip -- environment point clouds for edge validation, and their C ABI:
//   prrtc_cloud_bytes / prrtc_cloud_build   a uniform grid over the cloud, its points sorted by cell
//   prrtc_check_edges                       which edges [n_edges, G, dim] are free of the world,
//                                           the robot's self pairs and the cloud
//
// The reference's point-cloud planner drives RRT-Connect from the host and validates each
// round's edges against an environment point cloud on pyroffi's OptiX checker, the cloud's
// points taken as spheres of radius r_env (cuda-rrtc/jax/prrtc_robogpu.py:103-298,
// README.md:107-148).  That checker is not in the reference tree; here a robot sphere (c, r) hits
// point p iff |c - p| <= r + r_env, in the square-root-free form of the sphere test
// (d^2 <= rr |rr|, rr = r + r_env), the same expression the world spheres use.
//
// Grid: cells of edge h over the box [lo, lo + n h) per axis; a point outside goes to the
// nearest border cell (the query clamps the same way, so nothing is lost).  Cell (x, y, z) has
// index (z ny + y) nx + x, so the cells of one row x0..x1 hold one contiguous run of the sorted
// points: a query visits at most (y1 - y0 + 1)(z1 - z0 + 1) runs.  The grid is built on the
// device (count, scan, scatter), with no host synchronisation.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cmath>
#include <cstring>

#include "prrtc_amd.h"
#include "prrtc_cloud.h"
#include "prrtc_common.h"
#include "prrtc_host.h"

namespace prrtc {

constexpr long long kMaxCells = 1ll << 24;
constexpr int kScanThreads = 1024;
constexpr int kScanPer = 4;  // counts per thread of a scan block
constexpr int kScanBlock = kScanThreads * kScanPer;

__global__ void cloud_fill_kernel(int* a, long long n, int v) {
    for (long long i = static_cast<long long>(blockIdx.x) * blockDim.x + threadIdx.x; i < n;
         i += static_cast<long long>(gridDim.x) * blockDim.x)
        a[i] = v;
}

__global__ void cloud_count_kernel(const float* p, int n, CloudGrid g, int* counts) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        global_add(counts + cell_of(g, p[3 * i], p[3 * i + 1], p[3 * i + 2]), 1);
}

// Exclusive scan of counts[0, n) into start, one block of kScanBlock counts per work-group;
// bsum[block] receives the block's total.
__global__ __launch_bounds__(kScanThreads) void cloud_scan_blocks_kernel(const int* counts, int n, int* start,
                                                                          int* bsum) {
    __shared__ int scratch[20];
    const int base = blockIdx.x * kScanBlock + threadIdx.x * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = (base + k < n) ? counts[base + k] : 0;
        sum += v[k];
    }
    int total;
    int run = block_excl_scan(sum, scratch, &total);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < n) start[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) bsum[blockIdx.x] = total;
}

// Exclusive scan of the nb <= kScanBlock block totals in place (one work-group); start[n] = total.
__global__ __launch_bounds__(kScanThreads) void cloud_scan_sums_kernel(int* bsum, int nb, int* start, int n) {
    __shared__ int scratch[20];
    const int base = threadIdx.x * kScanPer;
    int v[kScanPer];
    int sum = 0;
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        v[k] = (base + k < nb) ? bsum[base + k] : 0;
        sum += v[k];
    }
    int total;
    int run = block_excl_scan(sum, scratch, &total);
#pragma unroll
    for (int k = 0; k < kScanPer; ++k) {
        if (base + k < nb) bsum[base + k] = run;
        run += v[k];
    }
    if (threadIdx.x == 0) start[n] = total;
}

__global__ void cloud_scan_add_kernel(int* start, int n, const int* bsum) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x)
        start[i] += bsum[i / kScanBlock];
}

// Points to their cell's run; the order inside a run follows the atomics (a verdict is an OR
// over the run, so it does not depend on that order).
__global__ void cloud_scatter_kernel(const float* p, int n, CloudGrid g, const int* start, int* cursor, float4* out) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) {
        const float x = p[3 * i], y = p[3 * i + 1], z = p[3 * i + 2];
        const int c = cell_of(g, x, y, z);
        out[start[c] + global_add(cursor + c, 1)] = make_float4(x, y, z, 0.0f);
    }
}

struct EdgeArgs {
    RobotDev robot;
    WorldDev world;
    ModelLayout ml;
    size_t off_scan;
    const float* waypoints;  // [n_edges * G, dim]
    int n_edges, G, dim, mode;
    float margin;
    CloudDev cloud;
    int has_cloud;
    int tile;                // waypoints per wave batch (FK one per lane), 1..kFB
    int parts;               // waves per waypoint when tile == 1 (they split its cloud rows), else 1
    uint8_t* out_free;       // [n_edges], 1 on entry
    unsigned long long* flops;
};

__global__ void edge_init_kernel(uint8_t* out_free, int n) {
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < n; i += gridDim.x * blockDim.x) out_free[i] = 1;
}

// Waves take batches of `tile` waypoints (FK one per lane), then check them one by one with the
// whole wave: the world and self pairs (the planner's culled verdict), then the cloud.  An edge
// already found colliding skips its remaining waypoints.  Small calls (a planner round's edges)
// take tiles of fewer waypoints, so that they still spread over the GPU's waves; the smallest
// give each waypoint `parts` waves, which share its cloud rows (part 0 also tests the world).
constexpr int kMaxParts = 16;

__global__ __launch_bounds__(1024) void edge_check_kernel(EdgeArgs a) {
    extern __shared__ __align__(16) char lds[];
    int* scan = reinterpret_cast<int*>(lds + a.off_scan);
    Model M;
    stage_model(M, lds, a.ml, a.robot, a.world, 0, a.dim, a.mode, a.margin, scan);
    const int lane = threadIdx.x & (kWave - 1);
    const WaveArea W = wave_area(lds, a.ml, M.n_joints);
    LaneScratch L;
    L.q = W.q + lane;
    L.T = W.T + lane;
    L.nt = kFBP;
    const int wpb = blockDim.x / kWave;
    const int total_waves = gridDim.x * wpb;
    const long long n = static_cast<long long>(a.n_edges) * a.G;
    const int tile = a.tile, parts = a.parts;
    const long long nb = (n + tile - 1) / tile * parts;
    unsigned long long flops = 0;
    for (long long bp = static_cast<long long>(blockIdx.x) * wpb + threadIdx.x / kWave; bp < nb; bp += total_waves) {
        const long long b = bp / parts;
        const int part = static_cast<int>(bp - b * parts);
        const long long i = b * tile + lane;
        if (lane < tile && i < n) {
            for (int d = 0; d < a.dim; ++d) L.q[d * kFBP] = a.waypoints[i * a.dim + d];
            if (M.n_joints > 0) fk_lane(M, L);
        }
        wave_sync();
        const int nc = static_cast<int>(min(static_cast<long long>(tile), n - b * tile));
        for (int c = 0; c < nc; ++c) {
            const int e = static_cast<int>((b * tile + c) / a.G);
            // lane 0's reading of the edge's flag decides for the wave
            const int open = __builtin_amdgcn_readfirstlane(
                static_cast<int>(*reinterpret_cast<volatile const uint8_t*>(a.out_free + e)));
            if (open == 0) continue;
            bool h = part == 0 && wave_in_collision(M, W.T + c, kFBP, flops);
            if (!h && a.has_cloud && M.n_joints > 0) h = wave_cloud_hit(M, W.T + c, kFBP, a.cloud, flops, part, parts);
            if (h && lane == 0) a.out_free[e] = 0;
        }
        wave_sync();
    }
    if (lane == 0 && a.flops && flops) global_add(a.flops, flops);
}

struct GridShape {
    int nx, ny, nz;
    long long cells;
};

int grid_shape(const float* lo, const float* hi, float cell, GridShape& s) {
    if (!lo || !hi) return fail_with(PRRTC_E_INVALID, "cloud bounds are null");
    if (!(cell > 0.0f) || !std::isfinite(cell)) return fail_with(PRRTC_E_INVALID, "cell must be a finite size > 0");
    long long n[3];
    for (int k = 0; k < 3; ++k) {
        if (!std::isfinite(lo[k]) || !std::isfinite(hi[k]) || !(hi[k] >= lo[k]))
            return fail_with(PRRTC_E_INVALID, "cloud bounds must be finite with hi >= lo");
        const double c = std::ceil((static_cast<double>(hi[k]) - lo[k]) / cell);
        if (c > static_cast<double>(kMaxCells)) return fail_with(PRRTC_E_INVALID, "cloud grid has too many cells");
        n[k] = std::max(1ll, static_cast<long long>(c));
    }
    s.nx = static_cast<int>(n[0]);
    s.ny = static_cast<int>(n[1]);
    s.nz = static_cast<int>(n[2]);
    s.cells = n[0] * n[1] * n[2];
    if (s.cells > kMaxCells)
        return fail_with(PRRTC_E_INVALID, "cloud grid has more than 2^24 cells: use a larger cell");
    return PRRTC_OK;
}

struct CloudMem {
    size_t pts, start, cursor, bsum, total;
};

CloudMem cloud_mem(long long n_points, long long cells) {
    const auto a256 = [](size_t x) { return (x + 255) & ~static_cast<size_t>(255); };
    CloudMem m;
    size_t o = 0;
    m.pts = o; o = a256(o + sizeof(float4) * static_cast<size_t>(n_points));
    m.start = o; o = a256(o + sizeof(int) * static_cast<size_t>(cells + 1));
    m.cursor = o; o = a256(o + sizeof(int) * static_cast<size_t>(cells));
    m.bsum = o; o = a256(o + sizeof(int) * static_cast<size_t>((cells + kScanBlock - 1) / kScanBlock));
    m.total = o;
    return m;
}

}  // namespace prrtc

using namespace prrtc;

extern "C" {

size_t prrtc_cloud_bytes(int32_t n_points, const float* lo, const float* hi, float cell) {
    GridShape s;
    if (n_points < 0 || grid_shape(lo, hi, cell, s)) return 0;
    return cloud_mem(n_points, s.cells).total;
}

int prrtc_cloud_build(prrtc_stream_t stream, const float* points, int32_t n_points, const float* lo, const float* hi,
                      float cell, float r_env, void* memory, size_t memory_bytes, prrtc_cloud* out) {
    if (!out) return fail_with(PRRTC_E_INVALID, "out is null");
    if (n_points < 0 || n_points > (1 << 30)) return fail_with(PRRTC_E_INVALID, "bad n_points");
    if (n_points > 0 && !points) return fail_with(PRRTC_E_INVALID, "points is null");
    if (!std::isfinite(r_env)) return fail_with(PRRTC_E_INVALID, "r_env must be finite");
    GridShape s;
    int rc = grid_shape(lo, hi, cell, s);
    if (rc) return rc;
    const CloudMem m = cloud_mem(n_points, s.cells);
    if (!memory || memory_bytes < m.total) return fail_with(PRRTC_E_WORKSPACE, "cloud memory too small");
    char* base = static_cast<char*>(memory);
    float4* pts = reinterpret_cast<float4*>(base + m.pts);
    int* start = reinterpret_cast<int*>(base + m.start);
    int* cursor = reinterpret_cast<int*>(base + m.cursor);
    int* bsum = reinterpret_cast<int*>(base + m.bsum);
    CloudGrid g;
    g.lx = lo[0]; g.ly = lo[1]; g.lz = lo[2];
    g.inv = 1.0f / cell;
    g.nx = s.nx; g.ny = s.ny; g.nz = s.nz;
    const int cells = static_cast<int>(s.cells);
    const int nblk = (cells + kScanBlock - 1) / kScanBlock;
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int fill_grid = static_cast<int>(std::min<long long>((s.cells + 255) / 256, 4096));
    const int pt_grid = std::max(1, std::min((n_points + 255) / 256, 4096));
    hipLaunchKernelGGL(cloud_fill_kernel, dim3(fill_grid), dim3(256), 0, st, cursor, s.cells, 0);
    if (n_points > 0) hipLaunchKernelGGL(cloud_count_kernel, dim3(pt_grid), dim3(256), 0, st, points, n_points, g, cursor);
    hipLaunchKernelGGL(cloud_scan_blocks_kernel, dim3(nblk), dim3(kScanThreads), 0, st, cursor, cells, start, bsum);
    hipLaunchKernelGGL(cloud_scan_sums_kernel, dim3(1), dim3(kScanThreads), 0, st, bsum, nblk, start, cells);
    hipLaunchKernelGGL(cloud_scan_add_kernel, dim3(fill_grid), dim3(256), 0, st, start, cells, bsum);
    hipLaunchKernelGGL(cloud_fill_kernel, dim3(fill_grid), dim3(256), 0, st, cursor, s.cells, 0);
    if (n_points > 0)
        hipLaunchKernelGGL(cloud_scatter_kernel, dim3(pt_grid), dim3(256), 0, st, points, n_points, g, start, cursor, pts);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    out->points = reinterpret_cast<const float*>(pts);
    out->cell_start = start;
    out->n_points = n_points;
    out->nx = s.nx; out->ny = s.ny; out->nz = s.nz;
    out->lo[0] = lo[0]; out->lo[1] = lo[1]; out->lo[2] = lo[2];
    out->cell = cell;
    out->r_env = r_env;
    return PRRTC_OK;
}

int prrtc_check_edges(prrtc_stream_t stream, const float* waypoints, int32_t n_edges, int32_t granularity, int32_t dim,
                      const prrtc_robot* robot, const prrtc_world* world, int32_t collision_mode,
                      float collision_margin, const prrtc_cloud* cloud, uint8_t* out_free, const prrtc_launch* launch,
                      uint64_t* work_flops) {
    int rc = check_robot(robot);
    if (rc) return rc;
    rc = check_world(world);
    if (rc) return rc;
    if (n_edges < 0 || granularity < 1 || dim < 1 || dim > PRRTC_MAX_DIM)
        return fail_with(PRRTC_E_INVALID, "bad n_edges, granularity or dim");
    if (static_cast<long long>(n_edges) * granularity > (1ll << 31) - 1)
        return fail_with(PRRTC_E_INVALID, "too many waypoints");
    if (collision_mode < 0 || collision_mode > 3) return fail_with(PRRTC_E_INVALID, "collision_mode must be 0..3");
    if (n_edges == 0) return PRRTC_OK;
    if (!waypoints || !out_free) return fail_with(PRRTC_E_INVALID, "waypoints / out_free is null");
    if (cloud) {
        const long long cells = static_cast<long long>(cloud->nx) * cloud->ny * cloud->nz;
        if (cloud->nx < 1 || cloud->ny < 1 || cloud->nz < 1 || cells > kMaxCells || cloud->n_points < 0 ||
            !(cloud->cell > 0.0f) || !std::isfinite(cloud->r_env) || !cloud->cell_start ||
            (cloud->n_points > 0 && !cloud->points))
            return fail_with(PRRTC_E_INVALID, "malformed prrtc_cloud (build it with prrtc_cloud_build)");
    }
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return fail_with(PRRTC_E_HIP, "hipGetDevice failed");
    const DevInfo& di = dev_info(dev);
    int nt = (launch && launch->threads > 0) ? launch->threads : 256;
    if (nt % kWave || nt > 1024) return fail_with(PRRTC_E_INVALID, "threads must be a multiple of 64, <= 1024");
    EdgeArgs a;
    std::memset(&a, 0, sizeof(a));
    size_t lds = 0;
    rc = model_lds(*robot, *world, nt, di.max_lds, a.ml, lds, false);
    if (rc) return rc;
    a.robot = robot_dev(*robot);
    a.world = world_dev(*world, false);
    a.off_scan = scan_offset(a.ml);
    a.waypoints = waypoints;
    a.n_edges = n_edges;
    a.G = granularity;
    a.dim = dim;
    a.mode = (collision_mode == 1) ? 1 : 0;
    a.margin = collision_margin;
    a.out_free = out_free;
    a.flops = reinterpret_cast<unsigned long long*>(work_flops);
    if (cloud && cloud->n_points > 0) {
        a.has_cloud = 1;
        a.cloud.pts = reinterpret_cast<const float4*>(cloud->points);
        a.cloud.start = cloud->cell_start;
        a.cloud.g.lx = cloud->lo[0]; a.cloud.g.ly = cloud->lo[1]; a.cloud.g.lz = cloud->lo[2];
        a.cloud.g.inv = 1.0f / cloud->cell;
        a.cloud.g.nx = cloud->nx; a.cloud.g.ny = cloud->ny; a.cloud.g.nz = cloud->nz;
        a.cloud.r_env = cloud->r_env;
    }
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(edge_init_kernel, dim3(std::max(1, std::min((n_edges + 255) / 256, 4096))), dim3(256), 0, st,
                       out_free, n_edges);
    // waypoints per wave batch: the largest power of two <= kFB that gives every CU 16 batches
    const long long nwp = static_cast<long long>(n_edges) * granularity;
    int tile = kFB;
    while (tile > 1 && (nwp + tile - 1) / tile < 16ll * di.cus) tile /= 2;
    a.tile = tile;
    int parts = 1;
    if (tile == 1 && a.has_cloud)
        while (parts < kMaxParts && nwp * parts * 2 <= 16ll * di.cus) parts *= 2;
    a.parts = parts;
    const long long batches = (nwp + tile - 1) / tile * parts;
    long long grid = (launch && launch->slots > 0) ? launch->slots : (batches + nt / kWave - 1) / (nt / kWave);
    grid = std::max(1ll, std::min(grid, static_cast<long long>(16 * di.cus)));
    hipError_t e = hipFuncSetAttribute(reinterpret_cast<const void*>(&edge_check_kernel),
                                       hipFuncAttributeMaxDynamicSharedMemorySize, static_cast<int>(lds));
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    hipLaunchKernelGGL(edge_check_kernel, dim3(static_cast<int>(grid)), dim3(nt), lds, st, a);
    e = hipGetLastError();
    if (e != hipSuccess) return fail_with(PRRTC_E_HIP, hipGetErrorString(e));
    return PRRTC_OK;
}

}  // extern "C"

// prrtc_cloud.h -- device side of the point-cloud check: the grid's cell arithmetic and the
// wave-level test of one configuration's robot spheres against the cloud (prrtc_cloud.hip holds
// the grid build and the edge-check kernel).
#pragma once
#include "prrtc_common.h"

namespace prrtc {

struct CloudGrid {
    float lx, ly, lz, inv;
    int nx, ny, nz;
};

struct CloudDev {
    const float4* pts;
    const int* start;
    CloudGrid g;
    float r_env;
};

__device__ __forceinline__ int cell_axis(float v, float lo, float inv, int n) {
    // clamped as a float first, so that a huge or NaN coordinate never reaches the conversion
    const float f = floorf((v - lo) * inv);
    return static_cast<int>(fminf(fmaxf(f, 0.0f), static_cast<float>(n - 1)));
}

__device__ __forceinline__ int cell_of(const CloudGrid& g, float x, float y, float z) {
    return (cell_axis(z, g.lz, g.inv, g.nz) * g.ny + cell_axis(y, g.ly, g.inv, g.ny)) * g.nx +
           cell_axis(x, g.lx, g.inv, g.nx);
}

// The cloud rows a sphere p (world frame, radius p.w) can reach: rows (y, z) for y0 <= y < y0 +
// nyr and z0 <= z, n of them in all (0 when the radius sum with r_env is negative).  The reach
// r + r_env is widened by 1e-5 m plus 1e-6 of the coordinates' magnitude, far above the rounding
// of the cell arithmetic, so every point within the reach lies in a visited cell.
struct CloudRows {
    float reach;
    int y0, z0, nyr, n;
};

__device__ __forceinline__ CloudRows cloud_rows(const float4 p, const CloudDev& C) {
    CloudRows R = {0.0f, 0, 0, 1, 0};
    const CloudGrid& g = C.g;
    const float rr = p.w + C.r_env;
    if (rr < 0.0f) return R;
    R.reach = rr + 1e-5f + 1e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
    R.y0 = cell_axis(p.y - R.reach, g.ly, g.inv, g.ny);
    R.z0 = cell_axis(p.z - R.reach, g.lz, g.inv, g.nz);
    const int y1 = cell_axis(p.y + R.reach, g.ly, g.inv, g.ny), z1 = cell_axis(p.z + R.reach, g.lz, g.inv, g.nz);
    R.nyr = y1 - R.y0 + 1;
    R.n = R.nyr * (z1 - R.z0 + 1);
    return R;
}

// The run [b, e) of sorted cloud points of row r of R that can lie within the reach: empty when
// the row's y-z rectangle lies beyond it, else the row's cells clipped in x to the reach's chord
// at that distance.  Border cells reach to infinity (they hold the clamped points).
__device__ __forceinline__ void cloud_run(const float4 p, const CloudRows& R, int r, const CloudDev& C, int& b,
                                          int& e) {
    const CloudGrid& g = C.g;
    const float h = 1.0f / g.inv;
    const int y = R.y0 + r % R.nyr, z = R.z0 + r / R.nyr;
    const float zl = (z == 0) ? -FLT_MAX : g.lz + static_cast<float>(z) * h;
    const float zh = (z == g.nz - 1) ? FLT_MAX : g.lz + static_cast<float>(z + 1) * h;
    const float yl = (y == 0) ? -FLT_MAX : g.ly + static_cast<float>(y) * h;
    const float yh = (y == g.ny - 1) ? FLT_MAX : g.ly + static_cast<float>(y + 1) * h;
    const float dz = fmaxf(fmaxf(zl - p.z, p.z - zh), 0.0f);
    const float dy = fmaxf(fmaxf(yl - p.y, p.y - yh), 0.0f);
    const float rest = R.reach * R.reach - (dy * dy + dz * dz);
    b = e = 0;
    if (rest >= 0.0f) {
        const float half = __builtin_sqrtf(rest) * 1.000001f + 1e-6f;
        const int x0 = cell_axis(p.x - half, g.lx, g.inv, g.nx);
        const int x1 = cell_axis(p.x + half, g.lx, g.inv, g.nx);
        const int row = (z * g.ny + y) * g.nx;
        b = C.start[row + x0];
        e = C.start[row + x1 + 1];
    }
}

// The lane j of the wave holding item t of a wave-wide list whose inclusive prefix counts are
// incl (nondecreasing over the lanes, t < the last lane's): the smallest j with incl_j > t.
// Every lane of the wave must call it.
__device__ __forceinline__ int wave_owner(int incl, int t) {
    int j = 0;
#pragma unroll
    for (int step = kWave / 2; step > 0; step >>= 1)
        if (__shfl(incl, j + step - 1, kWave) <= t) j += step;
    return j;
}

constexpr int kRunCap = 1 << 24;  // points of one run per scan, so that 64 runs' sum fits an int

// Whether any robot sphere of the configuration at T hits a point of the cloud (wave-uniform).
// The robot spheres are transformed lane-parallel into the wave's cull area.  Then the (sphere,
// row) items of up to 64 spheres are dealt to the lanes, 64 at a time: each lane finds its row's
// clipped run of sorted points, and the wave scans the concatenation of the 64 runs, 128 points
// per step (each lane finds its points' runs by a binary search over the runs' prefix counts), with
// one ballot per step, so a contact ends the check at once.  Every point of every run is tested
// until a contact: the verdict is whether any point within a sphere's reach exists.  With
// parts > 1, waves part = 0 .. parts - 1 test the same configuration and each takes every
// parts-th group of 64 items; the configuration hits iff one of them finds a contact.
__device__ bool wave_cloud_hit(const Model& M, const float* T, int ts, const CloudDev& C, unsigned long long& flops,
                               int part = 0, int parts = 1) {
    const int lane = threadIdx.x & (kWave - 1);
    float4* FP = reinterpret_cast<float4*>(M.cull_base + static_cast<size_t>(threadIdx.x / kWave) * M.cull_stride);
    for (int s0 = 0; s0 < M.n_spheres; s0 += kWave) {
        const int s = s0 + lane;
        bool act = false;
        if (s < M.n_spheres) {
            const int js = M.sphere_js[s];
            float4 o = make_float4(0.0f, 0.0f, 0.0f, -1.0f);
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
    wave_sync();
    unsigned tests = 0;
    bool hit = false;
    for (int s0 = 0; s0 < M.n_spheres && !hit; s0 += kWave) {
        // lane k counts the rows of sphere s0 + k
        int nrows = 0;
        if (s0 + lane < M.n_spheres && M.sphere_js[s0 + lane] >= 0) nrows = cloud_rows(FP[s0 + lane], C).n;
        const int incl = wave_incl_scan(nrows);
        const int excl = incl - nrows;
        const int total = __shfl(incl, kWave - 1, kWave);
        for (int t0 = part * kWave; t0 < total && !hit; t0 += parts * kWave) {
            // item t0 + lane: row r of sphere s0 + j
            const int t = min(t0 + lane, total - 1);
            const int j = wave_owner(incl, t);
            const int r = t - __shfl(excl, j, kWave);
            const int sph = s0 + j;
            int b = 0, e = 0;
            if (t0 + lane < total) {
                const float4 p = FP[sph];
                cloud_run(p, cloud_rows(p, C), r, C, b, e);
            }
            for (;;) {
                const int len = min(e - b, kRunCap);
                const int incl2 = wave_incl_scan(len);
                const int base = b - (incl2 - len);  // point of concatenated position k: base_j + k
                const int tot2 = __shfl(incl2, kWave - 1, kWave);
                for (int k0 = 0; k0 < tot2; k0 += 2 * kWave) {
                    // two points per lane per step: positions k0 + lane and k0 + 64 + lane
                    const int ka = min(k0 + lane, tot2 - 1), kb = min(k0 + kWave + lane, tot2 - 1);
                    const int ja = wave_owner(incl2, ka), jb = wave_owner(incl2, kb);
                    const int ia = __shfl(base, ja, kWave) + ka, ib = __shfl(base, jb, kWave) + kb;
                    const int sa = __shfl(sph, ja, kWave), sb = __shfl(sph, jb, kWave);
                    // clamped positions are valid points: both loads issue together, unmasked
                    const float4 qa = C.pts[ia], qb = C.pts[ib];
                    const float4 pa = FP[sa], pb = FP[sb];
                    const bool ha = hit_sphere(pa.x, pa.y, pa.z, pa.w, make_float4(qa.x, qa.y, qa.z, C.r_env));
                    const bool hb = hit_sphere(pb.x, pb.y, pb.z, pb.w, make_float4(qb.x, qb.y, qb.z, C.r_env));
                    const bool hk = (ha & (k0 + lane < tot2)) | (hb & (k0 + kWave + lane < tot2));
                    tests += static_cast<unsigned>(min(2 * kWave, tot2 - k0));
                    if (wave_any(hk)) {
                        hit = true;
                        break;
                    }
                }
                b += len;
                if (hit || !wave_any(e > b)) break;
            }
        }
    }
    flops += static_cast<unsigned long long>(tests) * kFlopsSS;
    wave_sync();
    return hit;
}

}  // namespace prrtc

// prrtc_cloud.h -- device side of the point-cloud check: the grid's cell arithmetic and the
// wave-level test of one configuration's robot spheres against the cloud (prrtc_cloud.hip holds
// the grid build and the edge-check kernel).
#pragma once
#include "prrtc_device.h"

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

// Whether any robot sphere of the configuration at T hits a point of the cloud (wave-uniform).
// The robot spheres are transformed lane-parallel into the wave's cull area; then, sphere by
// sphere, the lanes find the runs of the cloud rows the sphere can reach (one row per lane: a row
// (y, z) is skipped when its cells' y-z rectangle lies beyond the reach, and otherwise clipped in
// x to the reach's chord at that distance), and the whole wave scans the non-empty runs, 64
// consecutive points per step (coalesced loads) with one ballot per step, so a contact ends the
// check at once.  The reach r + r_env is widened by 1e-5 m plus 1e-6 of
// the coordinates' magnitude, far above the rounding of the cell arithmetic, so every point within
// the reach lies in a visited cell.
__device__ bool wave_cloud_hit(const Model& M, const float* T, int ts, const CloudDev& C, unsigned long long& flops) {
    const int lane = threadIdx.x & (kWave - 1);
    const CloudGrid& g = C.g;
    float4* FP = reinterpret_cast<float4*>(M.cull_base + static_cast<size_t>(threadIdx.x / kWave) * M.cull_stride);
    for (int s0 = 0; s0 < M.n_spheres; s0 += kWave) {
        const int s = s0 + lane;
        bool act = false;
        if (s < M.n_spheres) {
            const int js = M.sphere_js[s];
            float4 o = make_float4(0.0f, 0.0f, 0.0f, -1.0f);  // w < 0: no reach (skipped below)
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
    const float h = 1.0f / g.inv;
    unsigned tests = 0;
    bool hit = false;
    for (int s = 0; s < M.n_spheres && !hit; ++s) {
        if (M.sphere_js[s] < 0) continue;
        const float4 p = FP[s];
        const float rr = p.w + C.r_env;
        if (rr < 0.0f) continue;
        const float reach = rr + 1e-5f + 1e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
        const int y0 = cell_axis(p.y - reach, g.ly, g.inv, g.ny), y1 = cell_axis(p.y + reach, g.ly, g.inv, g.ny);
        const int z0 = cell_axis(p.z - reach, g.lz, g.inv, g.nz), z1 = cell_axis(p.z + reach, g.lz, g.inv, g.nz);
        const int nyr = y1 - y0 + 1, nrows = nyr * (z1 - z0 + 1);
        const float r2 = reach * reach;
        // lanes over the sphere's rows: each finds its row's clipped run [b, e) of sorted points
        for (int r0 = 0; r0 < nrows && !hit; r0 += kWave) {
            const int r = r0 + lane;
            int b = 0, e = 0;
            if (r < nrows) {
                const int y = y0 + r % nyr, z = z0 + r / nyr;
                // distance from the centre to the row's y-z rectangle (0 inside); border cells
                // reach to infinity
                const float zl = (z == 0) ? -FLT_MAX : g.lz + static_cast<float>(z) * h;
                const float zh = (z == g.nz - 1) ? FLT_MAX : g.lz + static_cast<float>(z + 1) * h;
                const float yl = (y == 0) ? -FLT_MAX : g.ly + static_cast<float>(y) * h;
                const float yh = (y == g.ny - 1) ? FLT_MAX : g.ly + static_cast<float>(y + 1) * h;
                const float dz = fmaxf(fmaxf(zl - p.z, p.z - zh), 0.0f);
                const float dy = fmaxf(fmaxf(yl - p.y, p.y - yh), 0.0f);
                const float rest = r2 - (dy * dy + dz * dz);
                if (rest >= 0.0f) {
                    const float half = __builtin_sqrtf(rest) * 1.000001f + 1e-6f;
                    const int x0 = cell_axis(p.x - half, g.lx, g.inv, g.nx);
                    const int x1 = cell_axis(p.x + half, g.lx, g.inv, g.nx);
                    const int row = (z * g.ny + y) * g.nx;
                    b = C.start[row + x0];
                    e = C.start[row + x1 + 1];
                }
            }
            // then the wave scans the non-empty runs one after another, 64 points per step
            unsigned long long rows = __ballot(e > b);
            while (rows != 0ull && !hit) {
                const int j = static_cast<int>(__builtin_ctzll(rows));
                rows &= rows - 1ull;
                const int rb = __shfl(b, j, kWave), re = __shfl(e, j, kWave);
                for (int k0 = rb; k0 < re; k0 += kWave) {
                    const int k = k0 + lane;
                    bool hk = false;
                    if (k < re) {
                        const float4 q = C.pts[k];
                        hk = hit_sphere(p.x, p.y, p.z, p.w, make_float4(q.x, q.y, q.z, C.r_env));
                    }
                    tests += static_cast<unsigned>(min(kWave, re - k0));
                    if (wave_any(hk)) {
                        hit = true;
                        break;
                    }
                }
            }
        }
    }
    flops += static_cast<unsigned long long>(tests) * kFlopsSS;
    wave_sync();
    return hit;
}

}  // namespace prrtc

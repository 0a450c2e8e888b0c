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
// Lane l takes robot spheres l, l + 64, ...: its sphere's reach r + r_env, widened by 1e-5 m plus
// 1e-6 of the coordinates' magnitude (far above the rounding of the cell arithmetic, so every
// point within the reach lies in a visited cell), bounds the cells it visits.
__device__ bool wave_cloud_hit(const Model& M, const float* T, int ts, const CloudDev& C, unsigned long long& flops) {
    const int lane = threadIdx.x & (kWave - 1);
    const CloudGrid& g = C.g;
    unsigned tests = 0;
    bool any = false;
    for (int s0 = 0; s0 < M.n_spheres && !any; s0 += kWave) {
        const int s = s0 + lane;
        bool h = false;
        bool act = false;
        if (s < M.n_spheres) {
            const int js = M.sphere_js[s];
            const float4 sp = M.spheres[s];
            const float rr = sp.w + C.r_env;
            if (js >= 0 && rr >= 0.0f) {
                act = true;
                Mat3 R;
                float tx, ty, tz;
                frame_from(T, ts, js, R, tx, ty, tz);
                const float3 p = mat_apply(R, tx, ty, tz, sp.x, sp.y, sp.z);
                const float reach = rr + 1e-5f + 1e-6f * (fabsf(p.x) + fabsf(p.y) + fabsf(p.z));
                const int x0 = cell_axis(p.x - reach, g.lx, g.inv, g.nx), x1 = cell_axis(p.x + reach, g.lx, g.inv, g.nx);
                const int y0 = cell_axis(p.y - reach, g.ly, g.inv, g.ny), y1 = cell_axis(p.y + reach, g.ly, g.inv, g.ny);
                const int z0 = cell_axis(p.z - reach, g.lz, g.inv, g.nz), z1 = cell_axis(p.z + reach, g.lz, g.inv, g.nz);
                const float4 me = make_float4(p.x, p.y, p.z, sp.w);
                for (int z = z0; z <= z1 && !h; ++z) {
                    for (int y = y0; y <= y1 && !h; ++y) {
                        const int row = (z * g.ny + y) * g.nx;
                        const int b = C.start[row + x0], e = C.start[row + x1 + 1];
                        for (int k = b; k < e; ++k) {
                            const float4 q = C.pts[k];
                            ++tests;
                            if (hit_sphere(me.x, me.y, me.z, me.w, make_float4(q.x, q.y, q.z, C.r_env))) {
                                h = true;
                                break;
                            }
                        }
                    }
                }
            }
        }
        flops += static_cast<unsigned long long>(__popcll(__ballot(act))) * kFlopsXf;
        any = wave_any(h);
    }
    flops += static_cast<unsigned long long>(wave_sum_u32(tests)) * kFlopsSS;
    return any;
}

}  // namespace prrtc

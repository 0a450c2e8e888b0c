// CPU build of the planner's wave-level collision checks (prrtc_device.h / prrtc_common.h), for
// bounds checking under AddressSanitizer.  One work-group of one wave: 64 threads run the device
// code of stage_model, fk_lane, wave_min_distance (flat sweep), wave_in_collision (culled) and
// wave_unit_in_collision (sub-units), exactly as the kernels call them, over the configurations
// of an input file.  The LDS of the work-group is one heap block of the model layout's size.
//
//   simt_check <input> <output>
//
// Input: little-endian records, each an int64 element count followed by the elements, in this
// order: fk_twists f32, fk_parent_tf f32, fk_parent_idx i32, fk_act_idx i32, fk_mimic_mul f32,
// fk_mimic_off f32, fk_mimic_act_idx i32, fk_topo_inv i32, sphere_link_idx i32, sphere_local f32,
// sphere_radius f32, world_spheres f32, world_capsules f32, world_boxes f32, world_halfspaces
// f32, self_pairs i32, params i32 [dim, mode, n_configs, n_units], margin f32 [1], configs f32
// [n_configs * dim], units i32 [n_units * 2] (first configuration, count; a unit lies inside one
// batch of 16 consecutive configurations).
// Output: hit_culled u8[n], min_distance f32[n], unit verdict i32[n_units],
// observed maxima i64[6] (culled queue depth, unit candidate list length, unit refinement queue
// depth, groups, surviving self pairs of a unit, units whose survivors overflowed the list).
#include <hip/hip_runtime.h>

#include <cstring>
#include <fstream>
#include <string>
#include <thread>
#include <vector>

#include "prrtc_common.h"

thread_local simt_dim3 threadIdx;
thread_local simt_dim3 blockIdx;
simt_dim3 blockDim;
simt_dim3 gridDim;

namespace simt {
static Wave g_wave;
static thread_local int g_lane = 0;
static std::atomic<long long> g_observed[6];
Wave& wave() { return g_wave; }
int lane() { return g_lane; }
void observe(int slot, long long value) {
    if (slot == 5) {  // a count, not a maximum
        g_observed[5] += value;
        return;
    }
    long long cur = g_observed[slot].load();
    while (value > cur && !g_observed[slot].compare_exchange_weak(cur, value)) {
    }
}
}  // namespace simt

namespace {

struct Reader {
    std::ifstream f;
    template <class T>
    std::vector<T> next() {
        int64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        if (!f || n < 0) {
            std::fprintf(stderr, "simt_check: truncated input\n");
            std::exit(2);
        }
        std::vector<T> v(static_cast<size_t>(n));
        if (n) f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
        if (!f) {
            std::fprintf(stderr, "simt_check: truncated input\n");
            std::exit(2);
        }
        return v;
    }
};

template <class T>
const T* ptr(const std::vector<T>& v) {
    return v.empty() ? nullptr : v.data();
}

}  // namespace

int main(int argc, char** argv) {
    using namespace prrtc;
    if (argc != 3) {
        std::fprintf(stderr, "usage: simt_check <input> <output>\n");
        return 2;
    }
    Reader rd{std::ifstream(argv[1], std::ios::binary)};
    const auto tw = rd.next<float>(), ptf = rd.next<float>();
    const auto par = rd.next<int>(), act = rd.next<int>();
    const auto mmul = rd.next<float>(), moff = rd.next<float>();
    const auto mact = rd.next<int>(), topo = rd.next<int>(), slink = rd.next<int>();
    const auto sloc = rd.next<float>(), srad = rd.next<float>();
    const auto ws = rd.next<float>(), wc = rd.next<float>(), wb = rd.next<float>(), wh = rd.next<float>();
    const auto sp = rd.next<int>();
    const auto prm = rd.next<int>();
    const auto mg = rd.next<float>();
    const auto cfg = rd.next<float>();
    const auto units = rd.next<int>();
    if (prm.size() != 4 || mg.size() != 1) {
        std::fprintf(stderr, "simt_check: bad parameter records\n");
        return 2;
    }
    const int dim = prm[0], mode = prm[1], n = prm[2], nun = prm[3];
    const int nj = static_cast<int>(par.size()), nrs = static_cast<int>(slink.size());
    if (dim < 1 || dim > kMaxDim || nj > kMaxJoints || cfg.size() != static_cast<size_t>(n) * dim ||
        units.size() != static_cast<size_t>(nun) * 2) {
        std::fprintf(stderr, "simt_check: inconsistent sizes\n");
        return 2;
    }
    for (int k = 0; k < nun; ++k) {
        const int c0 = units[2 * k], nu = units[2 * k + 1];
        if (nu < 1 || nu > kUnitMax || c0 < 0 || c0 + nu > n || (c0 % kFB) + nu > kFB) {
            std::fprintf(stderr, "simt_check: unit %d (%d, %d) does not lie inside one batch\n", k, c0, nu);
            return 2;
        }
    }
    RobotDev r{ptr(tw), ptr(ptf), ptr(par), ptr(act), ptr(mmul), ptr(moff), ptr(mact), ptr(topo), nj,
               ptr(slink), ptr(sloc), ptr(srad), nrs};
    WorldDev w{};
    w.ws = ptr(ws); w.wc = ptr(wc); w.wb = ptr(wb); w.wh = ptr(wh); w.sp = ptr(sp);
    w.n_ws = static_cast<int>(ws.size() / 4); w.n_wc = static_cast<int>(wc.size() / 7);
    w.n_wb = static_cast<int>(wb.size() / 15); w.n_wh = static_cast<int>(wh.size() / 6);
    w.n_sp = static_cast<int>(sp.size() / 2);
    const ModelLayout ml = model_layout(nj, nrs, w.n_sp, w.n_ws, w.n_wc, w.n_wb, w.n_wh, 1, kFBP, kUnitCap);
    std::vector<char> lds(ml.end);  // the work-group's dynamic LDS, exactly sized
    std::vector<int> scan(17);
    std::vector<uint8_t> hit_culled(n);
    std::vector<float> min_d(n);
    std::vector<int> unit_v(nun);
    blockDim.x = simt::kLanes;
    gridDim.x = 1;
    auto body = [&](int lane) {
        simt::g_lane = lane;
        threadIdx.x = static_cast<unsigned>(lane);
        blockIdx.x = 0;
        Model M;
        stage_model(M, lds.data(), ml, r, w, 0, dim, mode, mg[0], scan.data());
        simt::observe(3, M.n_groups);
        const WaveArea W = wave_area(lds.data(), ml, M.n_joints);
        LaneScratch L;
        L.q = W.q + lane;
        L.T = W.T + lane;
        L.nt = kFBP;
        unsigned long long flops = 0;
        int next_unit = 0;
        for (int b0 = 0; b0 < n; b0 += kFB) {
            const int i = b0 + lane;
            if (lane < kFB && i < n) {
                for (int d = 0; d < dim; ++d) L.q[d * kFBP] = cfg[static_cast<size_t>(i) * dim + d];
                if (M.n_joints > 0) fk_lane(M, L);
            }
            wave_sync();
            const int nc = min(kFB, n - b0);
            for (int c = 0; c < nc; ++c) {
                const float md = wave_min_distance(M, W.T + c, kFBP, flops);
                const bool hc = wave_in_collision(M, W.T + c, kFBP, flops);
                if (lane == 0) {
                    min_d[b0 + c] = md;
                    hit_culled[b0 + c] = hc ? 1 : 0;
                }
            }
            while (next_unit < nun && units[2 * next_unit] < b0 + kFB) {
                const int c0 = units[2 * next_unit] - b0, nu = units[2 * next_unit + 1];
                const int v = wave_unit_in_collision(M, W.T + c0, kFBP, nu, flops);
                if (lane == 0) unit_v[next_unit] = v;
                ++next_unit;
            }
            wave_sync();
        }
    };
    for (int k = 1; k < nun; ++k)
        if (units[2 * k] < units[2 * (k - 1)]) {
            std::fprintf(stderr, "simt_check: units must be sorted by first configuration\n");
            return 2;
        }
    std::vector<std::thread> lanes;
    lanes.reserve(simt::kLanes);
    for (int l = 0; l < simt::kLanes; ++l) lanes.emplace_back(body, l);
    for (auto& t : lanes) t.join();
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(hit_culled.data()), n);
    out.write(reinterpret_cast<const char*>(min_d.data()), static_cast<std::streamsize>(4 * n));
    out.write(reinterpret_cast<const char*>(unit_v.data()), static_cast<std::streamsize>(4 * nun));
    for (int s = 0; s < 6; ++s) {
        const long long v = simt::g_observed[s].load();
        out.write(reinterpret_cast<const char*>(&v), 8);
    }
    return out ? 0 : 1;
}

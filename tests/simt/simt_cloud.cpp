// CPU build of the point-cloud check (prrtc_cloud.h) for bounds checking under AddressSanitizer:
// one wave of 64 threads runs stage_model, fk_lane, wave_in_collision and wave_cloud_hit over
// the configurations of an input file, against a grid built here on the host with the device's
// cell arithmetic (cell_of), the points and cell starts in arrays of exactly their size.
//
//   simt_cloud <input> <output>
//
// Input records (int64 count, then elements): the robot and world records of simt_check, then
// params i32 [dim, mode, n_configs, nx, ny, nz], floats f32 [margin, lo_x, lo_y, lo_z, cell,
// r_env], points f32 [n_points * 3], configs f32 [n_configs * dim].
// Output: world verdict u8[n], cloud verdict u8[n] (computed for every configuration).  Exit
// status 3 when the cloud verdict split over three waves' shares of the rows differs from it
// (checked for every 4th configuration).
#include <hip/hip_runtime.h>

#include <cstdio>
#include <fstream>
#include <thread>
#include <vector>

#include "prrtc_cloud.h"
#include "prrtc_common.h"

thread_local simt_dim3 threadIdx;
thread_local simt_dim3 blockIdx;
simt_dim3 blockDim;
simt_dim3 gridDim;

namespace simt {
static Wave g_wave;
static thread_local int g_lane = 0;
Wave& wave() { return g_wave; }
int lane() { return g_lane; }
void observe(int, long long) {}
}  // namespace simt

namespace {

struct Reader {
    std::ifstream f;
    template <class T>
    std::vector<T> next() {
        int64_t n = 0;
        f.read(reinterpret_cast<char*>(&n), 8);
        if (!f || n < 0) {
            std::fprintf(stderr, "simt_cloud: truncated input\n");
            std::exit(2);
        }
        std::vector<T> v(static_cast<size_t>(n));
        if (n) f.read(reinterpret_cast<char*>(v.data()), static_cast<std::streamsize>(n * sizeof(T)));
        if (!f) {
            std::fprintf(stderr, "simt_cloud: truncated input\n");
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
        std::fprintf(stderr, "usage: simt_cloud <input> <output>\n");
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
    const auto fl = rd.next<float>();
    const auto pts_in = rd.next<float>();
    const auto cfg = rd.next<float>();
    if (prm.size() != 6 || fl.size() != 6 || pts_in.size() % 3) {
        std::fprintf(stderr, "simt_cloud: bad parameter records\n");
        return 2;
    }
    const int dim = prm[0], mode = prm[1], n = prm[2];
    const int nj = static_cast<int>(par.size()), nrs = static_cast<int>(slink.size());
    if (dim < 1 || dim > kMaxDim || nj > kMaxJoints || cfg.size() != static_cast<size_t>(n) * dim || prm[3] < 1 ||
        prm[4] < 1 || prm[5] < 1 || static_cast<long long>(prm[3]) * prm[4] * prm[5] > (1 << 24)) {
        std::fprintf(stderr, "simt_cloud: inconsistent sizes\n");
        return 2;
    }
    // the grid, as the build kernels make it (count, exclusive scan, scatter)
    CloudGrid g;
    g.nx = prm[3]; g.ny = prm[4]; g.nz = prm[5];
    g.lx = fl[1]; g.ly = fl[2]; g.lz = fl[3];
    g.inv = 1.0f / fl[4];
    const int np = static_cast<int>(pts_in.size() / 3);
    const int cells = g.nx * g.ny * g.nz;
    std::vector<int> cnt(cells, 0), start(cells + 1, 0);
    for (int i = 0; i < np; ++i) ++cnt[cell_of(g, pts_in[3 * i], pts_in[3 * i + 1], pts_in[3 * i + 2])];
    for (int c = 0; c < cells; ++c) start[c + 1] = start[c] + cnt[c];
    std::vector<float4> pts(np);
    std::vector<int> cur(cells, 0);
    for (int i = 0; i < np; ++i) {
        const int c = cell_of(g, pts_in[3 * i], pts_in[3 * i + 1], pts_in[3 * i + 2]);
        pts[start[c] + cur[c]++] = make_float4(pts_in[3 * i], pts_in[3 * i + 1], pts_in[3 * i + 2], 0.0f);
    }
    CloudDev C;
    C.pts = ptr(pts);
    C.start = start.data();
    C.g = g;
    C.r_env = fl[5];
    RobotDev r{ptr(tw), ptr(ptf), ptr(par), ptr(act), ptr(mmul), ptr(moff), ptr(mact), ptr(topo), nj,
               ptr(slink), ptr(sloc), ptr(srad), nrs};
    WorldDev w{};
    w.ws = ptr(ws); w.wc = ptr(wc); w.wb = ptr(wb); w.wh = ptr(wh); w.sp = ptr(sp);
    w.n_ws = static_cast<int>(ws.size() / 4); w.n_wc = static_cast<int>(wc.size() / 7);
    w.n_wb = static_cast<int>(wb.size() / 15); w.n_wh = static_cast<int>(wh.size() / 6);
    w.n_sp = static_cast<int>(sp.size() / 2);
    const ModelLayout ml = model_layout(nj, nrs, w.n_sp, w.n_ws, w.n_wc, w.n_wb, w.n_wh, 1, kFBP);
    std::vector<char> lds(ml.end);
    std::vector<int> scan(17);
    std::vector<uint8_t> hit_world(n), hit_cloud(n);
    constexpr int kParts = 3;
    bool split_mismatch = false;
    blockDim.x = simt::kLanes;
    gridDim.x = 1;
    auto body = [&](int lane) {
        simt::g_lane = lane;
        threadIdx.x = static_cast<unsigned>(lane);
        Model M;
        stage_model(M, lds.data(), ml, r, w, 0, dim, mode, fl[0], scan.data());
        const WaveArea W = wave_area(lds.data(), ml, M.n_joints);
        LaneScratch L;
        L.q = W.q + lane;
        L.T = W.T + lane;
        L.nt = kFBP;
        unsigned long long flops = 0;
        for (int b0 = 0; b0 < n; b0 += kFB) {
            const int i = b0 + lane;
            if (lane < kFB && i < n) {
                for (int d = 0; d < dim; ++d) L.q[d * kFBP] = cfg[static_cast<size_t>(i) * dim + d];
                if (M.n_joints > 0) fk_lane(M, L);
            }
            wave_sync();
            const int nc = min(kFB, n - b0);
            for (int c = 0; c < nc; ++c) {
                const bool hw = wave_in_collision(M, W.T + c, kFBP, flops);
                const bool hc = (M.n_joints > 0 && np > 0) ? wave_cloud_hit(M, W.T + c, kFBP, C, flops) : false;
                // every 4th configuration again, split over kParts waves (run here one after another)
                bool hs = hc;
                if ((b0 + c) % 4 == 0 && M.n_joints > 0 && np > 0) {
                    hs = false;
                    for (int part = 0; part < kParts; ++part)
                        hs = wave_cloud_hit(M, W.T + c, kFBP, C, flops, part, kParts) || hs;
                }
                if (lane == 0 && hs != hc) split_mismatch = true;
                if (lane == 0) {
                    hit_world[b0 + c] = hw ? 1 : 0;
                    hit_cloud[b0 + c] = hc ? 1 : 0;
                }
            }
            wave_sync();
        }
    };
    std::vector<std::thread> lanes;
    for (int l = 0; l < simt::kLanes; ++l) lanes.emplace_back(body, l);
    for (auto& t : lanes) t.join();
    if (split_mismatch) {
        std::fprintf(stderr, "cloud verdict split over %d parts differs from the whole\n", kParts);
        return 3;
    }
    std::ofstream out(argv[2], std::ios::binary);
    out.write(reinterpret_cast<const char*>(hit_world.data()), n);
    out.write(reinterpret_cast<const char*>(hit_cloud.data()), n);
    return out ? 0 : 1;
}

// Self-test of the host wave stand-in: each mode plants one mistake the bounds checks must catch.
//   harness_selftest ok        lockstep ballot and shuffles (exit 0)
//   harness_selftest diverge   lanes reach different ballots (abort: divergent collective)
//   harness_selftest skip      half the lanes skip a ballot (abort: only 32 of 64 lanes arrived)
//   harness_selftest overflow  a queue push past its capacity (abort: assertion failed)
//   harness_selftest heap      a write one element past a heap block (AddressSanitizer report)
#include <hip/hip_runtime.h>

#include <cstring>
#include <thread>
#include <vector>

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

int main(int argc, char** argv) {
    const char* mode = argc > 1 ? argv[1] : "ok";
    simt::wave().timeout_s = 2;
    std::vector<int> queue(8);
    std::vector<int>* heap = new std::vector<int>(64);
    unsigned long long seen = 0;
    auto body = [&](int lane) {
        simt::g_lane = lane;
        threadIdx.x = static_cast<unsigned>(lane);
        const unsigned long long m = __ballot(lane % 3 == 0);
        const int s = __shfl_xor(lane, 1, 64);
        if (lane == 0) seen = m ^ static_cast<unsigned long long>(s != 1);
        if (!std::strcmp(mode, "diverge")) {
            if (lane < 32) (void)__ballot(true);
            else (void)__ballot(false);
        } else if (!std::strcmp(mode, "skip")) {
            if (lane < 32) (void)__ballot(true);
        } else if (!std::strcmp(mode, "overflow")) {
            const int qn = static_cast<int>(__popcll(__ballot(true)));
            PRRTC_SIMT_ASSERT(qn <= static_cast<int>(queue.size()));
        } else if (!std::strcmp(mode, "heap")) {
            if (lane == 63) heap->data()[lane + 1] = 1;
        }
    };
    std::vector<std::thread> t;
    for (int l = 0; l < simt::kLanes; ++l) t.emplace_back(body, l);
    for (auto& x : t) x.join();
    delete heap;
    return seen == 0x9249249249249249ull ? 0 : 3;
}

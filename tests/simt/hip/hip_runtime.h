// Host stand-in for <hip/hip_runtime.h>, used only to build the planner's device headers with g++
// for the CPU bounds checks in tests/simt (never for the library).  One work-group of 64 lanes
// (one wave) runs as 64 threads; every cross-lane operation (ballot, shuffles, readfirstlane,
// barriers) is a rendezvous of all 64 threads, which reproduces the lockstep semantics the device
// code assumes when its control flow is wave-uniform.  A rendezvous checks that every lane
// reached the same source line (a divergent collective aborts with both lines) and aborts when a
// lane never arrives (a collective skipped by some lanes).
#pragma once

#include <float.h>
#include <math.h>
#include <stddef.h>
#include <stdint.h>

#include <algorithm>
#include <atomic>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <mutex>
#include <thread>

#define __device__
#define __host__
#define __global__
#define __forceinline__ inline
#define __launch_bounds__(...)
#define __shared__
#define __align__(n) alignas(n)

struct simt_dim3 {
    unsigned x = 0, y = 0, z = 0;
};
extern thread_local simt_dim3 threadIdx;
extern thread_local simt_dim3 blockIdx;
extern simt_dim3 blockDim;
extern simt_dim3 gridDim;

struct float3 {
    float x, y, z;
};
struct alignas(16) float4 {
    float x, y, z, w;
};
inline float3 make_float3(float x, float y, float z) { return float3{x, y, z}; }
inline float4 make_float4(float x, float y, float z, float w) { return float4{x, y, z, w}; }

inline int min(int a, int b) { return a < b ? a : b; }
inline int max(int a, int b) { return a > b ? a : b; }
inline unsigned min(unsigned a, unsigned b) { return a < b ? a : b; }
inline unsigned max(unsigned a, unsigned b) { return a > b ? a : b; }
inline long long min(long long a, long long b) { return a < b ? a : b; }
inline long long max(long long a, long long b) { return a > b ? a : b; }

inline int __popcll(unsigned long long v) { return __builtin_popcountll(v); }
inline float __builtin_amdgcn_sqrtf(float x) { return sqrtf(x); }

namespace simt {

constexpr int kLanes = 64;

// Rendezvous of the 64 lanes: each lane deposits a 64-bit word, and all lanes see all words.
// Arrivals count on an atomic; the last one checks the lines and opens the next generation, the
// others spin (yielding the CPU) until it does, or abort after timeout_s seconds.
struct Wave {
    std::atomic<int> arrived{0};
    std::atomic<unsigned long long> generation{0};
    unsigned long long words[2][kLanes] = {};
    int lines[2][kLanes] = {};
    int timeout_s = 20;
};
Wave& wave();
int lane();

// Deposit `word` (tag: the source line), wait for every lane, return the buffer of all words.
inline const unsigned long long* exchange(unsigned long long word, int line) {
    Wave& W = wave();
    const unsigned long long gen = W.generation.load(std::memory_order_acquire);
    const int buf = static_cast<int>(gen & 1ull);
    W.words[buf][lane()] = word;
    W.lines[buf][lane()] = line;
    if (W.arrived.fetch_add(1, std::memory_order_acq_rel) + 1 == kLanes) {
        for (int l = 1; l < kLanes; ++l)
            if (W.lines[buf][l] != W.lines[buf][0]) {
                std::fprintf(stderr, "simt: divergent collective: lane 0 at line %d, lane %d at line %d\n",
                             W.lines[buf][0], l, W.lines[buf][l]);
                std::abort();
            }
        W.arrived.store(0, std::memory_order_relaxed);
        W.generation.store(gen + 1, std::memory_order_release);
    } else {
        const auto t0 = std::chrono::steady_clock::now();
        for (unsigned spins = 0; W.generation.load(std::memory_order_acquire) == gen; ++spins) {
            std::this_thread::yield();
            if ((spins & 4095u) == 4095u &&
                std::chrono::steady_clock::now() - t0 > std::chrono::seconds(W.timeout_s)) {
                std::fprintf(stderr, "simt: collective at line %d: only %d of %d lanes arrived\n", line,
                             W.arrived.load(), kLanes);
                std::abort();
            }
        }
    }
    // Buffer `buf` is rewritten only by the collective after next, which no lane can enter
    // before every lane has left this one.
    return W.words[buf];
}

inline unsigned long long ballot(bool p, int line) {
    const unsigned long long* w = exchange(p ? 1ull : 0ull, line);
    unsigned long long m = 0;
    for (int l = 0; l < kLanes; ++l) m |= (w[l] & 1ull) << l;
    return m;
}

template <class T>
inline unsigned long long bits_of(T v) {
    static_assert(sizeof(T) <= 8, "shuffle of a value wider than 64 bits");
    unsigned long long b = 0;
    __builtin_memcpy(&b, &v, sizeof(T));
    return b;
}
template <class T>
inline T from_bits(unsigned long long b) {
    T v;
    __builtin_memcpy(&v, &b, sizeof(T));
    return v;
}

template <class T>
inline T shfl_xor(T v, int mask, int width, int line) {
    const unsigned long long* w = exchange(bits_of(v), line);
    const int l = lane();
    const int src = ((l ^ mask) & (width - 1)) | (l & ~(width - 1));
    return from_bits<T>(w[src]);
}
template <class T>
inline T shfl_up(T v, int delta, int width, int line) {
    const unsigned long long* w = exchange(bits_of(v), line);
    const int l = lane();
    const int base = l & ~(width - 1);
    return (l - base) >= delta ? from_bits<T>(w[l - delta]) : v;
}
template <class T>
inline T shfl(T v, int src, int width, int line) {
    const unsigned long long* w = exchange(bits_of(v), line);
    const int l = lane();
    return from_bits<T>(w[(l & ~(width - 1)) | (src & (width - 1))]);
}
template <class T>
inline T readfirstlane(T v, int line) {
    return from_bits<T>(exchange(bits_of(v), line)[0]);
}
inline void barrier(int line) { (void)exchange(0ull, line); }

template <class T>
inline T fetch_min(T* p, T v) {
    T cur = __atomic_load_n(p, __ATOMIC_RELAXED);
    while (v < cur && !__atomic_compare_exchange_n(p, &cur, v, false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)) {
    }
    return cur;
}

}  // namespace simt

#define __ballot(p) simt::ballot(static_cast<bool>(p), __LINE__)
#define __shfl_xor(v, m, w) simt::shfl_xor((v), (m), (w), __LINE__)
#define __shfl_up(v, d, w) simt::shfl_up((v), (d), (w), __LINE__)
#define __shfl(v, s, w) simt::shfl((v), (s), (w), __LINE__)
#define __builtin_amdgcn_readfirstlane(v) simt::readfirstlane((v), __LINE__)
#define __syncthreads() simt::barrier(__LINE__)
#define __builtin_amdgcn_wave_barrier() simt::barrier(__LINE__)
#define __builtin_amdgcn_fence(order, scope) std::atomic_thread_fence(std::memory_order_seq_cst)

#define __HIP_MEMORY_SCOPE_WORKGROUP 0
#define __HIP_MEMORY_SCOPE_AGENT 0
#define __hip_atomic_fetch_add(p, v, order, scope) __atomic_fetch_add((p), (v), __ATOMIC_RELAXED)
#define __hip_atomic_fetch_min(p, v, order, scope) simt::fetch_min((p), (v))
#define __hip_atomic_compare_exchange_strong(p, e, d, o1, o2, scope) \
    __atomic_compare_exchange_n((p), (e), (d), false, __ATOMIC_RELAXED, __ATOMIC_RELAXED)
#define __hip_atomic_load(p, order, scope) __atomic_load_n((p), __ATOMIC_RELAXED)
#define __hip_atomic_store(p, v, order, scope) __atomic_store_n((p), (v), __ATOMIC_RELAXED)

// Bounds checks of the device code's LDS queues and packed entries (no code on the device).
#define PRRTC_SIMT_ASSERT(cond)                                                                 \
    do {                                                                                        \
        if (!(cond)) {                                                                          \
            std::fprintf(stderr, "simt: %s:%d: assertion failed: %s\n", __FILE__, __LINE__, #cond); \
            std::abort();                                                                       \
        }                                                                                       \
    } while (0)
namespace simt {
void observe(int slot, long long value);
}
#define PRRTC_SIMT_OBSERVE(slot, value) simt::observe((slot), static_cast<long long>(value))

// common.hpp -- shared device/host helpers for the MaZe HIP library (gfx950).
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <climits>
#include <cstdint>
#include <cstdio>
#include <cstdlib>
#include <stdexcept>
#include <string>

namespace maze {

#define MAZE_HIP(call)                                                                   \
    do {                                                                                 \
        hipError_t e_ = (call);                                                          \
        if (e_ != hipSuccess) {                                                          \
            std::fprintf(stderr, "maze: HIP error %s at %s:%d: %s\n", hipGetErrorName(e_), \
                         __FILE__, __LINE__, #call);                                     \
            std::abort();                                                                \
        }                                                                                \
    } while (0)

// Kernel timing: when `start` is set, the next MAZE_LAUNCH records the
// kernel's own start and end in (start, stop) -- the dispatch's timestamps,
// the same interval a kernel trace reports -- and clears `start`, so a
// launcher that issues several kernels times only its first.
struct LaunchTimer {
    hipEvent_t start = nullptr, stop = nullptr;
};
extern thread_local LaunchTimer g_launch_timer;

#define MAZE_LAUNCH(kernel, grid, block, shmem, stream, ...)                                   \
    do {                                                                                       \
        if (::maze::g_launch_timer.start) {                                                    \
            const hipEvent_t a_ = ::maze::g_launch_timer.start;                                \
            ::maze::g_launch_timer.start = nullptr;                                            \
            hipExtLaunchKernelGGL(kernel, grid, block, shmem, stream, a_,                      \
                                  ::maze::g_launch_timer.stop, 0, __VA_ARGS__);                \
        } else {                                                                               \
            hipLaunchKernelGGL(kernel, grid, block, shmem, stream, __VA_ARGS__);               \
        }                                                                                      \
    } while (0)

constexpr int kBlock = 256;       // threads per work-group (4 wave64)
constexpr int kWave = 64;
constexpr double kPi = 3.14159265358979323846;

// Geometry of a periodic n^3 grid stored C-order [i][j][k], k fastest.
struct GridDims {
    int n;        // points per side
    int n2;       // n*n (points per i-plane)
    int chunks;   // ceil(n2 / kBlock): plane chunks, one per work-group column
    int slab;     // i-planes marched by one work-group
    int slabs;    // ceil(n / slab)
    int blocks;   // chunks * slabs
};

inline GridDims make_dims(int n, int slab) {
    GridDims d;
    d.n = n;
    d.n2 = n * n;
    d.chunks = (d.n2 + kBlock - 1) / kBlock;
    d.slab = slab < n ? slab : n;
    d.slabs = (n + d.slab - 1) / d.slab;
    d.blocks = d.chunks * d.slabs;
    return d;
}

// Work-group -> (chunk, slab) map that keeps neighbouring slabs on one XCD.
// Consecutive work-group ids are dispatched round-robin over the 8 XCDs, each
// with its own L2; regrouping ids so that one XCD walks a contiguous range of
// (slab, chunk) tiles keeps the i+-1 halo planes in the same L2.
__device__ __forceinline__ int xcd_tile(int b, int nblocks) {
    const int per = nblocks >> 3;
    if ((nblocks & 7) != 0) return b;
    return (b & 7) * per + (b >> 3);
}

// Work-group barrier that orders LDS only: the fences are restricted to the
// local address space, so the barrier waits for this wave's LDS operations
// (lgkmcnt) but not for its outstanding global loads and stores, which a
// __syncthreads() (workgroup fences over all address spaces) drains.  For
// kernels whose barriers only protect LDS and whose global loads are issued
// ahead of their use.
__device__ __forceinline__ void lds_barrier() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "workgroup", "local");
    __builtin_amdgcn_s_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "workgroup", "local");
}

// A load of solver state on the vector memory path (an agent-scope relaxed
// atomic load): the compiler turns a plain load of a uniform address into a
// scalar load, and a kernel that needs its scalar kernel arguments before it
// can issue more loads would then wait for the state's round trip as well.
template <class T>
__device__ __forceinline__ T vload_state(const T* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// x^3 rounded once (the reference's pow(x, 3)): x*x split exactly into
// hi + lo by an FMA, then hi*x + lo*x with one final rounding.
__device__ __forceinline__ double cube(double x) {
    const double hi = x * x;
    const double lo = fma(x, x, -hi);
    return fma(hi, x, lo * x);
}

// First index in the ascending a[0..n) whose value is >= key.
__device__ __forceinline__ int lower_bound_u32(const unsigned int* a, int n, unsigned int key) {
    int lo = 0, hi = n;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (a[mid] < key) lo = mid + 1;
        else hi = mid;
    }
    return lo;
}

// Neighbour offsets of a point inside its i-plane (periodic in j and k).
struct PlaneNb {
    int jm, jp, km, kp;
};

__device__ __forceinline__ PlaneNb plane_nb(int pl, int n, int n2) {
    const int j = pl / n;
    const int k = pl - j * n;
    PlaneNb o;
    o.jm = (j == 0) ? pl + n2 - n : pl - n;
    o.jp = (j == n - 1) ? pl - n2 + n : pl + n;
    o.km = (k == 0) ? pl + n - 1 : pl - 1;
    o.kp = (k == n - 1) ? pl - n + 1 : pl + 1;
    return o;
}

__device__ __forceinline__ double wave_sum(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v += __shfl_xor(v, o);
    return v;
}

__device__ __forceinline__ double wave_max(double v) {
#pragma unroll
    for (int o = kWave / 2; o > 0; o >>= 1) v = fmax(v, __shfl_xor(v, o));
    return v;
}

// Fixed-shape work-group reductions (deterministic: butterfly within each
// wave, then the NT/64 wave results in wave order).  Result valid in all
// threads; sh holds NT/64 doubles.
template <int NT = kBlock>
__device__ __forceinline__ double block_sum(double v, double* sh) {
    v = wave_sum(v);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < NT / kWave; i++) s += sh[i];
    return s;
}

__device__ __forceinline__ double block_max(double v, double* sh) {
    v = wave_max(v);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.0;
    for (int i = 0; i < kBlock / kWave; i++) s = fmax(s, sh[i]);
    return s;
}

// Sum and max of one value pair in a single LDS exchange.  sh holds 8 doubles.
__device__ __forceinline__ void block_sum_max(double& s, double& m, double* sh) {
    s = wave_sum(s);
    m = wave_max(m);
    const int w = threadIdx.x / kWave;
    __syncthreads();
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sh[w] = s;
        sh[4 + w] = m;
    }
    __syncthreads();
    double a = 0.0, b = 0.0;
    for (int i = 0; i < kBlock / kWave; i++) {
        a += sh[i];
        b = fmax(b, sh[4 + i]);
    }
    s = a;
    m = b;
}

// Per-work-group partials are exchanged through device-scope atomic loads and
// stores, which are performed at the level shared by all XCDs (their L2
// caches are not coherent with each other).
__device__ __forceinline__ void store_partial(double* p, int i, double v) {
    __hip_atomic_store(p + i, v, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ double load_partial(const double* p, int i) {
    return __hip_atomic_load(const_cast<double*>(p) + i, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Single-pass reduction hand-off: thread 0 publishes the work-group's
// partial(s), waits until those device-scope stores have completed, then
// arrives; the work-group that arrives last reduces all partials in index
// order (so the result does not depend on arrival order).  Only the partials
// need to be visible to the last work-group, so no device-wide release fence
// (an L2 write-back on this chip) is issued.
//
// Arrival is two-level: work-group b bumps the counter of its group b % 8
// (consecutive ids go round-robin to the 8 XCDs, so a group is one XCD's
// share), and the last of each group bumps the top counter.  All work-groups
// of a sweep finish within a short time, and device-scope atomics on one
// address are performed one after another, so a single counter serialises
// one atomic per work-group at the end of every reducing kernel; eight group
// counters on their own cache lines take them side by side.  The chain of
// returned counts orders every group's stores before the last arrival.
// counter: kArrivalWords words (the top counter, then one per group, each on
// its own 128-byte line), all zero between kernels.
constexpr int kArrivalStride = 32;
constexpr int kArrivalGroups = 8;
constexpr int kArrivalWords = (kArrivalGroups + 1) * kArrivalStride;

__device__ __forceinline__ bool publish_arrive(double* psum, double s, double* pmax, double m,
                                               unsigned int* counter, unsigned int nblocks) {
    __shared__ bool last;
    if (threadIdx.x == 0) {
        store_partial(psum, blockIdx.x, s);
        if (pmax) store_partial(pmax, blockIdx.x, m);
        __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
        const unsigned int g = blockIdx.x % kArrivalGroups;
        const unsigned int groups = nblocks < kArrivalGroups ? nblocks : kArrivalGroups;
        const unsigned int members = (nblocks - g + kArrivalGroups - 1) / kArrivalGroups;
        const unsigned int prev = __hip_atomic_fetch_add(counter + kArrivalStride * (1 + g), 1u,
                                                         __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        bool l = false;
        if (prev == members - 1) {
            const unsigned int top =
                __hip_atomic_fetch_add(counter, 1u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            l = (top == groups - 1);
        }
        last = l;
    }
    __syncthreads();
    return last;
}

// Zeroes the counters of publish_arrive (called by the last work-group, when
// every arrival of the kernel has been counted).
__device__ __forceinline__ void reset_arrivals(unsigned int* counter) {
    for (int g = 0; g <= kArrivalGroups; g++)
        __hip_atomic_store(counter + kArrivalStride * g, 0u, __ATOMIC_RELAXED,
                           __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ void reset_counter(unsigned int* counter) {
    __hip_atomic_store(counter, 0u, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

// Sum / max of partials[0..cnt) in a fixed order; valid in all threads.
// All loads of a thread are issued before they are combined (cnt <= 8 NT
// in one round trip).
template <int NT = kBlock>
__device__ __forceinline__ double reduce_partials_sum(const double* partials, int cnt, double* sh) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int i = threadIdx.x + u * NT;
        v[u] = i < cnt ? load_partial(partials, i) : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) s += v[u];
    for (int i = threadIdx.x + 8 * NT; i < cnt; i += NT) s += load_partial(partials, i);
    return block_sum<NT>(s, sh);
}

__device__ __forceinline__ double reduce_partials_max(const double* partials, int cnt, double* sh) {
    double v[8];
#pragma unroll
    for (int u = 0; u < 8; u++) {
        const int i = threadIdx.x + u * kBlock;
        v[u] = i < cnt ? load_partial(partials, i) : 0.0;
    }
    double s = 0.0;
#pragma unroll
    for (int u = 0; u < 8; u++) s = fmax(s, v[u]);
    for (int i = threadIdx.x + 8 * kBlock; i < cnt; i += kBlock) s = fmax(s, load_partial(partials, i));
    return block_max(s, sh);
}

// Both folds with the loads of the two arrays issued together.
__device__ __forceinline__ void reduce_partials_sum_max(const double* ps, const double* pm, int cnt,
                                                        double* sh, double& S, double& M) {
    double a[4], b[4];
#pragma unroll
    for (int u = 0; u < 4; u++) {
        const int i = threadIdx.x + u * kBlock;
        a[u] = i < cnt ? load_partial(ps, i) : 0.0;
        b[u] = i < cnt ? load_partial(pm, i) : 0.0;
    }
    double s = 0.0, m = 0.0;
#pragma unroll
    for (int u = 0; u < 4; u++) {
        s += a[u];
        m = fmax(m, b[u]);
    }
    for (int i = threadIdx.x + 4 * kBlock; i < cnt; i += kBlock) {
        s += load_partial(ps, i);
        m = fmax(m, load_partial(pm, i));
    }
    block_sum_max(s, m, sh);
    S = s;
    M = m;
}

}  // namespace maze

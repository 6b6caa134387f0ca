// noalloc.hip -- the FP64 implicit-Jacobi CG solve (library/src/plcg.c:28-90) on
// caller-owned device buffers, with every work buffer in one statically
// allocated device array: no device or host allocation at call time.
//
// For callers that cannot allocate device memory (code between graph captures,
// tight serving loops, harnesses that forbid hipMalloc).  The library keeps one
// copy of `scratch` per GPU for the life of the process; each call on a GPU
// reuses it, so calls that overlap on one GPU (two host threads, or two streams
// without ordering between them) race and must be serialised by the caller.
#include <cstdio>
#include <mutex>

#include "sim.hpp"

namespace maze {
namespace {

// Largest grid: 64^3 points (three FP64 work grids of 2 MiB each).
constexpr int kNoAllocMaxN = 64;
constexpr long kNoAllocMaxG = (long)kNoAllocMaxN * kNoAllocMaxN * kNoAllocMaxN;
constexpr long kNoAllocMaxPartials = 4096;

// Byte layout of the scratch array (each part 256-byte aligned).
constexpr size_t align256(size_t b) { return (b + 255) & ~size_t(255); }
constexpr size_t kOffState = 0;
constexpr size_t kOffCounter = align256(kOffState + sizeof(CGState));
constexpr size_t kOffPsum = align256(kOffCounter + kArrivalWords * sizeof(unsigned int));
constexpr size_t kOffPmax = align256(kOffPsum + kNoAllocMaxPartials * sizeof(double));
constexpr size_t kOffR = align256(kOffPmax + kNoAllocMaxPartials * sizeof(double));
constexpr size_t kOffP0 = align256(kOffR + kNoAllocMaxG * sizeof(double));
constexpr size_t kOffP1 = align256(kOffP0 + kNoAllocMaxG * sizeof(double));
constexpr size_t kScratchBytes = align256(kOffP1 + kNoAllocMaxG * sizeof(double));

static __device__ __align__(256) unsigned char scratch[kScratchBytes];

// The device address of `scratch` on each GPU, looked up once per device with
// that device current.  In host code the array's own name is not a device
// address; only these pointers are passed to kernels.
constexpr int kMaxDevices = 64;
std::mutex g_scratch_mu;
unsigned char* g_scratch_ptr[kMaxDevices] = {};

unsigned char* scratch_on_current_device() {
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kMaxDevices) return nullptr;
    std::lock_guard<std::mutex> g(g_scratch_mu);
    if (!g_scratch_ptr[dev]) {
        void* ptr = nullptr;
        const hipError_t e = hipGetSymbolAddress(&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr) {
            std::fprintf(stderr, "maze: hipGetSymbolAddress(scratch) failed on device %d: %s\n",
                         dev, hipGetErrorName(e));
            return nullptr;
        }
        g_scratch_ptr[dev] = static_cast<unsigned char*>(ptr);
    }
    return g_scratch_ptr[dev];
}

__global__ void k_clear_counters(unsigned int* c) {
    for (int i = threadIdx.x; i < kArrivalWords; i += blockDim.x) c[i] = 0u;
}

}  // namespace
}  // namespace maze

using namespace maze;

extern "C" int maze_conj_grad_noalloc(const double* b, double* x, double tol, int n,
                                      void* stream) {
    if (n < 2 || n > kNoAllocMaxN) {
        std::fprintf(stderr,
                     "maze_conj_grad_noalloc: n = %d is outside 2..%d (the statically allocated "
                     "work space holds grids of at most %d^3 points)\n",
                     n, kNoAllocMaxN, kNoAllocMaxN);
        return -2;
    }
    if (!b || !x) {
        std::fprintf(stderr, "maze_conj_grad_noalloc: b and x must be device pointers\n");
        return -2;
    }
    const GridDims d = grid_dims_for(n);
    if (d.blocks > kNoAllocMaxPartials) {
        std::fprintf(stderr, "maze_conj_grad_noalloc: %d work-groups exceed the %ld partials\n",
                     d.blocks, kNoAllocMaxPartials);
        return -2;
    }
    unsigned char* base = scratch_on_current_device();
    if (!base) return -3;
    hipStream_t s = static_cast<hipStream_t>(stream);
    CGState* st = reinterpret_cast<CGState*>(base + kOffState);
    Reduce red{reinterpret_cast<double*>(base + kOffPsum), reinterpret_cast<double*>(base + kOffPmax),
               reinterpret_cast<unsigned int*>(base + kOffCounter)};
    double* r = reinterpret_cast<double*>(base + kOffR);
    double* pa = reinterpret_cast<double*>(base + kOffP0);
    double* pb = reinterpret_cast<double*>(base + kOffP1);
    // The arrival counters must start at zero (a previous call leaves them at
    // zero unless it was interrupted); the work grids are written before read.
    MAZE_LAUNCH(k_clear_counters, dim3(1), dim3(64), 0, s, red.counter);
    const int limit = n * n;
    launch_cg_reset(s, st, tol, 0.0, limit, 0);
    launch_cg_init<double>(s, d, PBCoef{}, x, b, r, st, red);
    // Iterations in chunks (the first as long as a typical solve), the device's
    // stop flag read between chunks; launches after the stop are no-ops.
    int launched = 0, chunk = 33;
    CGState h{};
    for (;;) {
        for (int c = 0; c < chunk && launched < limit; c++, launched++) {
            launch_cg_dir<double>(s, d, PBCoef{}, pa, r, pb, st, red, launched == 0);
            launch_cg_update<double>(s, d, PBCoef{}, pb, x, r, st, red);
            std::swap(pa, pb);
        }
        MAZE_HIP(hipMemcpyAsync(&h, st, sizeof(CGState), hipMemcpyDeviceToHost, s));
        MAZE_HIP(hipStreamSynchronize(s));
        if (h.done || launched >= limit) break;
        chunk = std::max(4, chunk / 4);
    }
    return h.res;
}

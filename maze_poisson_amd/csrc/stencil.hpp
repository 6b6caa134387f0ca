// stencil.hpp -- the i-marching 7-point stencil sweep shared by the grid kernels.
//
// A work-group owns 256 consecutive points of an i-plane (coalesced along the
// contiguous plane) and walks `slab` planes, keeping the i-1 / i / i+1 values
// of its column in registers.  The value being differenced is produced on the
// fly by a source functor, so fused passes never materialise intermediates.
#pragma once

#include "grid.hpp"

namespace maze {

// ---- source functors -------------------------------------------------------
template <class Real>
struct SrcArr {
    const Real* a;
    __device__ __forceinline__ double operator()(int i) const { return (double)a[i]; }
};

struct SrcPred {  // phi* = 2 phi - phi_prev   (library/src/verlet.c:223-225)
    const double* phi;
    const double* phip;
    __device__ __forceinline__ double operator()(int i) const { return 2 * phi[i] - phip[i]; }
};

template <class Real>
struct SrcDir {  // p = beta p + r/6, or r/6 for the first direction (plcg.c:52-53, :79-80)
    const Real* p;
    const Real* r;
    double beta;
    int first;
    __device__ __forceinline__ double operator()(int i) const {
        const double c6 = 1.0 / 6.0;
        double v = first ? (double)r[i] * c6 : (double)p[i] * beta + c6 * (double)r[i];
        return (double)(Real)v;
    }
};

struct SrcSum {  // y + d
    const double* y;
    const float* d;
    int add;
    __device__ __forceinline__ double operator()(int i) const {
        return add ? y[i] + (double)d[i] : y[i];
    }
};

struct SrcZero {
    __device__ __forceinline__ double operator()(int) const { return 0.0; }
};

// ---- the i-marching stencil sweep -------------------------------------------
// Calls body(c, value_at_c, (A value)_c) for every point of this work-group's
// tile.  Threads past the end of the plane do nothing but stay for the
// work-group reductions.
template <bool PB, class Src, class Body>
__device__ __forceinline__ void march(const GridDims& d, const PBCoef& pb, const Src& src,
                                      Body& body) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int chunk = b % d.chunks;
    const int s = b / d.chunks;
    const int pl = chunk * kBlock + threadIdx.x;
    if (pl >= d.n2) return;
    const int n = d.n, n2 = d.n2;
    const PlaneNb nb = plane_nb(pl, n, n2);
    const int i0 = s * d.slab;
    const int i1 = min(n, i0 + d.slab);
    const int im0 = (i0 == 0) ? n - 1 : i0 - 1;
    double vm = src(im0 * n2 + pl);
    double vc = src(i0 * n2 + pl);
    double exm = 0.0;
    if (PB) exm = pb.ex[im0 * n2 + pl];
    for (int i = i0; i < i1; i++) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const int base = i * n2;
        const int c = base + pl;
        const double vp = src(ip * n2 + pl);
        const double vjm = src(base + nb.jm), vjp = src(base + nb.jp);
        const double vkm = src(base + nb.km), vkp = src(base + nb.kp);
        double lap;
        if (!PB) {
            lap = vp + vm + vjp + vjm + vkp + vkm - vc * 6.0;
        } else {
            const double exc = pb.ex[c], eyc = pb.ey[c], ezc = pb.ez[c];
            const double eym = pb.ey[base + nb.jm], ezm = pb.ez[base + nb.km];
            lap = vp * exc + vm * exm + vjp * eyc + vjm * eym + vkp * ezc + vkm * ezm -
                  vc * (exc + exm + eyc + eym + ezc + ezm + pb.k2[c]);
            exm = exc;
        }
        body(c, vc, lap);
        vm = vc;
        vc = vp;
    }
}

template <bool PB, class Src, class Body>
__global__ __launch_bounds__(kBlock) void k_pass(GridDims d, PBCoef pb, Src src, Body body,
                                                 Reduce red, CGState* st) {
    if (Body::kSkipWhenDone && st->done) return;
    Body bd = body;
    march<PB>(d, pb, src, bd);
    if constexpr (Body::kReduce) {
        __shared__ double sh[kBlock / kWave];
        const double s = block_sum(bd.s, sh);
        const double m = block_max(bd.m, sh);
        if (publish_arrive(red.psum, s, red.pmax, m, red.counter, gridDim.x)) {
            const double S = reduce_partials_sum(red.psum, gridDim.x, sh);
            const double M = reduce_partials_max(red.pmax, gridDim.x, sh);
            if (threadIdx.x == 0) {
                Body::finalize(st, S, M);
                reset_counter(red.counter);
            }
        }
    }
}

template <bool PB, class Src, class Body>
static void run_pass(hipStream_t s, const GridDims& d, const PBCoef& pb, const Src& src,
                     const Body& body, Reduce red, CGState* st) {
    hipLaunchKernelGGL((k_pass<PB, Src, Body>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb, src,
                       body, red, st);
}

#define DISPATCH_PB(pb, ...)                \
    do {                                    \
        if ((pb).ex) {                      \
            constexpr bool kPB = true;      \
            __VA_ARGS__;                    \
        } else {                            \
            constexpr bool kPB = false;     \
            __VA_ARGS__;                    \
        }                                   \
    } while (0)

}  // namespace maze

// stencil.hpp -- the i-marching 7-point stencil sweep shared by the grid kernels.
//
// A work-group owns 256 consecutive points of an i-plane (coalesced along the
// contiguous plane) and walks `slab` planes, keeping the i-1 / i / i+1 values
// of its column in registers.  The value being differenced is produced on the
// fly by a source functor, so fused passes never materialise intermediates.
//
// Source functors and bodies separate loading from arithmetic (`load` returns
// the raw operands, `val` / `apply` combine them): every load of a point --
// the five new neighbour values and the body's centre operands -- is issued
// before any of them is used, so a thread keeps several requests in flight.
// Branches that depend on launch-time constants are template parameters for
// the same reason (a data-independent branch between two loads still splits
// the loop into blocks that each wait for their loads).
#pragma once

#include "grid.hpp"

namespace maze {

struct NoRaw {};

// ---- source functors -------------------------------------------------------
template <class Real>
struct SrcArr {
    const Real* a;
    using Raw = Real;
    __device__ __forceinline__ Raw load(int i) const { return a[i]; }
    __device__ __forceinline__ double val(const Raw& r) const { return (double)r; }
};

struct SrcPred {  // phi* = 2 phi - phi_prev   (library/src/verlet.c:223-225)
    const double* phi;
    const double* phip;
    struct Raw { double a, b; };
    __device__ __forceinline__ Raw load(int i) const { return Raw{phi[i], phip[i]}; }
    __device__ __forceinline__ double val(const Raw& r) const { return 2 * r.a - r.b; }
};

// p = beta p + r/6, or r/6 for the first direction (library/src/plcg.c:52-53, :79-80),
// rounded to the storage type so that every thread sees the stored value.
template <class Real, bool First>
struct SrcDir {
    const Real* p;
    const Real* r;
    double beta;
    struct Raw { Real p, r; };
    __device__ __forceinline__ Raw load(int i) const {
        if (First) return Raw{(Real)0, r[i]};
        return Raw{p[i], r[i]};
    }
    __device__ __forceinline__ double val(const Raw& x) const {
        const double c6 = 1.0 / 6.0;
        const double v = First ? (double)x.r * c6 : (double)x.p * beta + c6 * (double)x.r;
        return (double)(Real)v;
    }
};

template <bool Add>
struct SrcSum {  // y + d
    const double* y;
    const float* d;
    struct Raw { double y; float d; };
    __device__ __forceinline__ Raw load(int i) const {
        if (Add) return Raw{y[i], d[i]};
        return Raw{y[i], 0.0f};
    }
    __device__ __forceinline__ double val(const Raw& r) const {
        return Add ? r.y + (double)r.d : r.y;
    }
};

struct SrcZero {
    using Raw = NoRaw;
    __device__ __forceinline__ Raw load(int) const { return Raw{}; }
    __device__ __forceinline__ double val(const Raw&) const { return 0.0; }
};

// ---- the i-marching stencil sweep -------------------------------------------
// Calls body.apply(c, body.load(c), value_at_c, (A value)_c) for every point of
// this work-group's tile.  Threads past the end of the plane do nothing but
// stay for the work-group reductions.
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
    double vm = src.val(src.load(im0 * n2 + pl));
    double vc = src.val(src.load(i0 * n2 + pl));
    double exm = 0.0;
    if (PB) exm = pb.ex[im0 * n2 + pl];
#pragma unroll 2
    for (int i = i0; i < i1; i++) {
        const int ip = (i + 1 == n) ? 0 : i + 1;
        const int base = i * n2;
        const int c = base + pl;
        const typename Src::Raw rp = src.load(ip * n2 + pl);
        const typename Src::Raw rjm = src.load(base + nb.jm);
        const typename Src::Raw rjp = src.load(base + nb.jp);
        const typename Src::Raw rkm = src.load(base + nb.km);
        const typename Src::Raw rkp = src.load(base + nb.kp);
        const typename Body::Raw bc = body.load(c);
        double exc = 0.0, eyc = 0.0, ezc = 0.0, eym = 0.0, ezm = 0.0, k2c = 0.0;
        if (PB) {
            exc = pb.ex[c];
            eyc = pb.ey[c];
            ezc = pb.ez[c];
            eym = pb.ey[base + nb.jm];
            ezm = pb.ez[base + nb.km];
            k2c = pb.k2[c];
        }
        const double vp = src.val(rp), vjm = src.val(rjm), vjp = src.val(rjp);
        const double vkm = src.val(rkm), vkp = src.val(rkp);
        double lap;
        if (!PB) {
            lap = vp + vm + vjp + vjm + vkp + vkm - vc * 6.0;
        } else {
            lap = vp * exc + vm * exm + vjp * eyc + vjm * eym + vkp * ezc + vkm * ezm -
                  vc * (exc + exm + eyc + eym + ezc + ezm + k2c);
            exm = exc;
        }
        body.apply(c, bc, vc, lap);
        vm = vc;
        vc = vp;
    }
}

// One sweep with an optional single-pass reduction of (sum, max) accumulated
// by the body; the last work-group calls Body::finalize with the totals.
template <bool PB, class Src, class Body>
__global__ __launch_bounds__(kBlock) void k_pass(GridDims d, PBCoef pb, Src src, Body body,
                                                 Reduce red, CGState* st) {
    if (Body::kSkipWhenDone && st->done) return;
    Body bd = body;
    march<PB>(d, pb, src, bd);
    if constexpr (Body::kReduce) {
        __shared__ double sh[2 * kBlock / kWave];
        double s = bd.s, m = 0.0;
        if constexpr (Body::kMax) {
            m = bd.m;
            block_sum_max(s, m, sh);
        } else {
            s = block_sum(s, sh);
        }
        if (publish_arrive(red.psum, s, Body::kMax ? red.pmax : nullptr, m, red.counter,
                           gridDim.x)) {
            double S, M = 0.0;
            if constexpr (Body::kMax)
                reduce_partials_sum_max(red.psum, red.pmax, gridDim.x, sh, S, M);
            else
                S = reduce_partials_sum(red.psum, gridDim.x, sh);
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

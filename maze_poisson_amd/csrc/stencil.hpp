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

// p = beta p + cr r, or cr r for the first direction, with cr = 1/6 for the
// implicit-Jacobi recurrence (library/src/plcg.c:52-53, :79-80) or -1 for plain
// CG (library/src/multigrid_pb.c:43-44, :89-90); rounded to the storage type so
// that every thread sees the stored value.
template <class Real, bool First>
struct SrcDir {
    const Real* p;
    const Real* r;
    double beta;
    double cr;
    struct Raw { Real p, r; };
    __device__ __forceinline__ Raw load(int i) const {
        if (First) return Raw{(Real)0, r[i]};
        return Raw{p[i], r[i]};
    }
    __device__ __forceinline__ double val(const Raw& x) const {
        const double v = First ? (double)x.r * cr : (double)x.p * beta + cr * (double)x.r;
        return (double)(Real)v;
    }
};

template <bool Add>
struct SrcSum {  // y + dscale d  (dscale 1, or 0 to ignore d)
    const double* y;
    const float* d;
    double dscale;
    struct Raw { double y; float d; };
    __device__ __forceinline__ Raw load(int i) const {
        if (Add) return Raw{y[i], d[i]};
        return Raw{y[i], 0.0f};
    }
    __device__ __forceinline__ double val(const Raw& r) const {
        return Add ? r.y + dscale * (double)r.d : r.y;
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
// stay for the work-group reductions.  Returns the hook's verdict (true
// without a hook).
//
// Software-pipelined by two planes: the operands of planes i+1 and i+2 are
// requested before plane i is computed, so each thread has up to three planes
// of loads in flight and the march is not one memory latency per plane.  (The
// planes fetched after the last one are in bounds -- indices wrap -- and
// unused.)  Every pass writes arrays other than the ones its source reads,
// so requesting the next plane before this plane's stores is safe.
struct PBPlane {
    double exc, eyc, ezc, eym, ezm, k2c;
};

template <bool PB, class C>
__device__ __forceinline__ PBPlane pb_load(const C& pb, int base, int pl, const PlaneNb& nb) {
    PBPlane r{0.0, 0.0, 0.0, 0.0, 0.0, 0.0};
    if (PB) {
        const int c = base + pl;
        r.exc = pb.ex[c];
        r.eyc = pb.ey[c];
        r.ezc = pb.ez[c];
        r.eym = pb.ey[base + nb.jm];
        r.ezm = pb.ez[base + nb.km];
        r.k2c = pb.k2[c];
    }
    return r;
}

struct NoHook {
    __device__ bool operator()() const { return true; }
};

// hook: called once by every thread of the work-group (also those past the
// end of the plane), at one point of uniform control flow, once the first
// plane's loads are in flight and before any value is computed: a kernel that
// derives a coefficient of its source or body from a reduction (deferred
// finalisation) sets it there, so the fold's latency overlaps the loads.  (A
// hook holds work-group barriers, so it must not sit in divergent code.)  It
// returns whether the sweep goes on; false (the same for the whole
// work-group) ends it before any value is computed or stored.
template <bool PB, class Src, class Body, class C, class Hook = NoHook>
__device__ __forceinline__ bool march(const GridDims& d, const C& pb, const Src& src, Body& body,
                                      Hook&& hook = Hook{}) {
    using SR = typename Src::Raw;
    using BR = typename Body::Raw;
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int chunk = b % d.chunks;
    const int s = b / d.chunks;
    const int pl0 = chunk * kBlock + threadIdx.x;
    const bool active = pl0 < d.n2;
    const int pl = active ? pl0 : 0;  // threads past the end compute nothing
    const int n = d.n, n2 = d.n2;
    const PlaneNb nb = plane_nb(pl, n, n2);
    const int i0 = s * d.slab;
    const int i1 = min(n, i0 + d.slab);
    const int im0 = (i0 == 0) ? n - 1 : i0 - 1;
    auto wrap = [n](int i) { return i >= n ? i - n : i; };
    // The operands of one plane: the i+1 neighbour, the four in-plane
    // neighbours, the body's centre operands and the PB coefficients.
    struct Ops {
        SR p, jm, jp, km, kp;
        BR c;
        PBPlane cf;
    };
    auto fetch = [&](int i) {
        const int base = wrap(i) * n2;
        Ops o;
        o.p = src.load(wrap(i + 1) * n2 + pl);
        o.jm = src.load(base + nb.jm);
        o.jp = src.load(base + nb.jp);
        o.km = src.load(base + nb.km);
        o.kp = src.load(base + nb.kp);
        o.c = body.load(base + pl);
        o.cf = pb_load<PB>(pb, base, pl, nb);
        return o;
    };
    // Threads past the end of the plane load point 0 of their planes (in
    // bounds, unused): unconditional loads leave the compiler no reason to
    // convert their values before the hook, which would wait for them there.
    const SR rm = src.load(im0 * n2 + pl);
    const SR rc = src.load(i0 * n2 + pl);
    double exm = PB ? pb.ex[im0 * n2 + pl] : 0.0;
    Ops A = fetch(i0);
    // A one-plane slab (small grids) reads its own plane again (cache hits)
    // instead of the next one; the select keeps the loads branch-free.
    Ops B = fetch(i0 + 1 < i1 ? i0 + 1 : i0);
    const bool go = hook();
    if (!active || !go) return go;
    double vm = src.val(rm);
    double vc = src.val(rc);
    auto compute = [&](int i, const Ops& o) {
        const double vp = src.val(o.p), vjm = src.val(o.jm), vjp = src.val(o.jp);
        const double vkm = src.val(o.km), vkp = src.val(o.kp);
        double lap;
        if (!PB) {
            lap = vp + vm + vjp + vjm + vkp + vkm - vc * 6.0;
        } else {
            const PBPlane& f = o.cf;
            lap = vp * f.exc + vm * exm + vjp * f.eyc + vjm * f.eym + vkp * f.ezc + vkm * f.ezm -
                  vc * (f.exc + exm + f.eyc + f.eym + f.ezc + f.ezm + f.k2c);
            exm = f.exc;
        }
        body.apply(i * n2 + pl, o.c, vc, lap);
        vm = vc;
        vc = vp;
    };
    // Three operand sets rotate (no register copies of pending loads), so two
    // planes of loads are in flight while a plane is computed.  The loop has
    // no exit between its thirds: with one there, the compiler's structured
    // control flow adds an edge on which later sets' loads are still pending
    // at the loop head, and the head then waits for the newest loads before
    // it issues the next plane's.  The scheduling barriers keep each plane's
    // loads ahead of the other planes' arithmetic.
    int i = i0;
    for (; i + 2 < i1; i += 3) {
        const Ops P = fetch(i + 2);
        __builtin_amdgcn_sched_barrier(0);
        compute(i, A);
        A = fetch(i + 3);
        __builtin_amdgcn_sched_barrier(0);
        compute(i + 1, B);
        B = fetch(i + 4);
        __builtin_amdgcn_sched_barrier(0);
        compute(i + 2, P);
    }
    if (i < i1) compute(i, A);
    if (i + 1 < i1) compute(i + 1, B);
    return true;
}

// Single-pass reduction of the (sum, max) a body accumulated: the last
// work-group calls bd.finalize(st, S, M) with the totals.
// NT: threads per work-group (the (sum, max) form needs NT == kBlock).
template <bool kMax, class Body, int NT = kBlock>
__device__ __forceinline__ void reduce_and_finalize(Body& bd, Reduce red, CGState* st) {
    static_assert(!kMax || NT == kBlock, "sum+max reduction assumes kBlock threads");
    __shared__ double sh[2 * NT / kWave];
    double s = bd.s, m = 0.0;
    if constexpr (kMax) {
        m = bd.m;
        block_sum_max(s, m, sh);
    } else {
        s = block_sum<NT>(s, sh);
    }
    if (publish_arrive(red.psum, s, kMax ? red.pmax : nullptr, m, red.counter, gridDim.x)) {
        double S, M = 0.0;
        if constexpr (kMax)
            reduce_partials_sum_max(red.psum, red.pmax, gridDim.x, sh, S, M);
        else
            S = reduce_partials_sum<NT>(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            bd.finalize(st, S, M);
            reset_arrivals(red.counter);
        }
    }
}

// x += alpha p; r += alpha A p; <r, r>, max|r|; stop test and beta
// (library/src/plcg.c:64-80; relative stop of library/src/multigrid.c:74-78).
// ZeroX: x is taken as 0 (first update of a solve started from zero), so it
// is written without being read.
template <class Real, bool ZeroX = false>
struct BodyUpdate {
    static constexpr bool kReduce = true, kMax = true, kSkipWhenDone = true;
    Real* x;
    Real* r;
    double alpha;
    double s = 0.0, m = 0.0;
    struct Raw { Real x, r; };
    __device__ Raw load(int c) const { return Raw{ZeroX ? (Real)0 : x[c], r[c]}; }
    __device__ void apply(int c, const Raw& o, double v, double lap) {
        x[c] = (Real)((double)o.x + alpha * v);
        const Real rn = (Real)((double)o.r + alpha * lap);
        r[c] = rn;
        const double rd = (double)rn;
        s += rd * rd;
        m = fmax(m, fabs(rd));
    }
    __device__ static void finalize(CGState* st, double S, double M) {
        st->rr = S;
        st->rmax = M;
        if (st->count_mode) {
            if (M <= st->rtol * st->r0) {
                st->done = 1;
                st->res = st->iter + 1;
                return;
            }
        } else if (M <= st->tol) {
            st->done = 1;
            st->res = st->iter;
            return;
        }
        if (!st->pcg) {
            // <r_new, v_new> with v_new = -r_new/6, or <r_new, r_new> for plain CG
            // (library/src/plcg.c:74-78, library/src/multigrid_pb.c:83-86).
            const double rn_dot_vn = st->plain ? S : -S / 6.0;
            st->beta = rn_dot_vn / st->rdv;
            st->rdv = rn_dot_vn;
        }
        st->iter += 1;
        if (st->iter >= st->limit) {
            st->done = 1;
            st->res = st->count_mode ? st->limit : -1;
        }
    }
};

// ---- deferred finalisation ----------------------------------------------------
// The producer writes one partial per work-group with a plain store and ends;
// every work-group of the consumer folds all partials itself, in the same
// fixed order, so all of them derive the same scalars.  The kernel boundary
// between producer and consumer makes the partials visible; no work-group of
// the producer waits for a store or an arrival counter.

// Work-group sum and max whose barrier orders LDS only (global stores issued
// before it stay in flight).  sh holds 2 NT / 64 doubles; results in all threads.
template <int NT>
__device__ __forceinline__ void block_sum_max_lds(double& s, double& m, double* sh) {
    s = wave_sum(s);
    m = wave_max(m);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sh[w] = s;
        sh[NT / kWave + w] = m;
    }
    lds_barrier();
    double a = 0.0, b = 0.0;
    for (int i = 0; i < NT / kWave; i++) {
        a += sh[i];
        b = fmax(b, sh[NT / kWave + i]);
    }
    s = a;
    m = b;
}

// Sum and max of partials[0..cnt) in a fixed order; valid in all threads.
// Called once per kernel, before any other use of sh.
template <int NT>
__device__ __forceinline__ void fold_sum_max(const double* __restrict__ ps,
                                             const double* __restrict__ pm, int cnt, double* sh,
                                             double& S, double& M) {
    double a = 0.0, b = 0.0;
    for (int i = threadIdx.x; i < cnt; i += NT) {
        a += ps[i];
        b = fmax(b, pm[i]);
    }
    block_sum_max_lds<NT>(a, b, sh);
    lds_barrier();
    S = a;
    M = b;
}

// The folds a deferred-finalisation consumer starts from, together: the sum
// over z[0..zcnt), and the sum over s and the maximum over m (both [0..ucnt);
// an array left out, HasZ or HasU false, leaves its result 0; a folded count
// is at least 1), bit for bit as fold_sum_max forms them.
// A thread reads its slots (t, t + NT, ...) in batches of kFoldBatch slots of
// every array and issues a batch's loads before it adds the first, so one
// memory round trip covers the batch (a loop that adds each slot as it
// arrives costs one round trip per slot and array).  flag: thread 0's value,
// handed to the whole work-group by the same barrier.  sh: its own
// 3 NT / 64 + 1 doubles (not reused, so no closing barrier); results in all
// threads.
constexpr int kFoldBatch = 4;

template <int NT, bool HasZ, bool HasU>
__device__ __forceinline__ void fold_partials(const double* __restrict__ pz, int zcnt,
                                              const double* __restrict__ ps,
                                              const double* __restrict__ pm, int ucnt,
                                              double* sh, double& Z, double& S, double& M,
                                              int& flag) {
    constexpr int NW = NT / kWave;
    double az = 0.0, as = 0.0, am = 0.0;
    if (!HasZ) zcnt = 0;
    if (!HasU) ucnt = 0;
    const int cnt = max(zcnt, ucnt);
    for (int i0 = threadIdx.x; i0 < cnt; i0 += kFoldBatch * NT) {
        double vz[kFoldBatch], vs[kFoldBatch], vm[kFoldBatch];
#pragma unroll
        for (int u = 0; u < kFoldBatch; u++) {
            const int i = i0 + u * NT;
            if (HasZ) vz[u] = pz[min(i, zcnt - 1)];
            if (HasU) {
                vs[u] = ps[min(i, ucnt - 1)];
                vm[u] = pm[min(i, ucnt - 1)];
            }
        }
#pragma unroll
        for (int u = 0; u < kFoldBatch; u++) {
            const int i = i0 + u * NT;
            if (HasZ && i < zcnt) az += vz[u];
            if (HasU && i < ucnt) {
                as += vs[u];
                am = fmax(am, vm[u]);
            }
        }
    }
    az = wave_sum(az);
    as = wave_sum(as);
    am = wave_max(am);
    const int w = threadIdx.x / kWave;
    if ((threadIdx.x & (kWave - 1)) == 0) {
        sh[w] = az;
        sh[NW + w] = as;
        sh[2 * NW + w] = am;
    }
    if (threadIdx.x == 0) sh[3 * NW] = (double)flag;
    lds_barrier();
    double a = 0.0, b = 0.0, c = 0.0;
    for (int i = 0; i < NW; i++) {
        a += sh[i];
        b += sh[NW + i];
        c = fmax(c, sh[2 * NW + i]);
    }
    flag = (int)sh[3 * NW];
    Z = a;
    S = b;
    M = c;
}

// Whether the solve has ended, read once per work-group (one barrier).  A
// kernel in which work-group 0 may record a stop while other work-groups are
// starting needs this form: were every wave to read st->done itself, the
// waves of one work-group could see different values, and those that went on
// would fold or use LDS contents that the waves that left were to write.
__device__ __forceinline__ bool wg_done(const CGState* st) {
    __shared__ int flag;
    if (threadIdx.x == 0) flag = st->done;
    lds_barrier();
    return flag != 0;
}

// The stop test of a multigrid-PCG update, from its partials (pus, pum): the
// same decision as BodyUpdate::finalize, with st->iter already advanced by the
// update.  Returns the result code if the test stops the solve (converged or
// out of budget), else INT_MIN; work-group 0 records a stop in st.  The test
// reads only fields that no consumer writes, so every work-group of a kernel
// decides alike whenever work-group 0 writes; a stop recorded by an earlier
// kernel (st->done, e.g. a breakdown) is the caller's to check.
// sh: 2 NT / 64 doubles.
// record: work-group 0 writes a stop into st.  A kernel whose own last
// work-group rewrites the solver state (the refinement pass) must not record:
// two work-groups writing the same fields in one kernel leave either value.
__device__ __forceinline__ int df_decide(CGState* st, double S, double M, bool record = true) {
    const int k = st->iter;
    int res = INT_MIN;
    if (st->count_mode) {
        if (M <= st->rtol * st->r0) res = k;
        else if (k >= st->limit) res = st->limit;
    } else {
        if (M <= st->tol) res = k - 1;
        else if (k >= st->limit) res = -1;
    }
    if (record && res != INT_MIN && blockIdx.x == 0 && threadIdx.x == 0) {
        st->rr = S;
        st->rmax = M;
        st->res = res;
        st->done = 1;
    }
    return res;
}

template <int NT>
__device__ __forceinline__ int df_stop_test(CGState* st, const double* pus, const double* pum,
                                            int cnt, double* sh, bool record = true) {
    double S, M;
    fold_sum_max<NT>(pus, pum, cnt, sh, S, M);
    return df_decide(st, S, M, record);
}

// One sweep with an optional single-pass reduction of (sum, max) accumulated
// by the body; the last work-group calls Body::finalize with the totals.
template <bool PB, class Src, class Body, class C>
__global__ __launch_bounds__(kBlock) void k_pass(GridDims d, C pb, Src src, Body body, Reduce red,
                                                 CGState* st) {
    if (Body::kSkipWhenDone && st->done) return;
    Body bd = body;
    march<PB>(d, pb, src, bd);
    if constexpr (Body::kReduce) reduce_and_finalize<Body::kMax>(bd, red, st);
}

template <bool PB, class Src, class Body, class C>
static void run_pass(hipStream_t s, const GridDims& d, const C& pb, const Src& src,
                     const Body& body, Reduce red, CGState* st) {
    MAZE_LAUNCH((k_pass<PB, Src, Body, C>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb, src, body,
                red, st);
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

// grid_kernels.hip -- 7-point Poisson / linear Poisson-Boltzmann kernels for gfx950.
//
// Every grid pass is one i-marching stencil sweep (stencil.hpp) with a fused
// body.  Reductions are single-pass: per-work-group partials plus a
// last-arriving work-group that folds them in index order (deterministic) and
// advances the solver scalars in device memory.
//
// Floating-point expressions follow the reference's operation order
// (library/src/laplace.c:301-309, :366-379; library/src/plcg.c:41-83;
// library/src/verlet.c:221-240), so every point-wise result matches the CPU
// oracle bit for bit; only the order of the global sums differs.
#include <cmath>

#include "stencil.hpp"

namespace maze {

// ---- bodies -----------------------------------------------------------------
// A body declares the centre operands it needs (Raw, load) and combines them
// with the source value and its stencil in apply; kReduce/kMax select the
// (sum, max) reduction handed to finalize.
struct BodyLaplace {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = false;
    double* out;
    double s = 0.0, m = 0.0;
    using Raw = NoRaw;
    __device__ Raw load(int) const { return Raw{}; }
    __device__ void apply(int c, const Raw&, double, double lap) { out[c] = lap; }
    __device__ static void finalize(CGState*, double, double) {}
};

// phi* and sigma = A phi* + coef q (library/src/verlet.c:221-230).
struct BodyPred {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = false;
    double* phistar;
    double* sigma;
    const double* q;
    double coef;
    double s = 0.0, m = 0.0;
    using Raw = double;
    __device__ Raw load(int c) const { return q[c]; }
    __device__ void apply(int c, const Raw& qc, double v, double lap) {
        phistar[c] = v;
        sigma[c] = lap + coef * qc;
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// r = A x - b; <r, r/6>; max|r|   (library/src/plcg.c:41-58)
// cr: 1/6 for the implicit-Jacobi recurrence, 1 for plain CG (st->plain):
// <r, v> = -<r, r/6> with v = -r/6, or <r, r> (library/src/multigrid_pb.c:55).
template <class Real>
struct BodyInit {
    // Skipped when the reset before it left the solve stopped (a gated solve
    // after the outer loop has converged).
    static constexpr bool kReduce = true, kMax = true, kSkipWhenDone = true;
    const Real* b;
    Real* r;
    Real* x_zero;  // when set, x is taken as 0 and written back as 0
    double cr;
    double s = 0.0, m = 0.0;
    using Raw = Real;
    __device__ Raw load(int c) const { return b[c]; }
    __device__ void apply(int c, const Raw& bc, double, double lap) {
        const Real rc = (Real)(lap + -1.0 * (double)bc);
        r[c] = rc;
        if (x_zero) x_zero[c] = (Real)0;
        const double rd = (double)rc;
        s += rd * (rd * cr);
        m = fmax(m, fabs(rd));
    }
    __device__ static void finalize(CGState* st, double S, double M) {
        st->rdv = st->plain ? S : -S;
        st->r0 = M;
        st->iter = 0;
        st->first = 1;
        st->beta = 0.0;
        st->alpha = 0.0;
        // An identically zero residual has no direction to search; the
        // reference would divide 0 by 0 here.  Stop with the input unchanged.
        if (M == 0.0) {
            st->done = 1;
            st->res = 0;
        } else {
            st->done = 0;
            st->res = -1;
        }
    }
};

// p_new = beta p + r/6 (written), <p_new, A p_new>   (library/src/plcg.c:61-63)
template <class Real>
struct BodyDir {
    static constexpr bool kReduce = true, kMax = false, kSkipWhenDone = true;
    Real* p_new;
    double s = 0.0, m = 0.0;
    using Raw = NoRaw;
    __device__ Raw load(int) const { return Raw{}; }
    __device__ void apply(int c, const Raw&, double v, double lap) {
        p_new[c] = (Real)v;
        s += v * lap;
    }
    __device__ static void finalize(CGState* st, double S, double) {
        st->pAp = S;
        st->first = 0;
        if (fabs(S) < 1e-300) {  // breakdown guard (library/src/multigrid.c:59-63)
            st->done = 1;
            st->res = st->iter;
            st->alpha = 0.0;
        } else {
            st->alpha = st->rdv / S;
        }
    }
};

// y_out = the new iterate (y + d, or the extrapolated start 2 y - y_prev);
// the FP64 residual b = sigma - A y_out; max|b|.  Advances the outer
// refinement loop and starts the next inner solve A x = b from x = 0:
// r32 = A 0 - b = -b rounded to FP32 and <r, r/6>, as the init pass of
// library/src/plcg.c:41-58 would compute them.
// Write: y_out is written; Add: a refinement pass (else the loop's start).
template <bool Write, bool Add>
struct BodyRefine {
    double* y_out;
    const double* sigma;
    float* r32;
    double otol, rtol;
    int limit, pcg;
    // Outcome of the inner solve of this pass (Add): stopped, and its result.
    int in_done = 0, in_res = 0;
    double s = 0.0, m = 0.0;
    using Raw = double;
    __device__ Raw load(int c) const { return sigma[c]; }
    __device__ void apply(int c, const Raw& sg, double v, double lap) {
        if (Write) y_out[c] = v;
        const double rr = sg - lap;
        const float rf = (float)(-rr);
        r32[c] = rf;
        const double rd = (double)rf;
        s += rd * (rd * (1.0 / 6.0));
        m = fmax(m, fabs(rr));
    }
    __device__ void finalize(CGState* st, double S, double M) const {
        st->resid = M;
        st->otol = otol;
        if (Add) {
            // The inner solve of this pass: its iteration count if it met its
            // relative stop, else minus the iterations it was given.
            const int it = in_done ? max(in_res, 0) : -st->iter;
            const int k = st->passes;
            if (k < kMaxPasses) st->pass_iters[k] = it;
            st->inner_total += it >= 0 ? it : -it;
            st->passes = k + 1;
            st->ysel ^= 1;
        } else {
            st->ysel = 0;
            st->passes = 0;
            st->inner_total = 0;
        }
        st->outer_done = M <= otol ? 1 : 0;
        if (st->outer_done) {
            st->done = 1;
            return;
        }
        st->tol = 0.0;
        st->rtol = rtol;
        st->limit = limit;
        st->count_mode = 1;
        st->pcg = pcg;
        st->plain = 0;
        st->rdv = -S;
        st->r0 = M;
        st->iter = 0;
        st->first = 1;
        st->alpha = 0.0;
        st->beta = 0.0;
        st->done = 0;
        st->res = -1;
    }
};

// max|A y - sigma|  (library/src/verlet.c:300-302)
struct BodyTrueRes {
    static constexpr bool kReduce = true, kMax = true, kSkipWhenDone = false;
    const double* sigma;
    double s = 0.0, m = 0.0;
    using Raw = double;
    __device__ Raw load(int c) const { return sigma[c]; }
    __device__ void apply(int, const Raw& sg, double, double lap) {
        const double t = lap + -1.0 * sg;
        m = fmax(m, fabs(t));
    }
    __device__ static void finalize(CGState* st, double, double M) { st->resid = M; }
};

// <r, v> for the explicit-preconditioner CG (library/src/plcg.c:127, :146-148).
struct BodyDotRV {
    static constexpr bool kReduce = true, kMax = false, kSkipWhenDone = false;
    const double* r;
    const double* v;
    int first;
    double s = 0.0, m = 0.0;
    struct Raw { double r, v; };
    __device__ Raw load(int c) const { return Raw{r[c], v[c]}; }
    __device__ void apply(int, const Raw& o, double, double) { s += o.r * o.v; }
    __device__ void finalize(CGState* st, double S, double) const {
        if (!first) st->beta = S / st->rdv;
        st->rdv = S;
    }
};

// ---- launchers --------------------------------------------------------------
void launch_laplace(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* u,
                    double* out) {
    BodyLaplace body{out};
    DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcArr<double>{u}, body, Reduce{}, nullptr));
}

void launch_pred_sigma(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* phi,
                       const double* phip, const double* q, double coef, double* phistar,
                       double* sigma) {
    BodyPred body{phistar, sigma, q, coef};
    DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcPred{phi, phip}, body, Reduce{}, nullptr));
}

template <class Real>
void launch_cg_init(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* x,
                    const Real* b, Real* r, CGState* st, Reduce red, Real* zero_x, int plain) {
    const double cr = plain ? 1.0 : 1.0 / 6.0;
    if (x) {
        BodyInit<Real> body{b, r, nullptr, cr};
        DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcArr<Real>{x}, body, red, st));
    } else {
        // x == nullptr means "start from zero": zero_x (if given) is cleared here.
        BodyInit<Real> body{b, r, zero_x, cr};
        DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcZero{}, body, red, st));
    }
}

template <bool PB, class Real, bool First>
__global__ __launch_bounds__(kBlock) void k_cg_dir(GridDims d, PBCoef pb, const Real* p_old,
                                                   const Real* r, Real* p_new, Reduce red,
                                                   CGState* st) {
    if (st->done) return;
    const SrcDir<Real, First> src{p_old, r, First ? 0.0 : st->beta, st->plain ? -1.0 : 1.0 / 6.0};
    BodyDir<Real> bd{p_new};
    march<PB>(d, pb, src, bd);
    reduce_and_finalize<false>(bd, red, st);
}

template <bool PB, class Real, bool ZeroX, class C = PBCoef>
__global__ __launch_bounds__(kBlock) void k_cg_update(GridDims d, C pb, const Real* p, Real* x,
                                                      Real* r, Reduce red, CGState* st) {
    if (st->done) return;
    BodyUpdate<Real, ZeroX> bd{x, r, st->alpha};
    march<PB>(d, pb, SrcArr<Real>{p}, bd);
    reduce_and_finalize<true>(bd, red, st);
}

// `first` is known to the host (the first direction of a solve is launched
// first) and selects the kernel, so the direction formula has no branch.
template <class Real>
void launch_cg_dir(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* p_old,
                   const Real* r, Real* p_new, CGState* st, Reduce red, int first) {
    if (first) {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_cg_dir<kPB, Real, true>), dim3(d.blocks),
                                           dim3(kBlock), 0, s, d, pb, p_old, r, p_new, red, st));
    } else {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_cg_dir<kPB, Real, false>), dim3(d.blocks),
                                           dim3(kBlock), 0, s, d, pb, p_old, r, p_new, red, st));
    }
}

template <class Real>
void launch_cg_update(hipStream_t s, const GridDims& d, const PBCoef& pb, const Real* p, Real* x,
                      Real* r, CGState* st, Reduce red, int zero_x) {
    if (zero_x) {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_cg_update<kPB, Real, true>), dim3(d.blocks),
                                           dim3(kBlock), 0, s, d, pb, p, x, r, red, st));
    } else {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_cg_update<kPB, Real, false>), dim3(d.blocks),
                                           dim3(kBlock), 0, s, d, pb, p, x, r, red, st));
    }
}

void launch_cg_update32(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* p,
                        float* x, float* r, CGState* st, Reduce red, int zero_x) {
    if (zero_x)
        MAZE_LAUNCH((k_cg_update<true, float, true, PBCoef32>), dim3(d.blocks), dim3(kBlock), 0, s,
                    d, pb, p, x, r, red, st);
    else
        MAZE_LAUNCH((k_cg_update<true, float, false, PBCoef32>), dim3(d.blocks), dim3(kBlock), 0,
                    s, d, pb, p, x, r, red, st);
}

template void launch_cg_init<double>(hipStream_t, const GridDims&, const PBCoef&, const double*,
                                     const double*, double*, CGState*, Reduce, double*, int);
template void launch_cg_init<float>(hipStream_t, const GridDims&, const PBCoef&, const float*,
                                    const float*, float*, CGState*, Reduce, float*, int);
template void launch_cg_dir<double>(hipStream_t, const GridDims&, const PBCoef&, const double*,
                                    const double*, double*, CGState*, Reduce, int);
template void launch_cg_dir<float>(hipStream_t, const GridDims&, const PBCoef&, const float*,
                                   const float*, float*, CGState*, Reduce, int);
template void launch_cg_update<double>(hipStream_t, const GridDims&, const PBCoef&, const double*,
                                       double*, double*, CGState*, Reduce, int);
template void launch_cg_update<float>(hipStream_t, const GridDims&, const PBCoef&, const float*,
                                      float*, float*, CGState*, Reduce, int);

void launch_dot_rv(hipStream_t s, const GridDims& d, const double* r, const double* v, CGState* st,
                   Reduce red, int first) {
    BodyDotRV body{r, v, first};
    run_pass<false>(s, d, PBCoef{}, SrcZero{}, body, red, st);
}

__global__ void k_cg_reset(CGState* st, double tol, double rtol, int limit, int count_mode,
                           int pcg, int gated, int plain) {
    if (threadIdx.x == 0 && blockIdx.x == 0) {
        if (gated && st->outer_done) {
            st->done = 1;
            return;
        }
        st->plain = plain;
        st->pcg = pcg;
        st->tol = tol;
        st->rtol = rtol;
        st->limit = limit;
        st->count_mode = count_mode;
        st->done = 0;
        st->res = -1;
        st->iter = 0;
        st->first = 1;
    }
}

void launch_cg_reset(hipStream_t s, CGState* st, double tol, double rtol, int limit,
                     int count_mode, int pcg, int gated, int plain) {
    MAZE_LAUNCH(k_cg_reset, dim3(1), dim3(64), 0, s, st, tol, rtol, limit, count_mode, pcg, gated,
                plain);
}

// The y buffers are chosen on the device (st->ysel), so a queued pass needs
// no host round trip; a pass after the outer loop converged returns at once.
// An inner solve that applied no update (stopped before its first one) left
// d untouched, so d then counts as zero.
// dps, dpm, dcnt: the partials of the inner solve's last update when its stop
// test is deferred (multigrid-PCG); every work-group applies it: a stop found
// there, or one recorded by an earlier kernel (st->done), ends the solve.  The
// outcome goes into the pass record of the last work-group's finalize, the
// only writer of the solver state in this kernel.
template <bool PB, bool Add>
__global__ __launch_bounds__(kBlock) void k_refine(GridDims d, PBCoef pb, double* y0, double* y1,
                                                   const float* dc, const double* sigma,
                                                   float* r32, RefineArgs a, Reduce red,
                                                   CGState* st, const double* dps,
                                                   const double* dpm, int dcnt) {
    int sel = 0;
    if (Add) {
        if (st->outer_done) return;
        sel = st->ysel;
    }
    // Read before the sweep: only the last work-group's finalize writes the
    // solver state in this kernel, after every work-group's sweep.
    const int done0 = Add ? st->done : 0, res0 = Add ? st->res : 0, iter0 = Add ? st->iter : 0;
    __shared__ double shf[3 * kBlock / kWave + 1];
    BodyRefine<Add, Add> bd{sel ? y0 : y1, sigma, r32, a.otol, a.rtol, a.limit, a.pcg};
    SrcSum<Add> src{sel ? y1 : y0, dc, 0.0};
    // The deferred stop test folds while the first plane's loads are in flight.
    march<PB>(d, pb, src, bd, [&] {
        if (Add) {
            int r = INT_MIN;
            if (dps) {
                double Z, S, M;
                int flag = 0;
                fold_partials<kBlock, false, true>(nullptr, 0, dps, dpm, dcnt, shf, Z, S, M,
                                                   flag);
                r = df_decide(st, S, M, /*record=*/false);
            }
            int in_done = 0, in_res = 0;
            if (r != INT_MIN) {
                in_done = 1;
                in_res = r;
            } else if (done0) {
                in_done = 1;
                in_res = res0;
            }
            bd.in_done = in_done;
            bd.in_res = in_res;
            src.dscale = (in_done ? in_res : iter0) > 0 ? 1.0 : 0.0;
        }
        return true;
    });
    reduce_and_finalize<true>(bd, red, st);
}

// Start of the loop from the extrapolated iterate y_out = 2 y - y_prev.
template <bool PB>
__global__ __launch_bounds__(kBlock) void k_refine_extrap(GridDims d, PBCoef pb, const double* y,
                                                          const double* y_prev, double* y_out,
                                                          const double* sigma, float* r32,
                                                          RefineArgs a, Reduce red, CGState* st) {
    BodyRefine<true, false> bd{y_out, sigma, r32, a.otol, a.rtol, a.limit, a.pcg};
    march<PB>(d, pb, SrcPred{y, y_prev}, bd);
    reduce_and_finalize<true>(bd, red, st);
}

void launch_refine_extrap(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* y,
                          const double* y_prev, double* y_out, const double* sigma, float* r32,
                          RefineArgs a, CGState* st, Reduce red) {
    DISPATCH_PB(pb, MAZE_LAUNCH((k_refine_extrap<kPB>), dim3(d.blocks), dim3(kBlock), 0, s,
                                       d, pb, y, y_prev, y_out, sigma, r32, a, red, st));
}

void launch_refine(hipStream_t s, const GridDims& d, const PBCoef& pb, double* y0, double* y1,
                   const float* dcorr, const double* sigma, float* r32, RefineArgs a, CGState* st,
                   Reduce red, int add_d, const DeferRed* defer) {
    const double* dps = add_d && defer ? defer->pus : nullptr;
    const double* dpm = add_d && defer ? defer->pum : nullptr;
    if (add_d) {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_refine<kPB, true>), dim3(d.blocks), dim3(kBlock), 0, s,
                                    d, pb, y0, y1, dcorr, sigma, r32, a, red, st, dps, dpm,
                                    d.blocks));
    } else {
        DISPATCH_PB(pb, MAZE_LAUNCH((k_refine<kPB, false>), dim3(d.blocks), dim3(kBlock), 0, s,
                                    d, pb, y0, y1, dcorr, sigma, r32, a, red, st, dps, dpm,
                                    d.blocks));
    }
}

void launch_true_residual_max(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* y,
                              const double* sigma, CGState* st, Reduce red) {
    BodyTrueRes body{sigma};
    DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcArr<double>{y}, body, red, st));
}

// ---- point-wise kernels ------------------------------------------------------
__global__ void k_axpy(long n, double alpha, const double* __restrict__ v, double* __restrict__ u) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        u[i] += alpha * v[i];
}
// phi -= y_final once the mixed-precision outer loop has converged, y_final
// being the y buffer the device selected (ysel); nothing otherwise.
__global__ void k_phi_update_gated(long n, const CGState* __restrict__ st,
                                   const double* __restrict__ ya, const double* __restrict__ yb,
                                   double* __restrict__ phi) {
    if (!st->outer_done) return;
    const double* __restrict__ v = st->ysel ? yb : ya;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        phi[i] += -1.0 * v[i];
}
__global__ void k_scale_copy(long n, double c, const double* __restrict__ in, double* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = in[i] * c;
}
// Copies the solver state and the charge sum to host-coherent memory, each
// thread fencing its stores at system scope before the work-group barrier, then
// publishes seq with a system-scope release store.
__global__ __launch_bounds__(kWave) void k_publish_state(const CGState* st, const double* q_sum,
                                                         StatePub* pub, unsigned int seq) {
    static_assert(sizeof(CGState) % 4 == 0, "CGState is copied as 32-bit words");
    const unsigned int* src = reinterpret_cast<const unsigned int*>(st);
    unsigned int* dst = reinterpret_cast<unsigned int*>(&pub->st);
    for (int i = threadIdx.x; i < (int)(sizeof(CGState) / 4); i += kWave) dst[i] = src[i];
    if (threadIdx.x == 0) pub->q_sum = *q_sum;
    __threadfence_system();
    __syncthreads();
    if (threadIdx.x == 0)
        __hip_atomic_store(&pub->seq, seq, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_SYSTEM);
}

void launch_publish_state(hipStream_t s, const CGState* st, const double* q_sum, StatePub* pub,
                          unsigned int seq) {
    MAZE_LAUNCH(k_publish_state, dim3(1), dim3(kWave), 0, s, st, q_sum, pub, seq);
}

__global__ void k_fill(long n, double v, double* out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = v;
}
__global__ void k_fill_f(long n, float v, float* out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = v;
}
__global__ void k_copy(long n, const double* __restrict__ in, double* __restrict__ out) {
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        out[i] = in[i];
}

static int pw_blocks(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)(b < 4096 ? (b < 1 ? 1 : b) : 4096);
}

void launch_axpy(hipStream_t s, long n, double alpha, const double* v, double* u) {
    MAZE_LAUNCH(k_axpy, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, alpha, v, u);
}

void launch_phi_update_gated(hipStream_t s, long n, const CGState* st, const double* ya,
                             const double* yb, double* phi) {
    MAZE_LAUNCH(k_phi_update_gated, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, st, ya, yb, phi);
}
void launch_scale_copy(hipStream_t s, long n, double c, const double* in, double* out) {
    MAZE_LAUNCH(k_scale_copy, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, c, in, out);
}
// ---- block-Jacobi preconditioner ------------------------------------------------
// v = A_B^-1 r on every 10^3 block of the grid (library/src/precond_blockjacobi.c:
// 14-69, 170-206), A_B the 7-point Laplacian of one block with zero values outside
// it.  The reference inverts A_B (1000 x 1000) and applies it as a dense product;
// here the block operator's eigenbasis applies the same inverse:
// A_B = T (x) I (x) I + I (x) T (x) I + I (x) I (x) T with T = tridiag(1, -2, 1) of
// order 10, whose orthonormal eigenvectors S[m][q] = sqrt(2/11) sin((m+1)(q+1) pi/11)
// form a symmetric orthogonal S, eigenvalues l_m = -4 sin^2((m+1) pi / 22).  So
// A_B^-1 = (S (x) S (x) S) diag(1 / (l_a + l_b + l_c)) (S (x) S (x) S): six passes of
// 10-term sums per point instead of a 1000-term one.  One work-group per block, one
// thread per point; tab holds S (100 values) then the 1000 inverse eigenvalue sums.
constexpr int kBJ = 10, kBJ3 = kBJ * kBJ * kBJ;

__global__ __launch_bounds__(1024) void k_blockjacobi(int n, const double* __restrict__ in,
                                                      double* __restrict__ out,
                                                      const double* __restrict__ tab) {
    __shared__ double u[kBJ3], w[kBJ3], S[kBJ * kBJ];
    const int t = threadIdx.x, nb = n / kBJ, b = blockIdx.x;
    const int bi = b / (nb * nb), bj = (b / nb) % nb, bk = b % nb;
    const bool act = t < kBJ3;
    const int a = act ? t / (kBJ * kBJ) : 0, c = act ? (t / kBJ) % kBJ : 0, e = act ? t % kBJ : 0;
    const long g = (long)(bi * kBJ + a) * n * n + (long)(bj * kBJ + c) * n + bk * kBJ + e;
    if (t < kBJ * kBJ) S[t] = tab[t];
    double il = 0.0;
    if (act) {
        u[t] = in[g];
        il = tab[kBJ * kBJ + t];
    }
    __syncthreads();
    // dst[a][c][e] = sum_q S[m][q] src[...q...], m the index of the pass's axis.
    auto pass = [&](const double* src, double* dst, int axis) {
        if (act) {
            double sum = 0.0;
            if (axis == 2) {
                const double* r = src + (a * kBJ + c) * kBJ;
#pragma unroll
                for (int q = 0; q < kBJ; q++) sum += S[e * kBJ + q] * r[q];
            } else if (axis == 1) {
                const double* r = src + a * kBJ * kBJ + e;
#pragma unroll
                for (int q = 0; q < kBJ; q++) sum += S[c * kBJ + q] * r[q * kBJ];
            } else {
                const double* r = src + c * kBJ + e;
#pragma unroll
                for (int q = 0; q < kBJ; q++) sum += S[a * kBJ + q] * r[q * kBJ * kBJ];
            }
            dst[t] = sum;
        }
        __syncthreads();
    };
    pass(u, w, 2);
    pass(w, u, 1);
    pass(u, w, 0);
    if (act) w[t] *= il;
    __syncthreads();
    pass(w, u, 2);
    pass(u, w, 1);
    pass(w, u, 0);
    if (act) out[g] = u[t];
}

void blockjacobi_tables(double* tab) {
    const double pi = 3.14159265358979323846;
    double lam[kBJ];
    for (int m = 0; m < kBJ; m++) {
        for (int q = 0; q < kBJ; q++)
            tab[m * kBJ + q] = std::sqrt(2.0 / (kBJ + 1)) * std::sin((m + 1) * (q + 1) * pi / (kBJ + 1));
        const double sn = std::sin((m + 1) * pi / (2.0 * (kBJ + 1)));
        lam[m] = -4.0 * sn * sn;
    }
    for (int a = 0; a < kBJ; a++)
        for (int c = 0; c < kBJ; c++)
            for (int e = 0; e < kBJ; e++)
                tab[kBJ * kBJ + (a * kBJ + c) * kBJ + e] = 1.0 / (lam[a] + lam[c] + lam[e]);
}

void launch_blockjacobi(hipStream_t s, int n, const double* in, double* out, const double* tab) {
    const int nb = n / kBJ;
    MAZE_LAUNCH(k_blockjacobi, dim3(nb * nb * nb), dim3(1024), 0, s, n, in, out, tab);
}

void launch_fill(hipStream_t s, long n, double v, double* out) {
    MAZE_LAUNCH(k_fill, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, v, out);
}
void launch_fill_f(hipStream_t s, long n, float v, float* out) {
    MAZE_LAUNCH(k_fill_f, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, v, out);
}
void launch_copy(hipStream_t s, long n, const double* in, double* out) {
    MAZE_LAUNCH(k_copy, dim3(pw_blocks(n)), dim3(kBlock), 0, s, n, in, out);
}

// out = sum(scale * v[i]) with the single-pass deterministic reduction; with
// scale 1/2 over phi this is grid_get_energy_elec (library/src/grid_general.c:304-316).
__global__ void k_sum_scaled(long n, double scale, const double* __restrict__ v, Reduce red,
                             double* out) {
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += scale * v[i];
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            *out = T;
            reset_arrivals(red.counter);
        }
    }
}

void launch_sum_scaled(hipStream_t s, long n, double scale, const double* v, double* out,
                       Reduce red) {
    long b = (n + 4L * kBlock - 1) / (4L * kBlock);
    if (b < 1) b = 1;
    if (b > 2048) b = 2048;
    MAZE_LAUNCH(k_sum_scaled, dim3((int)b), dim3(kBlock), 0, s, n, scale, v, red, out);
}

void launch_half_sum(hipStream_t s, long n, const double* phi, double* out, Reduce red) {
    launch_sum_scaled(s, n, 0.5, phi, out, red);
}

// out = scale * sum(a[i] * b[i]), single-pass deterministic reduction.
__global__ void k_dot_scaled(long n, double scale, const double* __restrict__ a,
                             const double* __restrict__ b, Reduce red, double* out) {
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < n; i += (long)gridDim.x * blockDim.x)
        s += a[i] * b[i];
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            *out = scale * T;
            reset_arrivals(red.counter);
        }
    }
}

void launch_dot_scaled(hipStream_t s, long n, double scale, const double* a, const double* b,
                       double* out, Reduce red) {
    long nb = (n + 4L * kBlock - 1) / (4L * kBlock);
    if (nb < 1) nb = 1;
    if (nb > 2048) nb = 2048;
    MAZE_LAUNCH(k_dot_scaled, dim3((int)nb), dim3(kBlock), 0, s, n, scale, a, b, red, out);
}

// ---- reference two-grid multigrid (library/src/multigrid.c) -----------------

// One colour of red-black Gauss-Seidel in place: points with (i+j+k)%2 ==
// color read only the other colour.  library/src/multigrid.c:663-727
// (neighbour order i-1, i+1, j-1, j+1, k-1, k+1).
__global__ __launch_bounds__(kBlock) void k_rbgs(GridDims d, const double* __restrict__ in,
                                                 double* out, int color) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int chunk = b % d.chunks;
    const int s = b / d.chunks;
    const int pl = chunk * kBlock + threadIdx.x;
    if (pl >= d.n2) return;
    const int n = d.n, n2 = d.n2;
    const int j = pl / n, k = pl - j * n;
    const PlaneNb nb = plane_nb(pl, n, n2);
    const int i0 = s * d.slab, i1 = min(n, i0 + d.slab);
    const double inv_diag = -1.0 / 6.0;
    for (int i = i0; i < i1; i++) {
        if (((i + j + k) & 1) != color) continue;
        const int base = i * n2, c = base + pl;
        const int im = (i == 0 ? n - 1 : i - 1) * n2 + pl;
        const int ip = (i == n - 1 ? 0 : i + 1) * n2 + pl;
        const double sum_nb = out[im] + out[ip] + out[base + nb.jm] + out[base + nb.jp] +
                              out[base + nb.km] + out[base + nb.kp];
        out[c] = (in[c] - sum_nb) * inv_diag;
    }
}

void launch_rbgs(hipStream_t s, const GridDims& d, const double* in, double* out, int iters) {
    for (int it = 0; it < iters; it++) {
        MAZE_LAUNCH(k_rbgs, dim3(d.blocks), dim3(kBlock), 0, s, d, in, out, 0);
        MAZE_LAUNCH(k_rbgs, dim3(d.blocks), dim3(kBlock), 0, s, d, in, out, 1);
    }
}

// One colour of the Poisson-Boltzmann RBGS in place: out = (sum of eps_f out_nb
// - in) / diag, diag = the six face coefficients + k2 floored at 1e-14.
// library/src/multigrid_pb.c:596-690 (term order :628-645).
__global__ __launch_bounds__(kBlock) void k_rbgs_pb(GridDims d, PBCoef pb,
                                                    const double* __restrict__ in, double* out,
                                                    int color) {
    const int b = xcd_tile(blockIdx.x, gridDim.x);
    const int chunk = b % d.chunks;
    const int s = b / d.chunks;
    const int pl = chunk * kBlock + threadIdx.x;
    if (pl >= d.n2) return;
    const int n = d.n, n2 = d.n2;
    const int j = pl / n, k = pl - j * n;
    const PlaneNb nb = plane_nb(pl, n, n2);
    const int i0 = s * d.slab, i1 = min(n, i0 + d.slab);
    for (int i = i0; i < i1; i++) {
        if (((i + j + k) & 1) != color) continue;
        const int base = i * n2, c = base + pl;
        const int im = (i == 0 ? n - 1 : i - 1) * n2 + pl;
        const int ip = (i == n - 1 ? 0 : i + 1) * n2 + pl;
        const int cy = base + nb.jm, cz = base + nb.km;
        const double ex_p = pb.ex[c], ex_m = pb.ex[im], ey_p = pb.ey[c], ey_m = pb.ey[cy];
        const double ez_p = pb.ez[c], ez_m = pb.ez[cz];
        double diag = ex_p + ex_m + ey_p + ey_m + ez_p + ez_m + pb.k2[c];
        if (diag < 1e-14) diag = 1e-14;
        const double rhs = ex_p * out[ip] + ex_m * out[im] + ey_p * out[base + nb.jp] +
                           ey_m * out[cy] + ez_p * out[base + nb.kp] + ez_m * out[cz] - in[c];
        out[c] = rhs / diag;
    }
}

void launch_rbgs_pb(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* in,
                    double* out, int iters) {
    for (int it = 0; it < iters; it++) {
        MAZE_LAUNCH(k_rbgs_pb, dim3(d.blocks), dim3(kBlock), 0, s, d, pb, in, out, 0);
        MAZE_LAUNCH(k_rbgs_pb, dim3(d.blocks), dim3(kBlock), 0, s, d, pb, in, out, 1);
    }
}

// Face coefficient of the coarse cell (I,J,K): 2x2 average over the plane
// normal to `axis` (x: fine plane 2I; y, z: fine planes 2I and 2I+1).
// library/src/multigrid_pb.c:375-500 (sum order per axis).
__global__ void k_restrict_eps(int n, int axis, const double* __restrict__ in,
                               double* __restrict__ out) {
    const int M = n / 2;
    const long total = (long)M * M * M;
    const long N = n, N2 = N * N;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int K = (int)(t % M), J = (int)((t / M) % M), I = (int)(t / ((long)M * M));
        const long i0 = 2L * I * N2, i1 = i0 + N2;
        const long j0 = 2L * J * N, j1 = (long)((2 * J + 1) % n) * N;
        const long k0 = 2 * K, k1 = (2 * K + 1) % n;
        double sm;
        if (axis == 0)
            sm = in[i0 + j0 + k0] + in[i0 + j0 + k1] + in[i0 + j1 + k0] + in[i0 + j1 + k1];
        else if (axis == 1)
            sm = in[i0 + j0 + k0] + in[i0 + j0 + k1] + in[i1 + j0 + k0] + in[i1 + j0 + k1];
        else
            sm = in[i0 + j0 + k0] + in[i1 + j0 + k0] + in[i0 + j1 + k0] + in[i1 + j1 + k0];
        out[t] = sm * 0.25;
    }
}

// Screening term of the coarse cell (I,J,K): 8-point average of the fine
// cells (2I+1, 2I+2) x (2J, 2J+1) x (2K, 2K+1) (the serial i offset is 1).
// library/src/multigrid_pb.c:503-568.
__global__ void k_restrict_k2(int n, const double* __restrict__ in, double* __restrict__ out) {
    const int M = n / 2;
    const long total = (long)M * M * M;
    const long N = n, N2 = N * N;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int K = (int)(t % M), J = (int)((t / M) % M), I = (int)(t / ((long)M * M));
        const long i0 = (2L * I + 1) * N2, i1 = (long)((2 * I + 2) % n) * N2;
        const long j0 = 2L * J * N, j1 = (long)((2 * J + 1) % n) * N;
        const long k0 = 2 * K, k1 = (2 * K + 1) % n;
        const double sum = in[i0 + j0 + k0] + in[i0 + j0 + k1] + in[i0 + j1 + k0] +
                           in[i0 + j1 + k1] + in[i1 + j0 + k0] + in[i1 + j0 + k1] +
                           in[i1 + j1 + k0] + in[i1 + j1 + k1];
        out[t] = (1.0 / 8.0) * sum;
    }
}

void launch_restrict_pb_coef(hipStream_t s, int n, const PBCoef& fine, double* ex, double* ey,
                             double* ez, double* k2) {
    const long M = n / 2;
    const int nb = pw_blocks(M * M * M);
    MAZE_LAUNCH(k_restrict_eps, dim3(nb), dim3(kBlock), 0, s, n, 0, fine.ex, ex);
    MAZE_LAUNCH(k_restrict_eps, dim3(nb), dim3(kBlock), 0, s, n, 1, fine.ey, ey);
    MAZE_LAUNCH(k_restrict_eps, dim3(nb), dim3(kBlock), 0, s, n, 2, fine.ez, ez);
    MAZE_LAUNCH(k_restrict_k2, dim3(nb), dim3(kBlock), 0, s, n, fine.k2, k2);
}

// r = in - A out, as (A out)(-1) + 1 in (library/src/multigrid.c:142-144).
struct BodyMgRes {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = false;
    const double* in;
    double* r;
    double s = 0.0, m = 0.0;
    using Raw = double;
    __device__ Raw load(int c) const { return in[c]; }
    __device__ void apply(int c, const Raw& ic, double, double lap) { r[c] = lap * -1.0 + 1.0 * ic; }
    __device__ static void finalize(CGState*, double, double) {}
};

void launch_mg_residual(hipStream_t s, const GridDims& d, const PBCoef& pb, const double* in,
                        const double* out, double* r) {
    BodyMgRes body{in, r};
    DISPATCH_PB(pb, run_pass<kPB>(s, d, pb, SrcArr<double>{out}, body, Reduce{}, nullptr));
}

// 27-point full weighting, coarse (I,J,K) <- fine (2I,2J,2K).
// library/src/multigrid.c:523-594 (sum order :571-590).
__global__ void k_restrict27(int n, const double* __restrict__ in, double* __restrict__ out) {
    const int M = n / 2;
    const long total = (long)M * M * M;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int K = (int)(t % M), J = (int)((t / M) % M), I = (int)(t / ((long)M * M));
        const int i = 2 * I, j = 2 * J, k = 2 * K;
        const int ip = (i + 1) % n, im = (i - 1 + n) % n;
        const int jp = (j + 1) % n, jm = (j - 1 + n) % n;
        const int kp = (k + 1) % n, km = (k - 1 + n) % n;
        const long N = n, N2 = N * N;
#define F(a, b, c) in[(a) * N2 + (b) * N + (c)]
        const double f = F(im, j, k) + F(ip, j, k) + F(i, jm, k) + F(i, jp, k) + F(i, j, km) + F(i, j, kp);
        const double e = F(ip, j, km) + F(ip, j, kp) + F(im, j, km) + F(im, j, kp) + F(ip, jm, k) +
                         F(ip, jp, k) + F(im, jm, k) + F(im, jp, k) + F(i, jm, km) + F(i, jm, kp) +
                         F(i, jp, km) + F(i, jp, kp);
        const double cc = F(ip, jp, km) + F(ip, jp, kp) + F(ip, jm, km) + F(ip, jm, kp) +
                          F(im, jp, km) + F(im, jp, kp) + F(im, jm, km) + F(im, jm, kp);
        out[t] = (8.0 * F(i, j, k) + 4.0 * f + 2.0 * e + cc) * (1.0 / 64.0);
#undef F
    }
}

void launch_restrict27(hipStream_t s, int n, const double* in, double* out) {
    const long M = n / 2;
    MAZE_LAUNCH(k_restrict27, dim3(pw_blocks(M * M * M)), dim3(kBlock), 0, s, n, in, out);
}

// out += trilinear(in) from an nc^3 grid to (2 nc)^3 (library/src/multigrid.c:393-463,
// fused with the correction daxpy of :158).
__global__ void k_prolong_add(int nc, const double* __restrict__ in, double* __restrict__ out) {
    const int N = 2 * nc;
    const long total = (long)N * N * N;
    const long M = nc, M2 = M * M;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % N), j = (int)((t / N) % N), i = (int)(t / ((long)N * N));
        const int a0 = (i / 2) % nc, a1 = (a0 + 1) % nc;
        const int b0 = (j / 2) % nc, b1 = (b0 + 1) % nc;
        const int c0 = (k / 2) % nc, c1 = (c0 + 1) % nc;
        const double wi0 = (i & 1) ? 0.5 : 1.0, wi1 = (i & 1) ? 0.5 : 0.0;
        const double wj0 = (j & 1) ? 0.5 : 1.0, wj1 = (j & 1) ? 0.5 : 0.0;
        const double wk0 = (k & 1) ? 0.5 : 1.0, wk1 = (k & 1) ? 0.5 : 0.0;
#define C(a, b, c) in[(a) * M2 + (b) * M + (c)]
        const double v = C(a0, b0, c0) * wi0 * wj0 * wk0 + C(a1, b0, c0) * wi1 * wj0 * wk0 +
                         C(a0, b1, c0) * wi0 * wj1 * wk0 + C(a1, b1, c0) * wi1 * wj1 * wk0 +
                         C(a0, b0, c1) * wi0 * wj0 * wk1 + C(a1, b0, c1) * wi1 * wj0 * wk1 +
                         C(a0, b1, c1) * wi0 * wj1 * wk1 + C(a1, b1, c1) * wi1 * wj1 * wk1;
#undef C
        out[t] += 1.0 * v;
    }
}

void launch_prolong_add(hipStream_t s, int nc, const double* in, double* out) {
    const long N = 2L * nc;
    MAZE_LAUNCH(k_prolong_add, dim3(pw_blocks(N * N * N)), dim3(kBlock), 0, s, nc, in, out);
}

}  // namespace maze

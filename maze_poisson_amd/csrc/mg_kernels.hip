// mg_kernels.hip -- FP32 geometric multigrid V-cycle used as the preconditioner
// of the mixed-precision CG (FP32 storage, FP64 arithmetic in registers).
//
// V(2,2) cycle for A d = b with A the unscaled periodic 7-point operator
// (negative semi-definite): damped Jacobi d <- d - w/6 (b - A d), 27-point
// full-weighting restriction with the (H/h)^2 = 4 factor of the unscaled
// coarse operator, trilinear prolongation, and an exact FP64 solve on the
// coarsest grid (<= 12^3) in the separable eigenbasis of the periodic
// operator, so the cycle is a fixed symmetric linear operator and can
// precondition CG.  Every kernel is a no-op once the solve has stopped.
#include "stencil.hpp"

namespace maze {

struct SrcNegScaled {  // first Jacobi sweep from zero: d1 = -c b (rounded to FP32)
    const float* b;
    double c;
    __device__ __forceinline__ double operator()(int i) const {
        return (double)(float)(-c * (double)b[i]);
    }
};

// Two pre-smoothing sweeps from d = 0 in one pass: d2 = d1 - c (b - A d1).
struct BodyMgPre {
    static constexpr bool kReduce = false, kSkipWhenDone = true;
    const float* b;
    float* d;
    double c;
    double s = 0.0, m = 0.0;
    __device__ void operator()(int i, double v, double lap) {
        d[i] = (float)(v - c * ((double)b[i] - lap));
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// r = b - A d
struct BodyMgResid {
    static constexpr bool kReduce = false, kSkipWhenDone = true;
    const float* b;
    float* r;
    double s = 0.0, m = 0.0;
    __device__ void operator()(int i, double, double lap) { r[i] = (float)((double)b[i] - lap); }
    __device__ static void finalize(CGState*, double, double) {}
};

// One post-smoothing sweep d_out = d - c (b - A d).
struct BodyMgPost {
    static constexpr bool kReduce = false, kSkipWhenDone = true;
    const float* b;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    __device__ void operator()(int i, double v, double lap) {
        dout[i] = (float)(v - c * ((double)b[i] - lap));
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// Last post-smoothing sweep on the finest level, fused with <r, v> (r is the
// CG residual = the cycle's right-hand side) and the PCG beta update.
struct BodyMgPostDot {
    static constexpr bool kReduce = true, kSkipWhenDone = true;
    const float* b;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    __device__ void operator()(int i, double v, double lap) {
        const float z = (float)(v - c * ((double)b[i] - lap));
        dout[i] = z;
        s += (double)b[i] * (double)z;
    }
    __device__ static void finalize(CGState* st, double S, double) {
        st->beta = st->first ? 0.0 : S / st->rdv;
        st->rdv = S;
    }
};

void launch_mg_pre(hipStream_t s, const GridDims& d, const float* b, float* dout, double c,
                   CGState* st) {
    run_pass<false>(s, d, PBCoef{}, SrcNegScaled{b, c}, BodyMgPre{b, dout, c}, Reduce{}, st);
}

void launch_mg_resid(hipStream_t s, const GridDims& d, const float* b, const float* din, float* r,
                     CGState* st) {
    run_pass<false>(s, d, PBCoef{}, SrcArr<float>{din}, BodyMgResid{b, r}, Reduce{}, st);
}

void launch_mg_post(hipStream_t s, const GridDims& d, const float* b, const float* din,
                    float* dout, double c, CGState* st, int dot, Reduce red) {
    if (dot)
        run_pass<false>(s, d, PBCoef{}, SrcArr<float>{din}, BodyMgPostDot{b, dout, c}, red, st);
    else
        run_pass<false>(s, d, PBCoef{}, SrcArr<float>{din}, BodyMgPost{b, dout, c}, Reduce{}, st);
}

static int pw_blocks_mg(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > 8192 ? 8192 : b));
}

// b_H = 4 * (27-point full weighting of r), coarse (I,J,K) <- fine (2I,2J,2K).
__global__ void k_mg_restrict(int n, const float* __restrict__ in, float* __restrict__ out,
                              const CGState* st) {
    if (st->done) return;
    const int M = n / 2;
    const long total = (long)M * M * M;
    const long N = n, N2 = N * N;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int K = (int)(t % M), J = (int)((t / M) % M), I = (int)(t / ((long)M * M));
        const int i = 2 * I, j = 2 * J, k = 2 * K;
        const int ii[3] = {(i - 1 + n) % n, i, (i + 1) % n};
        const int jj[3] = {(j - 1 + n) % n, j, (j + 1) % n};
        const int kk[3] = {(k - 1 + n) % n, k, (k + 1) % n};
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) {
                    const double w = (double)((a == 1 ? 2 : 1) * (bb == 1 ? 2 : 1) * (cc == 1 ? 2 : 1));
                    acc += w * (double)in[ii[a] * N2 + jj[bb] * N + kk[cc]];
                }
        out[t] = (float)(acc * (4.0 / 64.0));
    }
}

void launch_mg_restrict(hipStream_t s, int n, const float* r, float* bc, CGState* st) {
    const long M = n / 2;
    hipLaunchKernelGGL(k_mg_restrict, dim3(pw_blocks_mg(M * M * M)), dim3(kBlock), 0, s, n, r, bc, st);
}

// d += trilinear(e) from an nc^3 grid to (2 nc)^3.
__global__ void k_mg_prolong_add(int nc, const float* __restrict__ e, float* __restrict__ d,
                                 const CGState* st) {
    if (st->done) return;
    const int N = 2 * nc;
    const long total = (long)N * N * N;
    const long M = nc, M2 = M * M;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % N), j = (int)((t / N) % N), i = (int)(t / ((long)N * N));
        const int a0 = i >> 1, a1 = (a0 + 1) % nc;
        const int b0 = j >> 1, b1 = (b0 + 1) % nc;
        const int c0 = k >> 1, c1 = (c0 + 1) % nc;
        const double wi = (i & 1) ? 0.5 : 0.0, wj = (j & 1) ? 0.5 : 0.0, wk = (k & 1) ? 0.5 : 0.0;
#define C(a, b, c) (double)e[(a) * M2 + (b) * M + (c)]
        const double v00 = C(a0, b0, c0) + wk * (C(a0, b0, c1) - C(a0, b0, c0));
        const double v01 = C(a0, b1, c0) + wk * (C(a0, b1, c1) - C(a0, b1, c0));
        const double v10 = C(a1, b0, c0) + wk * (C(a1, b0, c1) - C(a1, b0, c0));
        const double v11 = C(a1, b1, c0) + wk * (C(a1, b1, c1) - C(a1, b1, c0));
#undef C
        const double v0 = v00 + wj * (v01 - v00);
        const double v1 = v10 + wj * (v11 - v10);
        d[t] = (float)((double)d[t] + v0 + wi * (v1 - v0));
    }
}

void launch_mg_prolong_add(hipStream_t s, int nc, const float* ec, float* d, CGState* st) {
    const long N = 2L * nc;
    hipLaunchKernelGGL(k_mg_prolong_add, dim3(pw_blocks_mg(N * N * N)), dim3(kBlock), 0, s, nc, ec,
                       d, st);
}

// Exact FP64 solve of A e = b on the coarsest c^3 grid (c <= 12): forward
// transform along k, j, i with the real orthonormal eigenbasis Qf (row t =
// eigenvector t), divide by the eigenvalue sum (the constant mode is
// dropped), transform back.  One work-group, data in LDS.
constexpr int kMaxCoarse = 12;

__global__ __launch_bounds__(kBlock) void k_mg_coarse(int c, const float* __restrict__ b,
                                                      float* __restrict__ e,
                                                      const double* __restrict__ Qf,
                                                      const double* __restrict__ ilam,
                                                      const CGState* st) {
    if (st->done) return;
    __shared__ double A[kMaxCoarse * kMaxCoarse * kMaxCoarse];
    __shared__ double B[kMaxCoarse * kMaxCoarse * kMaxCoarse];
    __shared__ double Q[kMaxCoarse * kMaxCoarse];
    const int N = c * c * c, c2 = c * c;
    for (int t = threadIdx.x; t < c2; t += kBlock) Q[t] = Qf[t];
    for (int t = threadIdx.x; t < N; t += kBlock) A[t] = (double)b[t];
    __syncthreads();
    // forward k: B[i][j][t] = sum_m Q[t][m] A[i][j][m]
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int t = x % c, line = x - t;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[line + m];
        B[x] = s;
    }
    __syncthreads();
    // forward j
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int k = x % c, t = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * B[i * c2 + m * c + k];
        A[x] = s;
    }
    __syncthreads();
    // forward i, scale by the inverse eigenvalue
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int jk = x % c2, t = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[m * c2 + jk];
        B[x] = s * ilam[x];
    }
    __syncthreads();
    // inverse i
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int jk = x % c2, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + i] * B[m * c2 + jk];
        A[x] = s;
    }
    __syncthreads();
    // inverse j
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int k = x % c, j = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + j] * A[i * c2 + m * c + k];
        B[x] = s;
    }
    __syncthreads();
    // inverse k
    for (int x = threadIdx.x; x < N; x += kBlock) {
        const int k = x % c, line = x - k;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + k] * B[line + m];
        e[x] = (float)s;
    }
}

void launch_mg_coarse(hipStream_t s, int c, const float* b, float* e, const double* Qf,
                      const double* ilam, CGState* st) {
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kBlock), 0, s, c, b, e, Qf, ilam, st);
}

// PCG direction p = -v (first) or beta p - v, with <p, A p>
// (library/src/plcg.c:130-155 form, v = M^-1 r).
struct SrcPcgDir {
    const float* p;
    const float* v;
    double beta;
    int first;
    __device__ __forceinline__ double operator()(int i) const {
        const double w = first ? -(double)v[i] : (double)p[i] * beta - (double)v[i];
        return (double)(float)w;
    }
};

struct BodyPcgDir {
    static constexpr bool kReduce = true, kSkipWhenDone = true;
    float* p_new;
    double s = 0.0, m = 0.0;
    __device__ void operator()(int c, double v, double lap) {
        p_new[c] = (float)v;
        s += v * lap;
    }
    __device__ static void finalize(CGState* st, double S, double) {
        st->pAp = S;
        st->first = 0;
        if (fabs(S) < 1e-300) {
            st->done = 1;
            st->res = st->iter;
            st->alpha = 0.0;
        } else {
            st->alpha = st->rdv / S;
        }
    }
};

__global__ __launch_bounds__(kBlock) void k_pcg_dir(GridDims d, const float* p_old, const float* v,
                                                    float* p_new, Reduce red, CGState* st) {
    if (st->done) return;
    const SrcPcgDir src{p_old, v, st->beta, st->first};
    BodyPcgDir bd{p_new};
    march<false>(d, PBCoef{}, src, bd);
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(bd.s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            BodyPcgDir::finalize(st, T, 0.0);
            reset_counter(red.counter);
        }
    }
}

void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red) {
    hipLaunchKernelGGL(k_pcg_dir, dim3(d.blocks), dim3(kBlock), 0, s, d, p_old, v, p_new, red, st);
}

}  // namespace maze

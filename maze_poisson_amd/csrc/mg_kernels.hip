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
    using Raw = float;
    __device__ __forceinline__ Raw load(int i) const { return b[i]; }
    __device__ __forceinline__ double val(const Raw& r) const {
        return (double)(float)(-c * (double)r);
    }
};

// Two pre-smoothing sweeps from d = 0 in one pass: d2 = d1 - c (b - A d1).
struct BodyMgPre {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = true;
    const float* b;
    float* d;
    double c;
    double s = 0.0, m = 0.0;
    using Raw = float;
    __device__ Raw load(int i) const { return b[i]; }
    __device__ void apply(int i, const Raw& bi, double v, double lap) {
        d[i] = (float)(v - c * ((double)bi - lap));
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// r = b - A d
struct BodyMgResid {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = true;
    const float* b;
    float* r;
    double s = 0.0, m = 0.0;
    using Raw = float;
    __device__ Raw load(int i) const { return b[i]; }
    __device__ void apply(int i, const Raw& bi, double, double lap) {
        r[i] = (float)((double)bi - lap);
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// One post-smoothing sweep d_out = d - c (b - A d).
struct BodyMgPost {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = true;
    const float* b;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    using Raw = float;
    __device__ Raw load(int i) const { return b[i]; }
    __device__ void apply(int i, const Raw& bi, double v, double lap) {
        dout[i] = (float)(v - c * ((double)bi - lap));
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// Last post-smoothing sweep on the finest level, fused with <r, v> (r is the
// CG residual = the cycle's right-hand side) and the PCG beta update.
struct BodyMgPostDot {
    static constexpr bool kReduce = true, kMax = false, kSkipWhenDone = true;
    const float* b;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    using Raw = float;
    __device__ Raw load(int i) const { return b[i]; }
    __device__ void apply(int i, const Raw& bi, double v, double lap) {
        const float z = (float)(v - c * ((double)bi - lap));
        dout[i] = z;
        s += (double)bi * (double)z;
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
        const int ii[3] = {(i == 0 ? n - 1 : i - 1), i, i + 1};
        const int jj[3] = {(j == 0 ? n - 1 : j - 1), j, j + 1};
        const int kk[3] = {(k == 0 ? n - 1 : k - 1), k, k + 1};
        float v[27];
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) v[(a * 3 + bb) * 3 + cc] = in[ii[a] * N2 + jj[bb] * N + kk[cc]];
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) {
                    const double w = (double)((a == 1 ? 2 : 1) * (bb == 1 ? 2 : 1) * (cc == 1 ? 2 : 1));
                    acc += w * (double)v[(a * 3 + bb) * 3 + cc];
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
        const int a0 = i >> 1, a1 = (a0 + 1 == nc) ? 0 : a0 + 1;
        const int b0 = j >> 1, b1 = (b0 + 1 == nc) ? 0 : b0 + 1;
        const int c0 = k >> 1, c1 = (c0 + 1 == nc) ? 0 : c0 + 1;
        const double wi = (i & 1) ? 0.5 : 0.0, wj = (j & 1) ? 0.5 : 0.0, wk = (k & 1) ? 0.5 : 0.0;
        const float e000 = e[a0 * M2 + b0 * M + c0], e001 = e[a0 * M2 + b0 * M + c1];
        const float e010 = e[a0 * M2 + b1 * M + c0], e011 = e[a0 * M2 + b1 * M + c1];
        const float e100 = e[a1 * M2 + b0 * M + c0], e101 = e[a1 * M2 + b0 * M + c1];
        const float e110 = e[a1 * M2 + b1 * M + c0], e111 = e[a1 * M2 + b1 * M + c1];
        const float dt = d[t];
        const double v00 = e000 + wk * ((double)e001 - e000);
        const double v01 = e010 + wk * ((double)e011 - e010);
        const double v10 = e100 + wk * ((double)e101 - e100);
        const double v11 = e110 + wk * ((double)e111 - e110);
        const double v0 = v00 + wj * (v01 - v00);
        const double v1 = v10 + wj * (v11 - v10);
        d[t] = (float)((double)dt + v0 + wi * (v1 - v0));
    }
}

void launch_mg_prolong_add(hipStream_t s, int nc, const float* ec, float* d, CGState* st) {
    const long N = 2L * nc;
    hipLaunchKernelGGL(k_mg_prolong_add, dim3(pw_blocks_mg(N * N * N)), dim3(kBlock), 0, s, nc, ec,
                       d, st);
}

// ---- exact FP64 solve on the coarsest grid -----------------------------------
// A e = b on a c^3 grid (c <= 12): forward transform along k, j, i with the
// real orthonormal eigenbasis Qf (row t = eigenvector t), divide by the
// eigenvalue sum (the constant mode is dropped), transform back.  A and B
// are c^3 double work arrays in LDS; `nt` threads cooperate.
constexpr int kMaxCoarse = 12;

__device__ void coarse_exact(int c, const double* Q, const double* __restrict__ ilam, double* A,
                             double* B, int tid, int nt) {
    const int N = c * c * c, c2 = c * c;
    for (int x = tid; x < N; x += nt) {  // forward k
        const int t = x % c, line = x - t;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[line + m];
        B[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < N; x += nt) {  // forward j
        const int k = x % c, t = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * B[i * c2 + m * c + k];
        A[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < N; x += nt) {  // forward i, scale by the inverse eigenvalue
        const int jk = x % c2, t = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[m * c2 + jk];
        B[x] = s * ilam[x];
    }
    __syncthreads();
    for (int x = tid; x < N; x += nt) {  // inverse i
        const int jk = x % c2, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + i] * B[m * c2 + jk];
        A[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < N; x += nt) {  // inverse j
        const int k = x % c, j = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + j] * A[i * c2 + m * c + k];
        B[x] = s;
    }
    __syncthreads();
    for (int x = tid; x < N; x += nt) {  // inverse k
        const int k = x % c, line = x - k;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + k] * B[line + m];
        A[x] = s;
    }
    __syncthreads();
}

__global__ __launch_bounds__(kBlock) void k_mg_coarse(int c, const float* __restrict__ b,
                                                      float* __restrict__ e,
                                                      const double* __restrict__ Qf,
                                                      const double* __restrict__ ilam,
                                                      const CGState* st) {
    if (st->done) return;
    __shared__ double A[kMaxCoarse * kMaxCoarse * kMaxCoarse];
    __shared__ double B[kMaxCoarse * kMaxCoarse * kMaxCoarse];
    __shared__ double Q[kMaxCoarse * kMaxCoarse];
    const int N = c * c * c;
    for (int t = threadIdx.x; t < c * c; t += kBlock) Q[t] = Qf[t];
    for (int t = threadIdx.x; t < N; t += kBlock) A[t] = (double)b[t];
    __syncthreads();
    coarse_exact(c, Q, ilam, A, B, threadIdx.x, kBlock);
    for (int t = threadIdx.x; t < N; t += kBlock) e[t] = (float)A[t];
}

void launch_mg_coarse(hipStream_t s, int c, const float* b, float* e, const double* Qf,
                      const double* ilam, CGState* st) {
    hipLaunchKernelGGL(k_mg_coarse, dim3(1), dim3(kBlock), 0, s, c, b, e, Qf, ilam, st);
}

// ---- the coarse tail of the V-cycle in one work-group --------------------------
// Levels of at most 16^3 points are launch-latency bound when each pass is its
// own kernel.  k_mg_tail copies the right-hand side of the first such level
// into LDS and runs the rest of the cycle there -- pre-smooth, residual and
// restriction down to the exact FP64 solve, prolongation and post-smoothing
// back up -- then writes that level's correction back; barriers separate the
// passes.  LDS holds b, d, t of every level (at most 3 * (16^3 + 8^3 + ...)
// floats); the exact solve's two double arrays reuse the first level's t.
constexpr int kTailThreads = 1024;
constexpr int kTailMaxLevels = 6;
constexpr int kTailMaxPts = 4096;  // 16^3
constexpr int kTailPool = 7168;    // doubles: 57344 bytes

struct TailLevels {
    int nlev;  // levels handled, the last one solved exactly
    int n[kTailMaxLevels];
};

__device__ __forceinline__ void nb6(int x, int m, int* o) {
    const int m2 = m * m;
    const int k = x % m, j = (x / m) % m, i = x / m2;
    o[0] = (i == 0 ? x + (m - 1) * m2 : x - m2);
    o[1] = (i == m - 1 ? x - (m - 1) * m2 : x + m2);
    o[2] = (j == 0 ? x + (m - 1) * m : x - m);
    o[3] = (j == m - 1 ? x - (m - 1) * m : x + m);
    o[4] = (k == 0 ? x + m - 1 : x - 1);
    o[5] = (k == m - 1 ? x - m + 1 : x + 1);
}

__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailLevels T, const float* __restrict__ bin,
                                                         float* __restrict__ dout, double c,
                                                         const double* __restrict__ Qf,
                                                         const double* __restrict__ ilam,
                                                         const CGState* st) {
    if (st->done) return;
    __shared__ double pool[kTailPool];
    __shared__ double Q[kMaxCoarse * kMaxCoarse];
    float* f = reinterpret_cast<float*>(pool);
    float *B[kTailMaxLevels], *D[kTailMaxLevels], *R[kTailMaxLevels];
    int off = 0;
    for (int l = 0; l < T.nlev; l++) {
        const int g = T.n[l] * T.n[l] * T.n[l];
        B[l] = f + off;
        D[l] = f + off + g;
        R[l] = f + off + 2 * g;
        off += 3 * g;
    }
    const int tid = threadIdx.x;
    {
        const int g0 = T.n[0] * T.n[0] * T.n[0];
        for (int x = tid; x < g0; x += kTailThreads) B[0][x] = bin[x];
    }
    __syncthreads();
    int o[6];
    // ---- down: pre-smooth, residual, restrict ----
    for (int l = 0; l < T.nlev - 1; l++) {
        const int m = T.n[l], G = m * m * m;
        const float* b = B[l];
        float* d = D[l];
        float* t = R[l];
        for (int x = tid; x < G; x += kTailThreads) {
            nb6(x, m, o);
            const double d1 = (double)(float)(-c * (double)b[x]);
            double lap = -6.0 * d1;
            for (int q = 0; q < 6; q++) lap += (double)(float)(-c * (double)b[o[q]]);
            d[x] = (float)(d1 - c * ((double)b[x] - lap));
        }
        __syncthreads();
        for (int x = tid; x < G; x += kTailThreads) {
            nb6(x, m, o);
            double lap = -6.0 * (double)d[x];
            for (int q = 0; q < 6; q++) lap += (double)d[o[q]];
            t[x] = (float)((double)b[x] - lap);
        }
        __syncthreads();
        const int M = m / 2, Gc = M * M * M;
        float* bc = B[l + 1];
        for (int x = tid; x < Gc; x += kTailThreads) {
            const int K = x % M, J = (x / M) % M, I = x / (M * M);
            const int i = 2 * I, j = 2 * J, k = 2 * K;
            double acc = 0.0;
            for (int a = -1; a <= 1; a++) {
                const int ii = (i + a + m) % m;
                for (int bb = -1; bb <= 1; bb++) {
                    const int jj = (j + bb + m) % m;
                    for (int cc = -1; cc <= 1; cc++) {
                        const int kk = (k + cc + m) % m;
                        const double w = (double)((a == 0 ? 2 : 1) * (bb == 0 ? 2 : 1) * (cc == 0 ? 2 : 1));
                        acc += w * (double)t[(ii * m + jj) * m + kk];
                    }
                }
            }
            bc[x] = (float)(acc * (4.0 / 64.0));
        }
        __syncthreads();
    }
    // ---- exact FP64 solve on the coarsest level ----
    {
        const int L = T.nlev - 1, cn = T.n[L], N = cn * cn * cn;
        double* A;
        double* Bd;
        if (T.nlev >= 2) {
            A = reinterpret_cast<double*>(R[0]);  // 4 m0^3 bytes >= 16 (m0/2)^3
        } else {
            A = pool + N;  // after b, d of the only level (2N floats = N doubles)
        }
        Bd = A + N;
        for (int x = tid; x < cn * cn; x += kTailThreads) Q[x] = Qf[x];
        for (int x = tid; x < N; x += kTailThreads) A[x] = (double)B[L][x];
        __syncthreads();
        coarse_exact(cn, Q, ilam, A, Bd, tid, kTailThreads);
        for (int x = tid; x < N; x += kTailThreads) D[L][x] = (float)A[x];
        __syncthreads();
    }
    // ---- up: prolong-correct, two post-smoothing sweeps ----
    for (int l = T.nlev - 2; l >= 0; l--) {
        const int m = T.n[l], G = m * m * m, nc = T.n[l + 1];
        const float* b = B[l];
        float* d = D[l];
        float* t = R[l];
        const float* e = D[l + 1];
        const int M2 = nc * nc;
        for (int x = tid; x < G; x += kTailThreads) {
            const int k = x % m, j = (x / m) % m, i = x / (m * m);
            const int a0 = i >> 1, a1 = (a0 + 1) % nc;
            const int b0 = j >> 1, b1 = (b0 + 1) % nc;
            const int c0 = k >> 1, c1 = (c0 + 1) % nc;
            const double wi = (i & 1) ? 0.5 : 0.0, wj = (j & 1) ? 0.5 : 0.0, wk = (k & 1) ? 0.5 : 0.0;
#define C(a, bb, cc) (double)e[(a) * M2 + (bb) * nc + (cc)]
            const double v00 = C(a0, b0, c0) + wk * (C(a0, b0, c1) - C(a0, b0, c0));
            const double v01 = C(a0, b1, c0) + wk * (C(a0, b1, c1) - C(a0, b1, c0));
            const double v10 = C(a1, b0, c0) + wk * (C(a1, b0, c1) - C(a1, b0, c0));
            const double v11 = C(a1, b1, c0) + wk * (C(a1, b1, c1) - C(a1, b1, c0));
#undef C
            const double v0 = v00 + wj * (v01 - v00);
            const double v1 = v10 + wj * (v11 - v10);
            d[x] = (float)((double)d[x] + v0 + wi * (v1 - v0));
        }
        __syncthreads();
        for (int x = tid; x < G; x += kTailThreads) {
            nb6(x, m, o);
            double lap = -6.0 * (double)d[x];
            for (int q = 0; q < 6; q++) lap += (double)d[o[q]];
            t[x] = (float)((double)d[x] - c * ((double)b[x] - lap));
        }
        __syncthreads();
        for (int x = tid; x < G; x += kTailThreads) {
            nb6(x, m, o);
            double lap = -6.0 * (double)t[x];
            for (int q = 0; q < 6; q++) lap += (double)t[o[q]];
            d[x] = (float)((double)t[x] - c * ((double)b[x] - lap));
        }
        __syncthreads();
    }
    {
        const int g0 = T.n[0] * T.n[0] * T.n[0];
        for (int x = tid; x < g0; x += kTailThreads) dout[x] = D[0][x];
    }
}

bool mg_tail_fits(int nlev, const int* n) {
    if (nlev < 1 || nlev > kTailMaxLevels) return false;
    if ((long)n[0] * n[0] * n[0] > kTailMaxPts) return false;
    long floats = 0;
    for (int l = 0; l < nlev; l++) floats += 3L * n[l] * n[l] * n[l];
    const long c3 = (long)n[nlev - 1] * n[nlev - 1] * n[nlev - 1];
    if (nlev == 1) floats = 2 * c3 + 2 + 4 * c3;  // b, d, then two double arrays
    return floats <= 2L * kTailPool && n[nlev - 1] <= kMaxCoarse;
}

void launch_mg_tail(hipStream_t s, int nlev, const int* n, const float* b, float* d, double c,
                    const double* Qf, const double* ilam, CGState* st) {
    TailLevels T;
    T.nlev = nlev;
    for (int l = 0; l < nlev; l++) T.n[l] = n[l];
    hipLaunchKernelGGL(k_mg_tail, dim3(1), dim3(kTailThreads), 0, s, T, b, d, c, Qf, ilam, st);
}

// PCG direction p = -v (first) or beta p - v, with <p, A p>
// (library/src/plcg.c:130-155 form, v = M^-1 r).
template <bool First>
struct SrcPcgDir {
    const float* p;
    const float* v;
    double beta;
    struct Raw { float p, v; };
    __device__ __forceinline__ Raw load(int i) const {
        if (First) return Raw{0.0f, v[i]};
        return Raw{p[i], v[i]};
    }
    __device__ __forceinline__ double val(const Raw& x) const {
        const double w = First ? -(double)x.v : (double)x.p * beta - (double)x.v;
        return (double)(float)w;
    }
};

struct BodyPcgDir {
    float* p_new;
    double s = 0.0, m = 0.0;
    using Raw = NoRaw;
    __device__ Raw load(int) const { return Raw{}; }
    __device__ void apply(int c, const Raw&, double v, double lap) {
        p_new[c] = (float)v;
        s += v * lap;
    }
    __device__ static void finalize(CGState* st, double S) {
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

template <bool First>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(GridDims d, const float* p_old, const float* v,
                                                    float* p_new, Reduce red, CGState* st) {
    if (st->done) return;
    const SrcPcgDir<First> src{p_old, v, First ? 0.0 : st->beta};
    BodyPcgDir bd{p_new};
    march<false>(d, PBCoef{}, src, bd);
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(bd.s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            BodyPcgDir::finalize(st, T);
            reset_counter(red.counter);
        }
    }
}

void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red, int first) {
    if (first)
        hipLaunchKernelGGL(k_pcg_dir<true>, dim3(d.blocks), dim3(kBlock), 0, s, d, p_old, v, p_new,
                           red, st);
    else
        hipLaunchKernelGGL(k_pcg_dir<false>, dim3(d.blocks), dim3(kBlock), 0, s, d, p_old, v,
                           p_new, red, st);
}

}  // namespace maze

// mg_kernels.hip -- FP32 geometric multigrid V-cycle used as the preconditioner
// of the mixed-precision CG (FP32 storage and arithmetic; dot products and
// CG scalars in FP64).
//
// V(2,2) cycle for A d = b with A the unscaled periodic 7-point operator
// (negative semi-definite): damped Jacobi d <- d - w/6 (b - A d), 27-point
// full-weighting restriction with the (H/h)^2 = 4 factor of the unscaled
// coarse operator, trilinear prolongation, and an exact solve on the
// coarsest grid in the separable eigenbasis of the periodic operator (FP32
// in one work-group up to 16^3, FP64 in the one-work-group tail and by axis
// passes above 16^3), so the cycle is a fixed symmetric linear operator and
// can precondition CG.  Every kernel is a no-op once the solve has stopped.
#include <cstdlib>

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
    MAZE_LAUNCH(k_mg_restrict, dim3(pw_blocks_mg(M * M * M)), dim3(kBlock), 0, s, n, r, bc, st);
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
    MAZE_LAUNCH(k_mg_prolong_add, dim3(pw_blocks_mg(N * N * N)), dim3(kBlock), 0, s, nc, ec,
                       d, st);
}

// ---- fused, LDS-tiled level kernels ------------------------------------------
// Each work-group owns a (TI, TJ, TK) tile of a fine level (even sizes) and
// loads its operands once, with the halo the fused steps need, into LDS; the
// steps then run out of LDS with redundant halo arithmetic (neighbouring tiles
// compute the same halo values from the same inputs).  Arithmetic is FP32,
// as is the storage: the cycle only preconditions the FP32 inner solve.
//
// down: d2 = two damped-Jacobi sweeps from zero on A d = b, which is the
//       stencil d2 = -2c b - c^2 A b; written for the owned points; then
//       b_H = 4 * full-weighting(b - A d2) for the tile's coarse points.
//       Halo: b on [-3, T+2), d2 on [-2, T+1), residual on [-1, T).
// up:   d = d2 + trilinear(e), then two sweeps d <- d - c (b - A d); the
//       finest level also forms <b, d> (PCG) and the beta update.
//       Halo: d on [-2, T+2), b and the first sweep on [-1, T+1), e on the
//       coarse points [T0/2 - 1, T0/2 + T/2 + 2).

__device__ __forceinline__ int wrap_mod(int v, int n) {
    v %= n;
    return v < 0 ? v + n : v;
}

// Calls f(x) for x = threadIdx.x + u NT < N (NT threads), every iteration
// unrolled (N is a compile-time constant) so that the loads of all
// iterations are in flight together.
template <int N, int NT, class F>
__device__ __forceinline__ void tile_for(F&& f) {
#pragma unroll
    for (int u = 0; u < (N + NT - 1) / NT; u++) {
        const int x = (int)threadIdx.x + u * NT;
        if (x < N) f(x);
    }
}

// A staging copy of N elements in tile_for's mapping, in two halves: load()
// into registers and store() to LDS, so that a kernel issues the loads of all
// its regions before the first store (a load-and-store loop waits for each
// load in turn).  f(x) must read a valid address for every x < N; the threads
// past N load element N - 1 again and skip the store.
template <int N, int NT, class T>
struct TileStage {
    static constexpr int U = (N + NT - 1) / NT;
    T v[U];
    template <class F>
    __device__ __forceinline__ void load(F&& f) {
#pragma unroll
        for (int u = 0; u < U; u++) v[u] = f(min((int)threadIdx.x + u * NT, N - 1));
    }
    __device__ __forceinline__ void store(T* lds) const {
#pragma unroll
        for (int u = 0; u < U; u++) {
            const int x = (int)threadIdx.x + u * NT;
            if (x < N) lds[x] = v[u];
        }
    }
};

template <int TI, int TJ, int TK>
struct MgTileMap {
    int i0, j0, k0;
    __device__ MgTileMap(int n) {
        const int ntj = (n + TJ - 1) / TJ, ntk = (n + TK - 1) / TK;
        const int t = xcd_tile(blockIdx.x, gridDim.x);
        const int tk = t % ntk, tj = (t / ntk) % ntj, ti = t / (ntk * ntj);
        i0 = ti * TI;
        j0 = tj * TJ;
        k0 = tk * TK;
    }
};

template <int TI, int TJ, int TK>
int mg_tiles(int n) {
    return ((n + TI - 1) / TI) * ((n + TJ - 1) / TJ) * ((n + TK - 1) / TK);
}

template <int TI, int TJ, int TK, int NT, class R = float>
__global__ __launch_bounds__(NT) void k_mg_down(int n, const R* __restrict__ b,
                                                    R* __restrict__ d2out,
                                                    R* __restrict__ bcoarse, R c1,
                                                    R c2, const CGState* st) {
    if (st->done) return;
    constexpr int BI = TI + 5, BJ = TJ + 5, BK = TK + 5;
    constexpr int DI = TI + 3, DJ = TJ + 3, DK = TK + 3;
    constexpr int RI = TI + 1, RJ = TJ + 1, RK = TK + 1;
    constexpr int CI = TI / 2, CJ = TJ / 2, CK = TK / 2;
    // All three arrays use b's row length BK, so that the threads of one wave
    // (consecutive positions along a row, rows in order) read and write
    // consecutive LDS words in every pass: d2 at (i, j, q) = D point (i, j,
    // q - 1), the residual at (i, j, q) = R point (i, j, q - 2); positions of a
    // row outside the region are skipped.
    __shared__ R sb[BI * BJ * BK];
    __shared__ R sd[DI * DJ * BK];
    __shared__ R sr[RI * RJ * BK];
    __shared__ int gi[BI], gj[BJ], gk[BK];
    const MgTileMap<TI, TJ, TK> T(n);
    const int tid = threadIdx.x;
    const int n2 = n * n;
    if (tid < BI) gi[tid] = wrap_mod(T.i0 - 3 + tid, n) * n2;
    if (tid >= 64 && tid < 64 + BJ) gj[tid - 64] = wrap_mod(T.j0 - 3 + tid - 64, n) * n;
    if (tid >= 128 && tid < 128 + BK) gk[tid - 128] = wrap_mod(T.k0 - 3 + tid - 128, n);
    lds_barrier();
    TileStage<BI * BJ * BK, NT, R> lb;
    lb.load([&](int x) {
        const int k = x % BK, j = (x / BK) % BJ, i = x / (BK * BJ);
        return b[gi[i] + gj[j] + gk[k]];
    });
    __asm__ volatile("" ::: "memory");  // every load is issued before the first store
    lb.store(sb);
    lds_barrier();
    // Sweeps from zero with weights c1 then c2: d2 = -(c1 + c2) b - c1 c2 A b.
    const R cs = c1 + c2, cp = c1 * c2;
    tile_for<DI * DJ * BK, NT>([&](int x) {
        const int q = x % BK, j = (x / BK) % DJ, i = x / (BK * DJ);
        if (q < 1 || q > DK) return;
        const int o = ((i + 1) * BJ + (j + 1)) * BK + q;
        const R bc = sb[o];
        const R lap = sb[o - BJ * BK] + sb[o + BJ * BK] + sb[o - BK] + sb[o + BK] + sb[o - 1] +
                          sb[o + 1] - R(6) * bc;
        sd[x] = -cs * bc - cp * lap;
    });
    lds_barrier();
    tile_for<RI * RJ * BK, NT>([&](int x) {
        const int q = x % BK, j = (x / BK) % RJ, i = x / (BK * RJ);
        if (q < 2 || q > RK + 1) return;
        const int o = ((i + 1) * DJ + (j + 1)) * BK + q;
        const R dc = sd[o];
        const R lap = sd[o - DJ * BK] + sd[o + DJ * BK] + sd[o - BK] + sd[o + BK] + sd[o - 1] +
                          sd[o + 1] - R(6) * dc;
        sr[x] = sb[((i + 2) * BJ + (j + 2)) * BK + q] - lap;
    });
    // Owned d2 (this loop and the residual pass only read sd).
    tile_for<TI * TJ * TK, NT>([&](int x) {
        const int k = x % TK, j = (x / TK) % TJ, i = x / (TK * TJ);
        if (T.i0 + i < n && T.j0 + j < n && T.k0 + k < n)
            d2out[(T.i0 + i) * n2 + (T.j0 + j) * n + T.k0 + k] =
                sd[((i + 2) * DJ + (j + 2)) * BK + (k + 3)];
    });
    lds_barrier();
    const int nc = n / 2;
    tile_for<CI * CJ * CK, NT>([&](int x) {
        const int K = x % CK, J = (x / CK) % CJ, I = x / (CK * CJ);
        const int gI = T.i0 / 2 + I, gJ = T.j0 / 2 + J, gK = T.k0 / 2 + K;
        if (gI >= nc || gJ >= nc || gK >= nc) return;
        const int o = ((2 * I + 1) * RJ + (2 * J + 1)) * BK + (2 * K + 3);
        R acc = R(0);
#pragma unroll
        for (int a = -1; a <= 1; a++)
#pragma unroll
            for (int bb = -1; bb <= 1; bb++)
#pragma unroll
                for (int cc = -1; cc <= 1; cc++) {
                    const R w = (R)((a == 0 ? 2 : 1) * (bb == 0 ? 2 : 1) * (cc == 0 ? 2 : 1));
                    acc += w * sr[o + (a * RJ + bb) * BK + cc];
                }
        bcoarse[(gI * nc + gJ) * nc + gK] = acc * (R(4) / R(64));
    });
}

template <int TI, int TJ, int TK, int NT, bool Dot, class R = float>
__global__ __launch_bounds__(NT) void k_mg_up(int n, const R* __restrict__ b,
                                                  const R* __restrict__ d2,
                                                  const R* __restrict__ e,
                                                  R* __restrict__ dout, R c1, R c2,
                                                  Reduce red,
                                                  CGState* st) {
    if (st->done) return;
    constexpr int DI = TI + 4, DJ = TJ + 4, DK = TK + 4;
    constexpr int PI = TI + 2, PJ = TJ + 2, PK = TK + 2;
    constexpr int EI = TI / 2 + 3, EJ = TJ / 2 + 3, EK = TK / 2 + 3;
    __shared__ R sd[DI * DJ * DK];
    __shared__ R sp[PI * PJ * PK];
    __shared__ R sb[PI * PJ * PK];
    __shared__ R se[EI * EJ * EK];
    __shared__ int gi[DI], gj[DJ], gk[DK], ci[EI], cj[EJ], ck[EK];
    const MgTileMap<TI, TJ, TK> T(n);
    const int tid = threadIdx.x;
    const int n2 = n * n, nc = n / 2;
    if (tid < DI) gi[tid] = wrap_mod(T.i0 - 2 + tid, n) * n2;
    if (tid >= 32 && tid < 32 + DJ) gj[tid - 32] = wrap_mod(T.j0 - 2 + tid - 32, n) * n;
    if (tid >= 64 && tid < 64 + DK) gk[tid - 64] = wrap_mod(T.k0 - 2 + tid - 64, n);
    if (tid >= 128 && tid < 128 + EI) ci[tid - 128] = wrap_mod(T.i0 / 2 - 1 + tid - 128, nc) * nc * nc;
    if (tid >= 160 && tid < 160 + EJ) cj[tid - 160] = wrap_mod(T.j0 / 2 - 1 + tid - 160, nc) * nc;
    if (tid >= 192 && tid < 192 + EK) ck[tid - 192] = wrap_mod(T.k0 / 2 - 1 + tid - 192, nc);
    lds_barrier();
    TileStage<DI * DJ * DK, NT, R> ld;
    TileStage<PI * PJ * PK, NT, R> lb;
    TileStage<EI * EJ * EK, NT, R> le;
    ld.load([&](int x) {
        const int k = x % DK, j = (x / DK) % DJ, i = x / (DK * DJ);
        return d2[gi[i] + gj[j] + gk[k]];
    });
    lb.load([&](int x) {
        const int k = x % PK, j = (x / PK) % PJ, i = x / (PK * PJ);
        return b[gi[i + 1] + gj[j + 1] + gk[k + 1]];
    });
    le.load([&](int x) {
        const int k = x % EK, j = (x / EK) % EJ, i = x / (EK * EJ);
        return e[ci[i] + cj[j] + ck[k]];
    });
    __asm__ volatile("" ::: "memory");  // every load is issued before the first store
    ld.store(sd);
    lb.store(sb);
    le.store(se);
    lds_barrier();
    // d = d2 + trilinear(e): fine local index l has coarse neighbours l>>1 and
    // (l>>1)+1 and weight 1/2 on odd l (tile origins are even).
    tile_for<DI * DJ * DK, NT>([&](int x) {
        const int k = x % DK, j = (x / DK) % DJ, i = x / (DK * DJ);
        const int a = i >> 1, bb = j >> 1, cc = k >> 1;
        const R wi = (i & 1) ? R(0.5) : R(0), wj = (j & 1) ? R(0.5) : R(0),
                    wk = (k & 1) ? R(0.5) : R(0);
        const int o = (a * EJ + bb) * EK + cc;
        const R e000 = se[o], e001 = se[o + 1], e010 = se[o + EK], e011 = se[o + EK + 1];
        const R e100 = se[o + EJ * EK], e101 = se[o + EJ * EK + 1];
        const R e110 = se[o + EJ * EK + EK], e111 = se[o + EJ * EK + EK + 1];
        const R v00 = e000 + wk * (e001 - e000);
        const R v01 = e010 + wk * (e011 - e010);
        const R v10 = e100 + wk * (e101 - e100);
        const R v11 = e110 + wk * (e111 - e110);
        const R v0 = v00 + wj * (v01 - v00);
        const R v1 = v10 + wj * (v11 - v10);
        sd[x] = sd[x] + (v0 + wi * (v1 - v0));
    });
    lds_barrier();
    tile_for<PI * PJ * PK, NT>([&](int x) {
        const int k = x % PK, j = (x / PK) % PJ, i = x / (PK * PJ);
        const int o = ((i + 1) * DJ + (j + 1)) * DK + (k + 1);
        const R dc = sd[o];
        const R lap = sd[o - DJ * DK] + sd[o + DJ * DK] + sd[o - DK] + sd[o + DK] + sd[o - 1] +
                          sd[o + 1] - R(6) * dc;
        sp[x] = dc - c2 * (sb[x] - lap);
    });
    lds_barrier();
    double s = 0.0;
    tile_for<TI * TJ * TK, NT>([&](int x) {
        const int k = x % TK, j = (x / TK) % TJ, i = x / (TK * TJ);
        const int o = ((i + 1) * PJ + (j + 1)) * PK + (k + 1);
        const R pc = sp[o];
        const R lap = sp[o - PJ * PK] + sp[o + PJ * PK] + sp[o - PK] + sp[o + PK] + sp[o - 1] +
                          sp[o + 1] - R(6) * pc;
        const R bc = sb[o];
        const R z = pc - c1 * (bc - lap);
        if (T.i0 + i < n && T.j0 + j < n && T.k0 + k < n) {
            dout[(T.i0 + i) * n2 + (T.j0 + j) * n + T.k0 + k] = z;
            if (Dot) s += (double)bc * (double)z;
        }
    });
    if constexpr (Dot) {  // <b, d> partial for the direction pass (deferred finalisation)
        __shared__ double sh[2 * NT / kWave];
        double m = 0.0;
        block_sum_max_lds<NT>(s, m, sh);
        if (tid == 0) red.psum[blockIdx.x] = s;
    }
}

// ---- the same fused level kernels with vector loads, column-mapped -------------
// For every even n >= 32.  Every row of a tile's region is read as the V-float
// groups of the aligned window [k0 - 4, k0 + TK + 4) (fine levels) or the
// VE-float groups of [k0/2 - 4, k0/2 + TK/2 + 4) (the coarse correction).  V
// and VE are the largest of 4, 2 and 1 that divide n and n / 2: group starts
// are multiples of the width and so is the row length, so a wrapped group
// stays contiguous and aligned (n % 8 == 0: 16-byte loads throughout; n = 82,
// 102, 130: 8-byte fine loads and scalar coarse loads).  LDS rows are
// RW = TK + 8 floats (a few unused columns).
// A thread owns one (j, k) column of a region and walks its planes i, so the
// region indices are formed once per thread instead of by division per
// element, the i-1 / i / i+1 values of a sweep stay in registers (5 LDS reads
// per point instead of 7), and the prolongation forms each coarse plane's
// bilinear value once per column.  The arithmetic per point is the scalar
// kernels', value for value and in the same order, so the level outputs are
// bit-identical to them (only the order in which the dot partial's terms are
// added differs).  At 128^3, 8x16x16 tiles: up 20.1 -> 15.0 us (512
// threads), down 14.2 -> 11.8 us (1024 threads).
__device__ __forceinline__ int wrap1(int v, int m) { return v < 0 ? v + m : (v >= m ? v - m : v); }

// One region of PI_ planes x PJ_ rows x RV_ groups of V floats, loaded in two
// halves: load() issues this thread's loads into registers, store() writes
// them to LDS.  Thread t takes group e = t % (PJ_ RV_) of planes h, h + G, ...
// with h = t / (PJ_ RV_).  A kernel issues the loads of all its regions before
// it stores any, so their latencies overlap; with a load and a store per
// region, each region's loads waited for the previous region's stores (one
// memory round trip per region).
template <int V>
struct VecOf {
    typedef float T __attribute__((ext_vector_type(V)));
};
template <>
struct VecOf<1> { using T = float; };

template <int PI_, int PJ_, int RV_, int NT, int V>
struct RegionLoad {
    static constexpr int PL = PJ_ * RV_;
    static_assert(PL <= NT, "every group of a plane row set has a thread");
    static constexpr int G = NT / PL, IT = (PI_ + G - 1) / G;
    using T = typename VecOf<V>::T;
    T v[IT];
    // f(i, j, g) -> global address of group g of row j of plane i.  Every
    // thread loads IT groups without a branch (threads past the region load
    // plane 0 again, which store() skips), so the values stay in registers.
    template <class F>
    __device__ __forceinline__ void load(F&& f) {
        const int t = threadIdx.x, h = t / PL, e = t % PL, g = e % RV_, j = e / RV_;
#pragma unroll
        for (int u = 0; u < IT; u++) {
            const int i = h + u * G;
            v[u] = *reinterpret_cast<const T*>(f(t < G * PL && i < PI_ ? i : 0, j, g));
        }
    }
    // Group g of row j of plane i goes to lds[(i PJ_ + j) RW + V g]; the extra
    // loads go to the 4 floats past the region (every store is made, so the
    // loads and stores form one branch-free block the compiler can order with
    // all loads first).  The LDS array holds PI_ PJ_ RW + 4 floats.
    template <int RW>
    __device__ __forceinline__ void store(float* lds) const {
        const int t = threadIdx.x, h = t / PL, e = t % PL, g = e % RV_, j = e / RV_;
#pragma unroll
        for (int u = 0; u < IT; u++) {
            const int i = h + u * G;
            const int o = t < G * PL && i < PI_ ? (i * PJ_ + j) * RW + V * g : PI_ * PJ_ * RW;
            *reinterpret_cast<T*>(&lds[o]) = v[u];
        }
    }
};

template <int TI, int TJ, int TK, int NT, bool Dot, int V, int VE>
__global__ __launch_bounds__(NT) void k_mg_up_vec(int n, const float* __restrict__ b,
                                                  const float* __restrict__ d2,
                                                  const float* __restrict__ e,
                                                  float* __restrict__ dout, float c1, float c2,
                                                  Reduce red, CGState* st) {
    // The stop flag is tested once the operand loads are in flight, so its
    // latency does not precede them; nothing writes it during this kernel, so
    // every wave of a work-group takes the same branch.
    const int done = st->done;
    constexpr int DI = TI + 4, DJ = TJ + 4, DK = TK + 4;
    constexpr int PI = TI + 2, PJ = TJ + 2, PK = TK + 2;
    constexpr int EI = TI / 2 + 3, EJ = TJ / 2 + 3;
    constexpr int RW = TK + 8, RV = RW / V;
    constexpr int RWE = TK / 2 + 8, RVE = RWE / VE;
    static_assert(DJ * DK <= NT && PJ * PK <= NT, "one thread per column");
    __shared__ __attribute__((aligned(16))) float sd[DI * DJ * RW + 4];
    __shared__ __attribute__((aligned(16))) float sb[PI * PJ * RW + 4];
    __shared__ __attribute__((aligned(16))) float se[EI * EJ * RWE + 4];
    __shared__ float sp[PI * PJ * PK];
    const MgTileMap<TI, TJ, TK> T(n);
    const int tid = threadIdx.x;
    const int n2 = n * n, nc = n / 2, nc2 = nc * nc;
    RegionLoad<DI, DJ, RV, NT, V> ld;
    RegionLoad<PI, PJ, RV, NT, V> lb;
    RegionLoad<EI, EJ, RVE, NT, VE> le;
    ld.load([&](int i, int j, int g) {
        const int gi = wrap1(T.i0 - 2 + i, n), gj = wrap1(T.j0 - 2 + j, n);
        return &d2[(long)gi * n2 + gj * n + wrap1(T.k0 - 4 + V * g, n)];
    });
    lb.load([&](int i, int j, int g) {
        const int gi = wrap1(T.i0 - 1 + i, n), gj = wrap1(T.j0 - 1 + j, n);
        return &b[(long)gi * n2 + gj * n + wrap1(T.k0 - 4 + V * g, n)];
    });
    le.load([&](int i, int j, int g) {
        const int gi = wrap1(T.i0 / 2 - 1 + i, nc), gj = wrap1(T.j0 / 2 - 1 + j, nc);
        return &e[(long)gi * nc2 + gj * nc + wrap1(T.k0 / 2 - 4 + VE * g, nc)];
    });
    __asm__ volatile("" ::: "memory");  // every load is issued before the first store
    ld.template store<RW>(sd);
    lb.template store<RW>(sb);
    le.template store<RWE>(se);
    if (done) return;
    lds_barrier();
    // d = d2 + trilinear(e) on the D region: column (j, k), coarse row bb =
    // j >> 1, coarse column (k >> 1) + 3; the bilinear value of every coarse
    // plane a once, then fine plane i takes planes i >> 1 and (i >> 1) + 1.
    if (tid < DJ * DK) {
        const int j = tid / DK, k = tid % DK;
        const float wj = (j & 1) ? 0.5f : 0.0f, wk = (k & 1) ? 0.5f : 0.0f;
        const int ob = (j >> 1) * RWE + (k >> 1) + 3;
        float B[EI];
#pragma unroll
        for (int a = 0; a < EI; a++) {
            const int o = a * EJ * RWE + ob;
            const float e00 = se[o], e01 = se[o + 1], e10 = se[o + RWE], e11 = se[o + RWE + 1];
            const float v0 = e00 + wk * (e01 - e00);
            const float v1 = e10 + wk * (e11 - e10);
            B[a] = v0 + wj * (v1 - v0);
        }
#pragma unroll
        for (int i = 0; i < DI; i++) {
            const float wi = (i & 1) ? 0.5f : 0.0f;
            const float v0 = B[i >> 1], v1 = B[(i >> 1) + 1];
            const int q = (i * DJ + j) * RW + k + 2;
            sd[q] = sd[q] + (v0 + wi * (v1 - v0));
        }
    }
    lds_barrier();
    // First sweep on the P region, walking i with the centre column in registers.
    if (tid < PJ * PK) {
        const int j = tid / PK, k = tid % PK;
        const int oc = (j + 1) * RW + (k + 3);  // D-plane offset of this column
        float dm = sd[oc], dc = sd[DJ * RW + oc];
#pragma unroll
        for (int i = 0; i < PI; i++) {
            const int o = (i + 1) * DJ * RW + oc;
            const float dp = sd[o + DJ * RW];
            const float lap = dm + dp + sd[o - RW] + sd[o + RW] + sd[o - 1] + sd[o + 1] - 6.0f * dc;
            sp[(i * PJ + j) * PK + k] = dc - c2 * (sb[(i * PJ + j) * RW + (k + 3)] - lap);
            dm = dc;
            dc = dp;
        }
    }
    lds_barrier();
    // Second sweep on the tile: G groups of TI / G planes per column.
    constexpr int CC = TJ * TK;
    constexpr int G = NT / CC < TI ? NT / CC : TI;
    static_assert(G >= 1 && TI % G == 0, "whole plane groups");
    constexpr int PG = TI / G;
    double s = 0.0;
    if (tid < G * CC) {
        const int h = tid / CC, c = tid % CC;
        const int j = c / TK, k = c % TK;
        const int i0 = h * PG;
        const int oc = (j + 1) * PK + (k + 1);
        float pm = sp[i0 * PJ * PK + oc], pc = sp[(i0 + 1) * PJ * PK + oc];
#pragma unroll
        for (int u = 0; u < PG; u++) {
            const int i = i0 + u;
            const int o = (i + 1) * PJ * PK + oc;
            const float pn = sp[o + PJ * PK];
            const float lap = pm + pn + sp[o - PK] + sp[o + PK] + sp[o - 1] + sp[o + 1] - 6.0f * pc;
            const float bc = sb[((i + 1) * PJ + (j + 1)) * RW + (k + 4)];
            const float z = pc - c1 * (bc - lap);
            if (T.i0 + i < n && T.j0 + j < n && T.k0 + k < n) {
                dout[(T.i0 + i) * n2 + (T.j0 + j) * n + T.k0 + k] = z;
                if (Dot) s += (double)bc * (double)z;
            }
            pm = pc;
            pc = pn;
        }
    }
    if constexpr (Dot) {
        __shared__ double sh[2 * NT / kWave];
        double m = 0.0;
        block_sum_max_lds<NT>(s, m, sh);
        if (tid == 0) red.psum[blockIdx.x] = s;
    }
}

// The down kernel's work; false if the solve has stopped (nothing written).
// Shared: the coarse right-hand side is stored at device scope, for a last
// work-group that reads it in the same kernel (k_mg_down_coarse).
template <int TI, int TJ, int TK, int NT, int V, bool Shared>
__device__ __forceinline__ bool mg_down_vec_body(int n, const float* __restrict__ b,
                                                 float* __restrict__ d2out,
                                                 float* __restrict__ bcoarse, float c1, float c2,
                                                 const CGState* st) {
    const int done = st->done;  // tested once the loads are in flight (as k_mg_up_vec)
    constexpr int BI = TI + 5, BJ = TJ + 5;
    constexpr int DI = TI + 3, DJ = TJ + 3, DK = TK + 3;
    constexpr int RI = TI + 1, RJ = TJ + 1, RK = TK + 1;
    constexpr int CI = TI / 2, CJ = TJ / 2, CK = TK / 2;
    constexpr int RW = TK + 8, RV = RW / V;
    static_assert(DJ * DK <= NT && RJ * RK <= NT, "one thread per column");
    __shared__ __attribute__((aligned(16))) float sb[BI * BJ * RW + 4];
    __shared__ float sd[DI * DJ * RW];
    __shared__ float sr[RI * RJ * RW];
    const MgTileMap<TI, TJ, TK> T(n);
    const int tid = threadIdx.x;
    const int n2 = n * n;
    RegionLoad<BI, BJ, RV, NT, V> lb;
    lb.load([&](int i, int j, int g) {
        const int gi = wrap1(T.i0 - 3 + i, n), gj = wrap1(T.j0 - 3 + j, n);
        return &b[(long)gi * n2 + gj * n + wrap1(T.k0 - 4 + V * g, n)];
    });
    __asm__ volatile("" ::: "memory");  // every load is issued before the first store
    lb.template store<RW>(sb);
    if (done) return false;
    lds_barrier();
    const float cs = c1 + c2, cp = c1 * c2;
    // d2 on the D region (LDS column k + 2), walking i.
    if (tid < DJ * DK) {
        const int j = tid / DK, k = tid % DK;
        const int oc = (j + 1) * RW + k + 2;  // B-plane offset of this column
        float bm = sb[oc], bc = sb[BJ * RW + oc];
#pragma unroll
        for (int i = 0; i < DI; i++) {
            const int o = (i + 1) * BJ * RW + oc;
            const float bp = sb[o + BJ * RW];
            const float lap = bm + bp + sb[o - RW] + sb[o + RW] + sb[o - 1] + sb[o + 1] - 6.0f * bc;
            sd[(i * DJ + j) * RW + k + 2] = -cs * bc - cp * lap;
            bm = bc;
            bc = bp;
        }
    }
    lds_barrier();
    // Residual on the R region (LDS column k + 3).
    if (tid < RJ * RK) {
        const int j = tid / RK, k = tid % RK;
        const int oc = (j + 1) * RW + k + 3;  // D-plane offset
        float dm = sd[oc], dc = sd[DJ * RW + oc];
#pragma unroll
        for (int i = 0; i < RI; i++) {
            const int o = (i + 1) * DJ * RW + oc;
            const float dp = sd[o + DJ * RW];
            const float lap = dm + dp + sd[o - RW] + sd[o + RW] + sd[o - 1] + sd[o + 1] - 6.0f * dc;
            sr[(i * RJ + j) * RW + k + 3] = sb[((i + 2) * BJ + (j + 2)) * RW + k + 3] - lap;
            dm = dc;
            dc = dp;
        }
    }
    // Owned d2, G groups of plane columns.
    {
        constexpr int CC = TJ * TK;
        constexpr int G = NT / CC < TI ? NT / CC : TI;
        static_assert(G >= 1 && TI % G == 0, "whole plane groups");
        constexpr int PG = TI / G;
        if (tid < G * CC) {
            const int h = tid / CC, c = tid % CC;
            const int j = c / TK, k = c % TK;
#pragma unroll
            for (int u = 0; u < PG; u++) {
                const int i = h * PG + u;
                if (T.i0 + i < n && T.j0 + j < n && T.k0 + k < n)
                    d2out[(T.i0 + i) * n2 + (T.j0 + j) * n + T.k0 + k] =
                        sd[((i + 2) * DJ + (j + 2)) * RW + k + 4];
            }
        }
    }
    lds_barrier();
    const int nc = n / 2;
    tile_for<CI * CJ * CK, NT>([&](int x) {
        const int K = x % CK, J = (x / CK) % CJ, I = x / (CK * CJ);
        const int gI = T.i0 / 2 + I, gJ = T.j0 / 2 + J, gK = T.k0 / 2 + K;
        if (gI >= nc || gJ >= nc || gK >= nc) return;
        const int o = ((2 * I + 1) * RJ + (2 * J + 1)) * RW + (2 * K + 4);
        float acc = 0.0f;
#pragma unroll
        for (int a = -1; a <= 1; a++)
#pragma unroll
            for (int bb = -1; bb <= 1; bb++)
#pragma unroll
                for (int cc = -1; cc <= 1; cc++) {
                    const float w = (float)((a == 0 ? 2 : 1) * (bb == 0 ? 2 : 1) * (cc == 0 ? 2 : 1));
                    acc += w * sr[o + (a * RJ + bb) * RW + cc];
                }
        const float v = acc * (4.0f / 64.0f);
        if constexpr (Shared)
            __hip_atomic_store(bcoarse + (gI * nc + gJ) * nc + gK, v, __ATOMIC_RELAXED,
                               __HIP_MEMORY_SCOPE_AGENT);
        else
            bcoarse[(gI * nc + gJ) * nc + gK] = v;
    });
    return true;
}

template <int TI, int TJ, int TK, int NT, int V>
__global__ __launch_bounds__(NT) void k_mg_down_vec(int n, const float* __restrict__ b,
                                                    float* __restrict__ d2out,
                                                    float* __restrict__ bcoarse, float c1,
                                                    float c2, const CGState* st) {
    mg_down_vec_body<TI, TJ, TK, NT, V, false>(n, b, d2out, bcoarse, c1, c2, st);
}

// Load widths of the column-mapped kernels: 0 where they do not apply (n odd
// or below 32: the halo windows then reach past one period), else the fine
// width V (4 or 2) times 8 plus the coarse width VE (4, 2 or 1).  scalar: the
// scalar-load tile kernels instead (tile ids or-ed with kMgScalarLoads).
inline int mg_vec_widths(int n, bool scalar) {
    static const bool all_scalar = [] {  // MAZE_MG_SCALAR=1: tuning override
        const char* e = std::getenv("MAZE_MG_SCALAR");
        return e && std::atoi(e) != 0;
    }();
    if (scalar || all_scalar || n < 32 || n % 2 != 0) return 0;
    const int nc = n / 2;
    const int v = n % 4 == 0 ? 4 : 2, ve = nc % 4 == 0 ? 4 : (nc % 2 == 0 ? 2 : 1);
    return 8 * v + ve;
}

// Tile shapes (fine points per work-group, threads) of the fused level kernels;
// Sim takes the defaults per level (mg_default_tile, mg_default_tile_up) unless
// MAZE_MG_TILES names one.
template <int TI, int TJ, int TK, int NT>
static void down_t(hipStream_t s, int n, const float* b, float* d2, float* bc, double c1,
                   double c2, CGState* st, bool scalar) {
    const dim3 g(mg_tiles<TI, TJ, TK>(n));
    const float f1 = (float)c1, f2 = (float)c2;
    switch (mg_vec_widths(n, scalar) / 8) {
        case 4: MAZE_LAUNCH((k_mg_down_vec<TI, TJ, TK, NT, 4>), g, dim3(NT), 0, s, n, b, d2, bc, f1, f2, st); break;
        case 2: MAZE_LAUNCH((k_mg_down_vec<TI, TJ, TK, NT, 2>), g, dim3(NT), 0, s, n, b, d2, bc, f1, f2, st); break;
        default: MAZE_LAUNCH((k_mg_down<TI, TJ, TK, NT>), g, dim3(NT), 0, s, n, b, d2, bc, f1, f2, st);
    }
}

template <int TI, int TJ, int TK, int NT, bool Dot>
static void up_dt(hipStream_t s, int n, const float* b, const float* d2, const float* e, float* d,
                  double c1, double c2, CGState* st, Reduce red, bool scalar) {
    const dim3 g(mg_tiles<TI, TJ, TK>(n));
    const float f1 = (float)c1, f2 = (float)c2;
    switch (mg_vec_widths(n, scalar)) {
        case 8 * 4 + 4:
            MAZE_LAUNCH((k_mg_up_vec<TI, TJ, TK, NT, Dot, 4, 4>), g, dim3(NT), 0, s, n, b, d2, e, d, f1, f2, red, st);
            break;
        case 8 * 4 + 2:
            MAZE_LAUNCH((k_mg_up_vec<TI, TJ, TK, NT, Dot, 4, 2>), g, dim3(NT), 0, s, n, b, d2, e, d, f1, f2, red, st);
            break;
        case 8 * 2 + 1:
            MAZE_LAUNCH((k_mg_up_vec<TI, TJ, TK, NT, Dot, 2, 1>), g, dim3(NT), 0, s, n, b, d2, e, d, f1, f2, red, st);
            break;
        default:
            MAZE_LAUNCH((k_mg_up<TI, TJ, TK, NT, Dot>), g, dim3(NT), 0, s, n, b, d2, e, d, f1, f2, red, st);
    }
}

template <int TI, int TJ, int TK, int NT>
static void up_t(hipStream_t s, int n, const float* b, const float* d2, const float* e, float* d,
                 double c1, double c2, CGState* st, int dot, Reduce red, bool scalar) {
    if (dot)
        up_dt<TI, TJ, TK, NT, true>(s, n, b, d2, e, d, c1, c2, st, red, scalar);
    else
        up_dt<TI, TJ, TK, NT, false>(s, n, b, d2, e, d, c1, c2, st, red, scalar);
}

// ---- plane-marching level kernels ---------------------------------------------
// The same down / up steps as the tiles above, with the same arithmetic in the
// same order (so the outputs are bit-identical), organised as a march along i:
// a 256-thread work-group owns a 16x16 column block of the (j, k) plane over S
// planes and keeps the planes the stencils need in LDS rings of three, each
// plane with its (j, k) halo.  Each plane's global loads are issued one plane
// ahead into registers, so loads stay in flight during the plane's LDS steps,
// and the halo is only in (j, k): 1.6x the points of the column block instead
// of 2.3x for an 8x16x16 tile.  Needs n a multiple of 16 and of S.
constexpr int kMarchC = 16;

__device__ __forceinline__ int ring3(int q) { return ((q % 3) + 3) % 3; }

template <int S>
__global__ __launch_bounds__(kBlock) void k_mg_down_march(int n, const float* __restrict__ b,
                                                          float* __restrict__ d2out,
                                                          float* __restrict__ bcoarse, float c1,
                                                          float c2, const CGState* st) {
    if (st->done) return;
    constexpr int WB = kMarchC + 6, WD = kMarchC + 4, WR = kMarchC + 2;
    constexpr int PB = WB * WB, PD = WD * WD, PR = WR * WR;
    static_assert(PB <= 2 * kBlock, "two B-plane points per thread");
    __shared__ float sb[3][PB], sd[3][PD], sr[3][PR];
    const int tid = threadIdx.x;
    const int nblk = n / kMarchC, n2 = n * n, nc = n / 2;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int cb = blk % (nblk * nblk), i0 = (blk / (nblk * nblk)) * S;
    const int j0 = (cb / nblk) * kMarchC, k0 = (cb % nblk) * kMarchC;
    int boff[2];
    bool bok[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int r = tid + u * kBlock;
        bok[u] = r < PB;
        const int jj = r / WB, kk = r % WB;
        boff[u] = bok[u] ? wrap_mod(j0 - 3 + jj, n) * n + wrap_mod(k0 - 3 + kk, n) : 0;
    }
    // Global loads run three planes ahead of the plane being stored: register
    // sets ra, rb, rc rotate (loop unrolled by three), so each plane's loads
    // have three planes' worth of LDS work to arrive.
    auto load_b = [&](int q, float (&v)[2]) {
        const int pl = wrap_mod(q, n) * n2;
#pragma unroll
        for (int u = 0; u < 2; u++) v[u] = bok[u] ? b[pl + boff[u]] : 0.0f;
    };
    const float cs = c1 + c2, cp = c1 * c2;
    constexpr int U = S + 5;  // B planes i0-3 .. i0+S+1
    // B planes i0-3 .. i0+S+1, d2 planes i0-2 .. i0+S, residual planes i0-1 .. i0+S-1.
    auto step = [&](int u, float (&nb)[2]) {
        if (u >= U) return;
        const int qb = i0 - 3 + u;
        float* B = sb[ring3(qb)];
#pragma unroll
        for (int v = 0; v < 2; v++)
            if (bok[v]) B[tid + v * kBlock] = nb[v];
        if (u + 3 < U) load_b(qb + 3, nb);
        lds_barrier();
        if (u >= 2) {
            const int qd = qb - 1;
            const float* Bm = sb[ring3(qd - 1)];
            const float* B0 = sb[ring3(qd)];
            const float* Bp = sb[ring3(qd + 1)];
            float* D = sd[ring3(qd)];
            const bool own = qd >= i0 && qd < i0 + S;
#pragma unroll 2
            for (int r = tid; r < PD; r += kBlock) {
                const int jj = r / WD, kk = r % WD;
                const int o = (jj + 1) * WB + (kk + 1);
                const float bc = B0[o];
                const float lap = Bm[o] + Bp[o] + B0[o - WB] + B0[o + WB] + B0[o - 1] + B0[o + 1] -
                                  6.0f * bc;
                const float v = -cs * bc - cp * lap;
                D[r] = v;
                if (own && jj >= 2 && jj < WD - 2 && kk >= 2 && kk < WD - 2)
                    d2out[qd * n2 + (j0 + jj - 2) * n + k0 + kk - 2] = v;
            }
        }
        lds_barrier();
        if (u >= 4) {
            const int qr = qb - 2;
            const float* Dm = sd[ring3(qr - 1)];
            const float* D0 = sd[ring3(qr)];
            const float* Dp = sd[ring3(qr + 1)];
            const float* B0 = sb[ring3(qr)];
            float* R = sr[ring3(qr)];
#pragma unroll 2
            for (int r = tid; r < PR; r += kBlock) {
                const int jj = r / WR, kk = r % WR;
                const int o = (jj + 1) * WD + (kk + 1);
                const float dc = D0[o];
                const float lap = Dm[o] + Dp[o] + D0[o - WD] + D0[o + WD] + D0[o - 1] + D0[o + 1] -
                                  6.0f * dc;
                R[r] = B0[(jj + 2) * WB + (kk + 2)] - lap;
            }
            lds_barrier();
            // Coarse plane Q of this slab from residual planes 2Q-1, 2Q, 2Q+1 =
            // qr-2, qr-1, qr (qr odd relative to the even i0, and qr > i0 so that
            // all three planes were computed here).
            if (qr > i0 && ((qr - i0) & 1) && tid < (kMarchC / 2) * (kMarchC / 2)) {
                const int Q = (qr - 1) / 2;
                const int J = tid / (kMarchC / 2), K = tid % (kMarchC / 2);
                const int o = (2 * J + 1) * WR + (2 * K + 1);
                const float* Rp[3] = {sr[ring3(qr - 2)], sr[ring3(qr - 1)], sr[ring3(qr)]};
                float acc = 0.0f;
#pragma unroll
                for (int a = -1; a <= 1; a++)
#pragma unroll
                    for (int bb = -1; bb <= 1; bb++)
#pragma unroll
                        for (int cc = -1; cc <= 1; cc++) {
                            const float w =
                                (float)((a == 0 ? 2 : 1) * (bb == 0 ? 2 : 1) * (cc == 0 ? 2 : 1));
                            acc += w * Rp[a + 1][o + bb * WR + cc];
                        }
                bcoarse[(Q * nc + j0 / 2 + J) * nc + k0 / 2 + K] = acc * (4.0f / 64.0f);
            }
        }
    };
    float ra[2], rb[2], rc[2];
    load_b(i0 - 3, ra);
    load_b(i0 - 2, rb);
    load_b(i0 - 1, rc);
    for (int u = 0; u < U; u += 3) {
        step(u, ra);
        step(u + 1, rb);
        step(u + 2, rc);
    }
}

template <int S, bool Dot>
__global__ __launch_bounds__(kBlock) void k_mg_up_march(int n, const float* __restrict__ b,
                                                        const float* __restrict__ d2,
                                                        const float* __restrict__ e,
                                                        float* __restrict__ dout, float c1,
                                                        float c2, Reduce red, CGState* st) {
    if (st->done) return;
    constexpr int WD = kMarchC + 4, WT = kMarchC + 2, WE = kMarchC / 2 + 3;
    constexpr int PD = WD * WD, PT = WT * WT, PE = WE * WE;
    static_assert(PD <= 2 * kBlock && PE <= kBlock, "plane points per thread");
    __shared__ float sd[3][PD], st_[3][PT], sbp[3][PT], se[3][PE];
    const int tid = threadIdx.x;
    const int nblk = n / kMarchC, n2 = n * n, nc = n / 2;
    const int blk = xcd_tile(blockIdx.x, gridDim.x);
    const int cb = blk % (nblk * nblk), i0 = (blk / (nblk * nblk)) * S;
    const int j0 = (cb / nblk) * kMarchC, k0 = (cb % nblk) * kMarchC;
    int doff[2], toff[2];
    bool dok[2], tok[2];
#pragma unroll
    for (int u = 0; u < 2; u++) {
        const int r = tid + u * kBlock;
        dok[u] = r < PD;
        doff[u] = dok[u] ? wrap_mod(j0 - 2 + r / WD, n) * n + wrap_mod(k0 - 2 + r % WD, n) : 0;
        tok[u] = r < PT;
        toff[u] = tok[u] ? wrap_mod(j0 - 1 + r / WT, n) * n + wrap_mod(k0 - 1 + r % WT, n) : 0;
    }
    const bool eok = tid < PE;
    const int eoff = eok ? wrap_mod((j0 - 2) / 2 + tid / WE, nc) * nc +
                               wrap_mod((k0 - 2) / 2 + tid % WE, nc)
                         : 0;
    // Registers of a plane three planes ahead (sets ra, rb, rc rotate, loop
    // unrolled by three): d2 plane q, b plane q - 1 and, when q is even, coarse
    // plane q / 2 + 1.
    struct Regs {
        float d[2], b[2], e;
    };
    auto load = [&](int q, Regs& g) {
        const int pd = wrap_mod(q, n) * n2, pb = wrap_mod(q - 1, n) * n2;
#pragma unroll
        for (int u = 0; u < 2; u++) {
            g.d[u] = dok[u] ? d2[pd + doff[u]] : 0.0f;
            g.b[u] = tok[u] ? b[pb + toff[u]] : 0.0f;
        }
        g.e = ((q & 1) == 0 && eok) ? e[wrap_mod((q >> 1) + 1, nc) * nc * nc + eoff] : 0.0f;
    };
    // The first D plane (even) also needs coarse plane q >> 1.
    if (eok) se[ring3(((i0 - 2) >> 1))][tid] = e[wrap_mod((i0 - 2) >> 1, nc) * nc * nc + eoff];
    double s = 0.0;
    constexpr int U = S + 4;
    // D planes i0-2 .. i0+S+1, T planes i0-1 .. i0+S, Z planes i0 .. i0+S-1.
    auto step = [&](int u, Regs& g) {
        if (u >= U) return;
        const int q = i0 - 2 + u;
        const float dreg[2] = {g.d[0], g.d[1]};
        float* Bq = sbp[ring3(q - 1)];
#pragma unroll
        for (int v = 0; v < 2; v++)
            if (tok[v]) Bq[tid + v * kBlock] = g.b[v];
        if ((q & 1) == 0 && eok) se[ring3((q >> 1) + 1)][tid] = g.e;
        if (u + 3 < U) load(q + 3, g);
        lds_barrier();
        {  // D plane q = d2 + trilinear(e)
            const int a = q >> 1;
            const float wi = (q & 1) ? 0.5f : 0.0f;
            const float* e0 = se[ring3(a)];
            const float* e1 = se[ring3(a + 1)];
            float* D = sd[ring3(q)];
#pragma unroll
            for (int v = 0; v < 2; v++) {
                if (!dok[v]) continue;
                const int r = tid + v * kBlock;
                const int j = r / WD, k = r % WD;
                const int bb = j >> 1, cc = k >> 1;
                const float wj = (j & 1) ? 0.5f : 0.0f, wk = (k & 1) ? 0.5f : 0.0f;
                const int o = bb * WE + cc;
                const float e000 = e0[o], e001 = e0[o + 1], e010 = e0[o + WE], e011 = e0[o + WE + 1];
                const float e100 = e1[o], e101 = e1[o + 1];
                const float e110 = e1[o + WE], e111 = e1[o + WE + 1];
                const float v00 = e000 + wk * (e001 - e000);
                const float v01 = e010 + wk * (e011 - e010);
                const float v10 = e100 + wk * (e101 - e100);
                const float v11 = e110 + wk * (e111 - e110);
                const float v0 = v00 + wj * (v01 - v00);
                const float v1 = v10 + wj * (v11 - v10);
                D[r] = dreg[v] + (v0 + wi * (v1 - v0));
            }
        }
        lds_barrier();
        if (u >= 2) {  // T plane q - 1: first sweep (weight c2)
            const int qt = q - 1;
            const float* Dm = sd[ring3(qt - 1)];
            const float* D0 = sd[ring3(qt)];
            const float* Dp = sd[ring3(qt + 1)];
            const float* B0 = sbp[ring3(qt)];
            float* T = st_[ring3(qt)];
#pragma unroll 2
            for (int r = tid; r < PT; r += kBlock) {
                const int jj = r / WT, kk = r % WT;
                const int o = (jj + 1) * WD + (kk + 1);
                const float dc = D0[o];
                const float lap = Dm[o] + Dp[o] + D0[o - WD] + D0[o + WD] + D0[o - 1] + D0[o + 1] -
                                  6.0f * dc;
                T[r] = dc - c2 * (B0[r] - lap);
            }
        }
        lds_barrier();
        if (u >= 4) {  // Z plane q - 2: second sweep (weight c1), written out
            const int qz = q - 2;
            const float* Tm = st_[ring3(qz - 1)];
            const float* T0 = st_[ring3(qz)];
            const float* Tp = st_[ring3(qz + 1)];
            const float* B0 = sbp[ring3(qz)];
            const int jj = tid / kMarchC, kk = tid % kMarchC;
            const int o = (jj + 1) * WT + (kk + 1);
            const float pc = T0[o];
            const float lap = Tm[o] + Tp[o] + T0[o - WT] + T0[o + WT] + T0[o - 1] + T0[o + 1] -
                              6.0f * pc;
            const float bc = B0[o];
            const float z = pc - c1 * (bc - lap);
            dout[qz * n2 + (j0 + jj) * n + k0 + kk] = z;
            if (Dot) s += (double)bc * (double)z;
        }
    };
    Regs ra, rb, rc;
    load(i0 - 2, ra);
    load(i0 - 1, rb);
    load(i0, rc);
    for (int u = 0; u < U; u += 3) {
        step(u, ra);
        step(u + 1, rb);
        step(u + 2, rc);
    }
    if constexpr (Dot) {  // <b, d> partial for the direction pass (deferred finalisation)
        __shared__ double shd[2 * kBlock / kWave];
        double m = 0.0;
        block_sum_max_lds<kBlock>(s, m, shd);
        if (threadIdx.x == 0) red.psum[blockIdx.x] = s;
    }
}

bool mg_tile_ok(int tile, int n) {
    if (tile >= kMgScalarLoads && tile < kMgScalarLoads + kMgMarch16) tile -= kMgScalarLoads;
    if (tile < 0 || tile >= kMgTiles) return false;
    if (tile == kMgMarch16) return n >= 32 && n % 16 == 0;
    if (tile == kMgMarch8) return n >= 32 && n % 16 == 0;
    return true;
}

template <int S>
static void down_m(hipStream_t s, int n, const float* b, float* d2, float* bc, double c1,
                   double c2, CGState* st) {
    const int nblk = n / kMarchC;
    MAZE_LAUNCH((k_mg_down_march<S>), dim3(nblk * nblk * (n / S)), dim3(kBlock), 0, s, n, b, d2,
                bc, (float)c1, (float)c2, st);
}

template <int S>
static void up_m(hipStream_t s, int n, const float* b, const float* d2, const float* e, float* d,
                 double c1, double c2, CGState* st, int dot, Reduce red) {
    const int nblk = n / kMarchC;
    const dim3 g(nblk * nblk * (n / S));
    if (dot)
        MAZE_LAUNCH((k_mg_up_march<S, true>), g, dim3(kBlock), 0, s, n, b, d2, e, d, (float)c1,
                    (float)c2, red, st);
    else
        MAZE_LAUNCH((k_mg_up_march<S, false>), g, dim3(kBlock), 0, s, n, b, d2, e, d, (float)c1,
                    (float)c2, red, st);
}

const char* mg_tile_name(int tile) {
    static const char* names[kMgTiles] = {"8x16x16/1024", "8x16x16/512", "8x8x16/512",
                                          "4x8x8/512",    "4x8x8/256",   "march16x16/S16",
                                          "march16x16/S8"};
    if (tile >= kMgScalarLoads && tile < kMgScalarLoads + kMgMarch16)
        tile -= kMgScalarLoads;  // the same shape with scalar loads
    return tile >= 0 && tile < kMgTiles ? names[tile] : "?";
}

// Levels of 128^3 and more: 8x16x16 tiles (1024 or more work-groups), with
// 1024 threads for the down kernel and 512 for the up kernel (column-mapped
// at 128^3: down 11.8 against 12.4 us, up 15.0 against 15.9 us); smaller
// levels: 4x8x8 tiles with 512 threads, so that a 64^3 level still has 1024
// work-groups (8x16x16 would leave half the CUs idle).
int mg_default_tile(int n) { return n >= 128 ? kMgT8x16x16_1024 : kMgT4x8x8_512; }
int mg_default_tile_up(int n) {  // 64^3: 5.55 us with 256 threads against 6.22 with 512
    return n >= 128 ? kMgT8x16x16_512 : (n >= 64 ? kMgT4x8x8_256 : kMgT4x8x8_512);
}

void launch_mg_down(hipStream_t s, int n, int tile, const float* b, float* d2, float* bc,
                    double c1, double c2, CGState* st) {
    if (!mg_tile_ok(tile, n)) tile = mg_default_tile(n);
    const bool sc = tile >= kMgScalarLoads;
    switch (sc ? tile - kMgScalarLoads : tile) {
        case kMgMarch16: down_m<16>(s, n, b, d2, bc, c1, c2, st); break;
        case kMgMarch8: down_m<8>(s, n, b, d2, bc, c1, c2, st); break;
        case kMgT8x16x16_1024: down_t<8, 16, 16, 1024>(s, n, b, d2, bc, c1, c2, st, sc); break;
        case kMgT8x16x16_512: down_t<8, 16, 16, 512>(s, n, b, d2, bc, c1, c2, st, sc); break;
        case kMgT8x8x16_512: down_t<8, 8, 16, 512>(s, n, b, d2, bc, c1, c2, st, sc); break;
        case kMgT4x8x8_512: down_t<4, 8, 8, 512>(s, n, b, d2, bc, c1, c2, st, sc); break;
        default: down_t<4, 8, 8, 256>(s, n, b, d2, bc, c1, c2, st, sc); break;
    }
}

// Work-groups of the up kernel at a level (the partials its dot writes).
int mg_up_blocks(int n, int tile) {
    if (!mg_tile_ok(tile, n)) tile = mg_default_tile_up(n);
    switch (tile >= kMgScalarLoads ? tile - kMgScalarLoads : tile) {
        case kMgMarch16: return (n / kMarchC) * (n / kMarchC) * (n / 16);
        case kMgMarch8: return (n / kMarchC) * (n / kMarchC) * (n / 8);
        case kMgT8x16x16_1024:
        case kMgT8x16x16_512: return mg_tiles<8, 16, 16>(n);
        case kMgT8x8x16_512: return mg_tiles<8, 8, 16>(n);
        default: return mg_tiles<4, 8, 8>(n);
    }
}

void launch_mg_up(hipStream_t s, int n, int tile, const float* b, const float* d2, const float* e,
                  float* d, double c1, double c2, CGState* st, int dot, Reduce red) {
    if (!mg_tile_ok(tile, n)) tile = mg_default_tile_up(n);
    const bool sc = tile >= kMgScalarLoads;
    switch (sc ? tile - kMgScalarLoads : tile) {
        case kMgMarch16: up_m<16>(s, n, b, d2, e, d, c1, c2, st, dot, red); break;
        case kMgMarch8: up_m<8>(s, n, b, d2, e, d, c1, c2, st, dot, red); break;
        case kMgT8x16x16_1024: up_t<8, 16, 16, 1024>(s, n, b, d2, e, d, c1, c2, st, dot, red, sc); break;
        case kMgT8x16x16_512: up_t<8, 16, 16, 512>(s, n, b, d2, e, d, c1, c2, st, dot, red, sc); break;
        case kMgT8x8x16_512: up_t<8, 8, 16, 512>(s, n, b, d2, e, d, c1, c2, st, dot, red, sc); break;
        case kMgT4x8x8_512: up_t<4, 8, 8, 512>(s, n, b, d2, e, d, c1, c2, st, dot, red, sc); break;
        default: up_t<4, 8, 8, 256>(s, n, b, d2, e, d, c1, c2, st, dot, red, sc); break;
    }
}

// The level kernels in FP64 (the FP64 multigrid-PCG): the scalar kernels with
// double operands in two tile shapes, 8x8x16 with 512 threads from 128^3 (62
// and 56 KB of LDS for down and up) and 4x8x8 with 512 threads below.
int mg_up_blocks64(int n) { return n >= 128 ? mg_tiles<8, 8, 16>(n) : mg_tiles<4, 8, 8>(n); }

void launch_mg_down(hipStream_t s, int n, const double* b, double* d2, double* bc, double c1,
                    double c2, CGState* st) {
    if (n >= 128)
        MAZE_LAUNCH((k_mg_down<8, 8, 16, 512, double>), dim3(mg_tiles<8, 8, 16>(n)), dim3(512), 0,
                    s, n, b, d2, bc, c1, c2, st);
    else
        MAZE_LAUNCH((k_mg_down<4, 8, 8, 512, double>), dim3(mg_tiles<4, 8, 8>(n)), dim3(512), 0, s,
                    n, b, d2, bc, c1, c2, st);
}

void launch_mg_up(hipStream_t s, int n, const double* b, const double* d2, const double* e,
                  double* d, double c1, double c2, CGState* st, int dot, Reduce red) {
    const dim3 blk(512);
    if (n >= 128) {
        const dim3 g(mg_tiles<8, 8, 16>(n));
        if (dot)
            MAZE_LAUNCH((k_mg_up<8, 8, 16, 512, true, double>), g, blk, 0, s, n, b, d2, e, d, c1,
                        c2, red, st);
        else
            MAZE_LAUNCH((k_mg_up<8, 8, 16, 512, false, double>), g, blk, 0, s, n, b, d2, e, d, c1,
                        c2, red, st);
    } else {
        const dim3 g(mg_tiles<4, 8, 8>(n));
        if (dot)
            MAZE_LAUNCH((k_mg_up<4, 8, 8, 512, true, double>), g, blk, 0, s, n, b, d2, e, d, c1,
                        c2, red, st);
        else
            MAZE_LAUNCH((k_mg_up<4, 8, 8, 512, false, double>), g, blk, 0, s, n, b, d2, e, d, c1,
                        c2, red, st);
    }
}

// ---- exact FP64 solve on the coarsest grid -----------------------------------
// A e = b on a c^3 grid (c <= 12): forward transform along k, j, i with the
// real orthonormal eigenbasis Qf (row t = eigenvector t), divide by the
// eigenvalue sum (the constant mode is dropped), transform back.  A and B
// are c^3 double work arrays in LDS; `nt` threads cooperate.
constexpr int kMaxCoarse = 12;  // coarsest level of the one-work-group tail

__device__ void coarse_exact(int c, const double* Q, const double* __restrict__ ilam, double* A,
                             double* B, int tid, int nt) {
    const int N = c * c * c, c2 = c * c;
    for (int x = tid; x < N; x += nt) {  // forward k
        const int t = x % c, line = x - t;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[line + m];
        B[x] = s;
    }
    lds_barrier();
    for (int x = tid; x < N; x += nt) {  // forward j
        const int k = x % c, t = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * B[i * c2 + m * c + k];
        A[x] = s;
    }
    lds_barrier();
    for (int x = tid; x < N; x += nt) {  // forward i, scale by the inverse eigenvalue
        const int jk = x % c2, t = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[t * c + m] * A[m * c2 + jk];
        B[x] = s * ilam[x];
    }
    lds_barrier();
    for (int x = tid; x < N; x += nt) {  // inverse i
        const int jk = x % c2, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + i] * B[m * c2 + jk];
        A[x] = s;
    }
    lds_barrier();
    for (int x = tid; x < N; x += nt) {  // inverse j
        const int k = x % c, j = (x / c) % c, i = x / c2;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + j] * A[i * c2 + m * c + k];
        B[x] = s;
    }
    lds_barrier();
    for (int x = tid; x < N; x += nt) {  // inverse k
        const int k = x % c, line = x - k;
        double s = 0.0;
        for (int m = 0; m < c; m++) s += Q[m * c + k] * B[line + m];
        A[x] = s;
    }
    lds_barrier();
}

// Exact solve of a coarsest level of at most kMgCoarseLds^3 points in one
// work-group: the separable transform of coarse_exact, one axis per pass
// between two LDS arrays.  Four threads share one line of the level along
// the axis being transformed (the same input address: a broadcast) and each
// forms four of its c outputs.  Lines are padded to c + 1 values, which puts
// the lines of one wave on different LDS banks in all three orientations.
// The eigenbasis sits in LDS twice, as QF[m][t] = Q[t][m] (forward) and
// QI[m][t] = Q[m][t] (inverse), zero-padded to kMgCoarseLds^2.  T is the
// arithmetic: float, as in the rest of the V-cycle (measured 3 % faster per MD
// step at 128^3 than double, with the same cycle count; the FP64 refinement
// and stop test are unchanged).
// CT > 0 fixes c at compile time (fully unrolled line loops); CT = 0 reads it
// from c_arg.  The inverse eigenvalues of a thread's points are loaded first
// and kept in registers.
template <int CT, class T, class IO = float>
__global__ __launch_bounds__(4 * kMgCoarseLds * kMgCoarseLds) void k_mg_coarse(
    int c_arg, const IO* __restrict__ b, IO* __restrict__ e, const double* __restrict__ Qf,
    const double* __restrict__ ilam, const CGState* st) {
    if (st->done) return;
    constexpr int CM = kMgCoarseLds, NT = 4 * CM * CM, PER = CM * CM * CM / NT;
    __shared__ T A[CM * CM * (CM + 1)], B[CM * CM * (CM + 1)];
    __shared__ T QF[CM * CM], QI[CM * CM];
    const int c = CT > 0 ? CT : c_arg;
    const int tid = threadIdx.x, c2 = c * c, N = c2 * c, pitch = c + 1;
    T il[PER];
#pragma unroll
    for (int k = 0; k < PER; k++) il[k] = tid + k * NT < N ? (T)ilam[tid + k * NT] : T(0);
    if (tid < CM * CM) {
        const int m = tid / CM, t = tid % CM;
        const bool in = m < c && t < c;
        QF[tid] = in ? (T)Qf[t * c + m] : T(0);
        QI[tid] = in ? (T)Qf[m * c + t] : T(0);
    }
    for (int x = tid; x < N; x += NT) A[(x / c) * pitch + x % c] = (T)b[x];
    lds_barrier();
    // axis 0: lines along k at (i, j) = (u, v); 1: along j at (i, k); 2: along i at (j, k).
    auto pass = [&](const T* src, T* dst, int axis, bool inv) {
        const int line = tid >> 2, t0 = (tid & 3) * (CM / 4);
        if (line < c2) {
            const int u = line / c, v = line % c;
            const int base = axis == 0 ? (u * c + v) * pitch : (axis == 1 ? u * c * pitch + v : u * pitch + v);
            const int stride = axis == 0 ? 1 : (axis == 1 ? pitch : c * pitch);
            const T* Qm = (inv ? QI : QF) + t0;
            T s[CM / 4];
#pragma unroll
            for (int j = 0; j < CM / 4; j++) s[j] = T(0);
#pragma unroll
            for (int m = 0; m < (CT > 0 ? CT : CM); m++) {
                if (CT == 0 && m >= c) break;
                const T am = src[base + m * stride];
#pragma unroll
                for (int j = 0; j < CM / 4; j++) s[j] = fma(Qm[m * CM + j], am, s[j]);
            }
#pragma unroll
            for (int j = 0; j < CM / 4; j++)
                if (t0 + j < c) dst[base + (t0 + j) * stride] = s[j];
        }
        lds_barrier();
    };
    pass(A, B, 0, false);
    pass(B, A, 1, false);
    pass(A, B, 2, false);
#pragma unroll
    for (int k = 0; k < PER; k++) {
        const int x = tid + k * NT;
        if (x < N) B[(x / c) * pitch + x % c] *= il[k];
    }
    lds_barrier();
    pass(B, A, 2, true);
    pass(A, B, 1, true);
    pass(B, A, 0, true);
    for (int x = tid; x < N; x += NT) e[x] = (IO)A[(x / c) * pitch + x % c];
}

// The same solve with each axis pass as a matrix product on the matrix cores:
// a pass maps the c^2 lines of the level (rows) through the c x c eigenbasis,
// out[line][t] = sum_m in[line][m] M[m][t], as 16-line tiles of
// v_mfma_f32_16x16x4_f32 (four K-steps of 4 cover m < 16).  Wave w owns
// lines 16 w .. 16 w + 15 in every pass; each lane keeps the eigenbasis
// operands of both directions in registers for all six passes.  Operand
// layout of the 16x16x4 f32 instruction (wave64): A[i][k] from lane
// i + 16 k, B[k][j] from lane j + 16 k, and result register r of lane
// j + 16 q is D[4 q + r][j].  Values past c (padding) enter as exact zeros.
// FP32 products with FP32 accumulation, as the VALU form.
typedef float mg_v4f __attribute__((ext_vector_type(4)));
typedef double mg_v4d __attribute__((ext_vector_type(4)));
constexpr int kMgCoarseMfmaThreads = kMgCoarseLds * kMgCoarseLds * kMgCoarseLds / 4;

// The solve on NT threads (any multiple of 64 up to 1024): wave w takes the
// 16-line tiles w, w + NT / 64, ... of every pass.  load() issues the loads of
// the eigenbasis operands and the inverse eigenvalues into registers; solve()
// runs the six passes and the scaling on the right-hand side in A (rows of
// c + 1 floats) and writes the correction e.  C (the level's side) is a
// compile-time constant, so every index division by it is a multiply-shift (a
// run-time divisor cost more than the transforms).
template <int C, int NT>
struct CoarseMfma {
    static constexpr int CM = kMgCoarseLds;
    static_assert(CM == 16, "16-line tiles with 16-wide operands");
    static_assert(C >= 2 && C <= CM, "coarsest level of 2..16 points per side");
    static_assert(NT % 64 == 0 && NT <= 1024, "whole waves");
    static constexpr int c = C, c2 = c * c, N = c2 * c, pitch = c + 1;
    static constexpr int NW = NT / 64, TPW = ((c2 + 15) / 16 + NW - 1) / NW;  // tiles per wave
    static constexpr int kLds = CM * CM * (CM + 1);  // floats of A and of B
    // Inverse eigenvalues of this lane's results in the pass along i: tile t,
    // register r is line L0 + 4 (lane / 16) + r (the point (j, k)) at i = lane % 16.
    float il[TPW][4];
    // B operands: forward M[m][t] = Q[t][m], inverse M[m][t] = Q[m][t], with
    // m = 4 kk + lane / 16 and t = lane % 16.
    float qf[4], qi[4];

    __device__ __forceinline__ void load(const double* Qf, const double* ilam) {
        const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
        const int col = lane & 15, krow = lane >> 4;
#pragma unroll
        for (int t = 0; t < TPW; t++)
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int Ld = 16 * (wave + t * NW) + 4 * krow + r;
                il[t][r] = Ld < c2 && col < c ? (float)ilam[(col * c + Ld / c) * c + Ld % c] : 0.0f;
            }
#pragma unroll
        for (int kk = 0; kk < 4; kk++) {
            const int m = 4 * kk + krow;
            const bool in = m < c && col < c;
            qf[kk] = in ? (float)Qf[col * c + m] : 0.0f;
            qi[kk] = in ? (float)Qf[m * c + col] : 0.0f;
        }
    }

    static __device__ __forceinline__ int rhs_index(int x) { return (x / c) * pitch + x % c; }

    // The right-hand side is in A; the barrier that makes it visible comes first.
    __device__ __forceinline__ void solve(float* A, float* B, float* __restrict__ e) {
        const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
        const int col = lane & 15, krow = lane >> 4;
        lds_barrier();
        // axis 0: lines along k at (i, j) = (u, v); 1: along j at (i, k); 2: along i at (j, k).
        auto line_base = [&](int L, int axis) {
            const int u = L / c, v = L - (L / c) * c;
            return axis == 0 ? (u * c + v) * pitch : (axis == 1 ? u * c * pitch + v : u * pitch + v);
        };
        // One 16-line tile of a pass: the transform of lines L0 .. L0 + 15 from src.
        auto tile = [&](const float* src, int L0, int axis, bool inv) {
            const int stride = axis == 0 ? 1 : (axis == 1 ? pitch : c * pitch);
            const int La = L0 + col;  // A row of this lane
            const int ba = La < c2 ? line_base(La, axis) : 0;
            mg_v4f acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < 4; kk++) {
                const int m = 4 * kk + krow;
                const float a = (La < c2 && m < c) ? src[ba + m * stride] : 0.0f;
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, inv ? qi[kk] : qf[kk], acc, 0, 0, 0);
            }
            return acc;
        };
        auto put = [&](float* dst, int L0, int axis, const mg_v4f& acc) {
            const int stride = axis == 0 ? 1 : (axis == 1 ? pitch : c * pitch);
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int Ld = L0 + 4 * krow + r;  // D row of register r
                if (Ld < c2 && col < c) dst[line_base(Ld, axis) + col * stride] = acc[r];
            }
        };
        auto pass = [&](const float* src, float* dst, int axis, bool inv) {
            for (int L0 = 16 * wave; L0 < c2; L0 += 16 * NW) put(dst, L0, axis, tile(src, L0, axis, inv));
            lds_barrier();
        };
        pass(A, B, 0, false);
        pass(B, A, 1, false);
        // Along i: forward, scaled by the inverse eigenvalues, and back.  A wave
        // reads and writes only the positions of its own lines here, and its
        // LDS operations complete in order, so no barrier is needed inside.
#pragma unroll
        for (int t = 0; t < TPW; t++) {
            const int L0 = 16 * (wave + t * NW);
            if (L0 < c2) {
                mg_v4f acc = tile(A, L0, 2, false);
#pragma unroll
                for (int r = 0; r < 4; r++) acc[r] *= il[t][r];
                put(B, L0, 2, acc);
                put(A, L0, 2, tile(B, L0, 2, true));
            }
        }
        lds_barrier();
        pass(A, B, 1, true);
        pass(B, A, 0, true);
        for (int x = tid; x < N; x += NT) e[x] = A[rhs_index(x)];
    }
};

template <int C>
__global__ __launch_bounds__(kMgCoarseMfmaThreads) void k_mg_coarse_mfma(
    const float* __restrict__ b, float* __restrict__ e, const double* __restrict__ Qf,
    const double* __restrict__ ilam, const CGState* st) {
    using CS = CoarseMfma<C, kMgCoarseMfmaThreads>;
    const int done = st->done;  // tested once the loads are in flight (as k_mg_up_vec)
    __shared__ float A[CS::kLds], B[CS::kLds];
    CS cs;
    cs.load(Qf, ilam);
    for (int x = threadIdx.x; x < CS::N; x += kMgCoarseMfmaThreads) A[CS::rhs_index(x)] = b[x];
    if (done) return;
    cs.solve(A, B, e);
}

// The level above a coarsest grid of C points, with the coarse solve in the
// same kernel: every work-group stores its part of the coarse right-hand side
// at device scope (the XCDs' L2 caches are not coherent with each other),
// waits until those stores have completed and counts its arrival; the last to
// arrive reads the whole right-hand side back at device scope, solves the
// coarse system (CoarseMfma on the kernel's threads) and writes its
// correction, then clears the counter for the next cycle.  The same hand-off
// as the single-pass reductions; it saves the coarse kernel's launch and its
// wait for a free CU.
template <int TI, int TJ, int TK, int NT, int V, int C>
__global__ __launch_bounds__(NT) void k_mg_down_coarse(int n, const float* __restrict__ b,
                                                       float* __restrict__ d2out,
                                                       float* __restrict__ bcoarse, float c1,
                                                       float c2, const CGState* st,
                                                       CoarseTail tail) {
    if (!mg_down_vec_body<TI, TJ, TK, NT, V, true>(n, b, d2out, bcoarse, c1, c2, st)) return;
    __asm__ volatile("s_waitcnt vmcnt(0)" ::: "memory");
    __syncthreads();
    __shared__ bool last;
    if (threadIdx.x == 0)
        last = __hip_atomic_fetch_add(tail.count, 1u, __ATOMIC_RELAXED,
                                      __HIP_MEMORY_SCOPE_AGENT) == gridDim.x - 1;
    __syncthreads();
    if (!last) return;
    using CS = CoarseMfma<C, NT>;
    __shared__ float A[CS::kLds], B[CS::kLds];
    CS cs;
    cs.load(tail.Qf, tail.ilam);
    for (int x = threadIdx.x; x < CS::N; x += NT)
        A[CS::rhs_index(x)] =
            __hip_atomic_load(bcoarse + x, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    if (threadIdx.x == 0) reset_counter(tail.count);
    cs.solve(A, B, tail.e);
}

template <int TI, int TJ, int TK, int NT>
static void down_coarse_t(hipStream_t s, int n, const float* b, float* d2, float* bc, double c1,
                          double c2, CGState* st, const CoarseTail& tail) {
    MAZE_LAUNCH((k_mg_down_coarse<TI, TJ, TK, NT, 4, kMgCoarseLds>), dim3(mg_tiles<TI, TJ, TK>(n)),
                dim3(NT), 0, s, n, b, d2, bc, (float)c1, (float)c2, st, tail);
}

// Built for the 32^3 level above the 16^3 coarsest (16-byte loads) in the five
// tile shapes; the plane-marching and scalar-load shapes keep the separate
// coarse kernel.
bool mg_down_coarse_ok(int n, int tile, int c) {
    return c == kMgCoarseLds && n == 2 * c && mg_vec_widths(n, false) == 8 * 4 + 4 && tile >= 0 &&
           tile < kMgMarch16;
}

bool launch_mg_down_coarse(hipStream_t s, int n, int tile, const float* b, float* d2, float* bc,
                           double c1, double c2, CGState* st, int c, const CoarseTail& tail) {
    if (!mg_down_coarse_ok(n, tile, c)) return false;
    switch (tile) {
        case kMgT8x16x16_1024: down_coarse_t<8, 16, 16, 1024>(s, n, b, d2, bc, c1, c2, st, tail); return true;
        case kMgT8x16x16_512: down_coarse_t<8, 16, 16, 512>(s, n, b, d2, bc, c1, c2, st, tail); return true;
        case kMgT8x8x16_512: down_coarse_t<8, 8, 16, 512>(s, n, b, d2, bc, c1, c2, st, tail); return true;
        case kMgT4x8x8_512: down_coarse_t<4, 8, 8, 512>(s, n, b, d2, bc, c1, c2, st, tail); return true;
        case kMgT4x8x8_256: down_coarse_t<4, 8, 8, 256>(s, n, b, d2, bc, c1, c2, st, tail); return true;
        default: return false;  // plane-marching and scalar-load shapes
    }
}

void launch_mg_coarse(hipStream_t s, int c, const float* b, float* e, const double* Qf,
                      const double* ilam, CGState* st) {
    if (c > kMgCoarseLds) {
        std::fprintf(stderr, "maze: coarsest level %d above %d for the one-work-group solve\n", c,
                     kMgCoarseLds);
        std::exit(1);
    }
    // MAZE_MG_COARSE_VALU=1 selects the vector-ALU form (the same transform).
    static const bool valu = [] {
        const char* v = std::getenv("MAZE_MG_COARSE_VALU");
        return v && std::atoi(v) != 0;
    }();
    if (!valu) {
        switch (c) {
#define MAZE_COARSE_CASE(C)                                                                 \
    case C:                                                                                 \
        MAZE_LAUNCH(k_mg_coarse_mfma<C>, dim3(1), dim3(kMgCoarseMfmaThreads), 0, s, b, e, Qf, \
                    ilam, st);                                                              \
        return;
            MAZE_COARSE_CASE(2) MAZE_COARSE_CASE(3) MAZE_COARSE_CASE(4) MAZE_COARSE_CASE(5)
            MAZE_COARSE_CASE(6) MAZE_COARSE_CASE(7) MAZE_COARSE_CASE(8) MAZE_COARSE_CASE(9)
            MAZE_COARSE_CASE(10) MAZE_COARSE_CASE(11) MAZE_COARSE_CASE(12) MAZE_COARSE_CASE(13)
            MAZE_COARSE_CASE(14) MAZE_COARSE_CASE(15) MAZE_COARSE_CASE(16)
#undef MAZE_COARSE_CASE
            default:
                break;
        }
    }
    const dim3 grid(1), block(4 * kMgCoarseLds * kMgCoarseLds);
    if (c == kMgCoarseLds)
        MAZE_LAUNCH((k_mg_coarse<kMgCoarseLds, float>), grid, block, 0, s, c, b, e, Qf, ilam, st);
    else
        MAZE_LAUNCH((k_mg_coarse<0, float>), grid, block, 0, s, c, b, e, Qf, ilam, st);
}

// FP64 form (the FP64 multigrid-PCG): the vector-ALU transform in FP64.
void launch_mg_coarse(hipStream_t s, int c, const double* b, double* e, const double* Qf,
                      const double* ilam, CGState* st) {
    if (c > kMgCoarseLds) {
        std::fprintf(stderr, "maze: coarsest level %d above %d for the one-work-group solve\n", c,
                     kMgCoarseLds);
        std::exit(1);
    }
    const dim3 grid(1), block(4 * kMgCoarseLds * kMgCoarseLds);
    if (c == kMgCoarseLds)
        MAZE_LAUNCH((k_mg_coarse<kMgCoarseLds, double, double>), grid, block, 0, s, c, b, e, Qf,
                    ilam, st);
    else
        MAZE_LAUNCH((k_mg_coarse<0, double, double>), grid, block, 0, s, c, b, e, Qf, ilam, st);
}

// ---- exact solve on a larger coarsest grid (kMgCoarseLds < c <= kMgCoarseMax) --
// The separable transform of coarse_exact with one axis per kernel over
// global work arrays.  A pass transforms the last axis of its input and
// writes it first, out[t][i][j] = sum_m M[t][m] in[i][j][m] with M = Qf
// (forward) or Qf^T (inverse), so three passes cover the three axes and
// restore the order.  A work-group stages R consecutive rows of one plane
// (contiguous in memory) in LDS.  Mode 1 reads the FP32 right-hand side, mode
// 2 scales by the inverse eigenvalues (after the third forward pass), mode 3
// writes the FP32 correction; mode 0 is double to double.  Every output sums
// its c products in the same order for any R.
template <int Mode, bool Inv, int R>
__global__ __launch_bounds__(kBlock) void k_mg_coarse_axis(int c, const void* __restrict__ in_,
                                                           void* __restrict__ out_,
                                                           const double* __restrict__ Qf,
                                                           const double* __restrict__ ilam,
                                                           const CGState* st) {
    if (st->done) return;
    __shared__ double rows[R * kMgCoarseMax];
    const int nj = (c + R - 1) / R;
    const int i = (int)blockIdx.x / nj, j0 = ((int)blockIdx.x % nj) * R;
    const int J = min(R, c - j0);
    const long base = ((long)i * c + j0) * c;
    for (int x = threadIdx.x; x < J * c; x += kBlock) {
        if constexpr (Mode == 1) rows[x] = (double)static_cast<const float*>(in_)[base + x];
        else rows[x] = static_cast<const double*>(in_)[base + x];
    }
    lds_barrier();
    for (int x = threadIdx.x; x < J * c; x += kBlock) {
        const int jj = x % J, t = x / J;
        const double* r = rows + jj * c;
        double s = 0.0;
        if constexpr (Inv) {
            for (int m = 0; m < c; m++) s += Qf[m * c + t] * r[m];
        } else {
            for (int m = 0; m < c; m++) s += Qf[t * c + m] * r[m];
        }
        const long o = ((long)t * c + i) * c + j0 + jj;
        if constexpr (Mode == 2) s *= ilam[o];
        if constexpr (Mode == 3) static_cast<float*>(out_)[o] = (float)s;
        else static_cast<double*>(out_)[o] = s;
    }
}

// The axis passes on the matrix cores (v_mfma_f64_16x16x4f64): a pass is the
// product OUT[line][t] = sum_m IN[line][m] M[m][t] of the c^2 lines of the
// level (rows of c values, contiguous) with the c x c eigenbasis M = Qf^T
// (forward) or Qf (inverse).  A work-group takes 16 consecutive lines of one
// plane (i, j0 .. j0 + 15), stages them in LDS zero-padded to a multiple of 4
// values, and each of its four waves forms 16 x 16 output tiles (t0 .. t0 +
// 15) in K-steps of 4.  Operand layout of the 16x16x4 f64 instruction
// (wave64, checked on the GPU): A[i][k] from lane i + 16 k, B[k][j] from lane
// j + 16 k, result register r of lane j + 16 q is D[q + 4 r][j] (the f32 form
// has D[4 q + r][j]); values past c enter as exact zeros.  The products are summed four at a time into the accumulator, so the
// sums differ from the vector-ALU form in the last bits.
template <int Mode, bool Inv>
__global__ __launch_bounds__(kBlock) void k_mg_coarse_axis_mfma(int c, const void* __restrict__ in_,
                                                                void* __restrict__ out_,
                                                                const double* __restrict__ Qf,
                                                                const double* __restrict__ ilam,
                                                                const CGState* st) {
    if (st->done) return;
    constexpr int PMAX = kMgCoarseMax + 5;  // padded row pitch (odd, against bank conflicts)
    __shared__ double rows[16 * PMAX];
    const int ng = (c + 15) / 16;
    const int i = (int)blockIdx.x / ng, j0 = ((int)blockIdx.x % ng) * 16;
    const int J = min(16, c - j0);
    const int K4 = (c + 3) & ~3, P = K4 + 1;
    const long base = ((long)i * c + j0) * c;
    for (int x = threadIdx.x; x < 16 * K4; x += kBlock) {
        const int line = x / K4, m = x % K4;
        double v = 0.0;
        if (line < J && m < c) {
            if constexpr (Mode == 1) v = (double)static_cast<const float*>(in_)[base + (long)line * c + m];
            else v = static_cast<const double*>(in_)[base + (long)line * c + m];
        }
        rows[line * P + m] = v;
    }
    lds_barrier();
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    const int lr = lane & 15, lq = lane >> 4;
    for (int tt = wave; tt < (c + 15) / 16; tt += kBlock / 64) {
        const int t0 = tt * 16, tb = t0 + lr;
        mg_v4d acc = {0.0, 0.0, 0.0, 0.0};
        for (int m0 = 0; m0 < K4; m0 += 4) {
            const int m = m0 + lq;
            const double a = rows[lr * P + m];
            const double b = (m < c && tb < c) ? (Inv ? Qf[m * c + tb] : Qf[tb * c + m]) : 0.0;
            acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a, b, acc, 0, 0, 0);
        }
        if (tb < c) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int line = lq + 4 * r;
                if (line < J) {
                    const long o = ((long)tb * c + i) * c + j0 + line;
                    double s = acc[r];
                    if constexpr (Mode == 2) s *= ilam[o];
                    if constexpr (Mode == 3) static_cast<float*>(out_)[o] = (float)s;
                    else static_cast<double*>(out_)[o] = s;
                }
            }
        }
    }
}

// ---- the same solve in three kernels for c <= kMgPlaneMax -------------------------
// The six axis passes regrouped by plane: the forward transform of the two
// in-plane axes of every plane i (Z_i = Qf X_i Qf^T), then per line along i the
// forward transform, the scaling by the inverse eigenvalues and the inverse
// transform (the scaling is point-wise, so the i axis goes forward and back in
// one kernel), then the inverse transform of the two in-plane axes (X_i =
// Qf^T Z_i Qf).  Every product runs on the FP64 matrix cores with its operands
// in LDS (the same operand layout as k_mg_coarse_axis_mfma); arrays keep the
// natural [i][j][k] order, so each kernel reads and writes contiguous rows.
// One plane of up to 80 x 80 values, its intermediate and the eigenbasis take
// 152 KB of LDS.
constexpr int kMgPlaneMax = 80;
constexpr int kMgPlaneThreads = 1024;
constexpr int kMgPlanePitch = kMgPlaneMax + 1;

// One 16 x 16 tile of the product A B (K-steps of 4 over kq values) on the
// matrix cores; a(row, m) and b(m, col) give the operands, tile-relative.
template <class FA, class FB>
__device__ __forceinline__ mg_v4d mfma_tile_f64(int kq, FA&& a, FB&& b) {
    const int lane = threadIdx.x & 63, lr = lane & 15, lk = lane >> 4;
    mg_v4d acc = {0.0, 0.0, 0.0, 0.0};
    for (int m0 = 0; m0 < kq; m0 += 4)
        acc = __builtin_amdgcn_mfma_f64_16x16x4f64(a(lr, m0 + lk), b(m0 + lk, lr), acc, 0, 0, 0);
    return acc;
}

// Both in-plane axes of plane blockIdx.x: forward Z = Qf X Qf^T, inverse
// Z = Qf^T X Qf.  In: the FP32 right-hand side (forward) or doubles; Out:
// doubles, or the FP32 correction (inverse).
template <bool Fwd, class In, class Out>
__global__ __launch_bounds__(kMgPlaneThreads) void k_mg_coarse_plane(int c,
                                                                    const In* __restrict__ in,
                                                                    Out* __restrict__ out,
                                                                    const double* __restrict__ Qf,
                                                                    const CGState* st) {
    const int done = st->done;  // tested once the loads are in flight (as k_mg_up_vec)
    constexpr int P = kMgPlanePitch;
    __shared__ double X[kMgPlaneMax * P], Y[kMgPlaneMax * P], Q[kMgPlaneMax * P];
    const int CP = (c + 15) & ~15, K4 = (c + 3) & ~3;
    const long plane = (long)blockIdx.x * c * c;
    // All of a thread's loads are issued before the first store (every load
    // goes to a valid address; values past c are zeroed afterwards), so the
    // staging takes one memory round trip.
    constexpr int U = (kMgPlaneMax * kMgPlaneMax + kMgPlaneThreads - 1) / kMgPlaneThreads;
    double vq[U], vx[U];
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int x = min((int)threadIdx.x + u * kMgPlaneThreads, CP * CP - 1);
        const int r = x / CP, q = x % CP;
        const bool in_c = r < c && q < c;
        const double a = Qf[in_c ? r * c + q : 0];
        const double v = (double)in[plane + (in_c ? (long)r * c + q : 0)];
        vq[u] = in_c ? a : 0.0;
        vx[u] = in_c ? v : 0.0;
    }
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < U; u++) {
        const int x = (int)threadIdx.x + u * kMgPlaneThreads;
        if (x < CP * CP) {
            const int r = x / CP, q = x % CP;
            Q[r * P + q] = vq[u];
            X[r * P + q] = vx[u];
        }
    }
    if (done) return;
    lds_barrier();
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int nt = CP / 16, rows1 = (K4 + 15) / 16;  // Y rows past K4 are not needed
    // Y = X B1 along the columns: B1[m][t] = Qf[t][m] (forward) or Qf[m][t].
    for (int tt = wave; tt < rows1 * nt; tt += kMgPlaneThreads / 64) {
        const int r0 = (tt / nt) * 16, t0 = (tt % nt) * 16;
        const mg_v4d acc = mfma_tile_f64(
            K4, [&](int r, int m) { return X[(r0 + r) * P + m]; },
            [&](int m, int t) { return Fwd ? Q[(t0 + t) * P + m] : Q[m * P + t0 + t]; });
#pragma unroll
        for (int r = 0; r < 4; r++) Y[(r0 + lq + 4 * r) * P + t0 + lr] = acc[r];
    }
    lds_barrier();
    // Z = A2 Y along the rows: A2[s][j] = Qf[s][j] (forward) or Qf[j][s].
    for (int tt = wave; tt < nt * nt; tt += kMgPlaneThreads / 64) {
        const int s0 = (tt / nt) * 16, t0 = (tt % nt) * 16;
        const mg_v4d acc = mfma_tile_f64(
            K4, [&](int s, int j) { return Fwd ? Q[(s0 + s) * P + j] : Q[j * P + s0 + s]; },
            [&](int j, int t) { return Y[j * P + t0 + t]; });
        const int t = t0 + lr;
        if (t < c) {
#pragma unroll
            for (int r = 0; r < 4; r++) {
                const int sr = s0 + lq + 4 * r;
                if (sr < c) out[plane + (long)sr * c + t] = (Out)acc[r];
            }
        }
    }
}

// The i axis of 16 lines (j = blockIdx.x / ng, k = k0 .. k0 + 15): forward,
// scaled by the inverse eigenvalues, inverse; in place on w.
constexpr int kMgLineThreads = 256;
__global__ __launch_bounds__(kMgLineThreads) void k_mg_coarse_line(int c, double* __restrict__ w,
                                                                   const double* __restrict__ Qf,
                                                                   const double* __restrict__ ilam,
                                                                   const CGState* st) {
    const int done = st->done;  // tested once the loads are in flight (as k_mg_up_vec)
    constexpr int P = kMgPlanePitch;
    __shared__ double L[16 * P], O[16 * P], Q[kMgPlaneMax * P];
    const int CP = (c + 15) & ~15, K4 = (c + 3) & ~3, c2 = c * c;
    const int ng = (c + 15) / 16;
    const int j = (int)blockIdx.x / ng, k0 = ((int)blockIdx.x % ng) * 16;
    const int wave = threadIdx.x >> 6, lane = threadIdx.x & 63, lr = lane & 15, lq = lane >> 4;
    const int nt = CP / 16;
    // All of a thread's loads (the eigenbasis, the 16 lines and the inverse
    // eigenvalues its result tiles take) are issued before the first store
    // (every load goes to a valid address; values past c are zeroed after), so
    // the staging takes one memory round trip.
    constexpr int UQ = (kMgPlaneMax * kMgPlaneMax + kMgLineThreads - 1) / kMgLineThreads;
    constexpr int UL = (16 * kMgPlaneMax + kMgLineThreads - 1) / kMgLineThreads;
    constexpr int TW = (kMgPlaneMax / 16 + kMgLineThreads / 64 - 1) / (kMgLineThreads / 64);
    double vq[UQ], vl[UL], il[TW][4];
#pragma unroll
    for (int u = 0; u < UQ; u++) {
        const int x = min((int)threadIdx.x + u * kMgLineThreads, CP * CP - 1);
        const int r = x / CP, q = x % CP;
        const bool in_c = r < c && q < c;
        const double a = Qf[in_c ? r * c + q : 0];
        vq[u] = in_c ? a : 0.0;
    }
#pragma unroll
    for (int u = 0; u < UL; u++) {
        const int x = min((int)threadIdx.x + u * kMgLineThreads, 16 * CP - 1);
        const int l = x & 15, i = x >> 4;  // consecutive threads: consecutive k
        const bool in_c = i < c && k0 + l < c;
        const double a = w[in_c ? (long)i * c2 + j * c + k0 + l : 0];
        vl[u] = in_c ? a : 0.0;
    }
#pragma unroll
    for (int v = 0; v < TW; v++) {
        const int t = (wave + v * (kMgLineThreads / 64)) * 16 + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) {
            const int l = lq + 4 * r;
            const bool in_c = t < c && k0 + l < c;
            const double a = ilam[in_c ? (long)t * c2 + j * c + k0 + l : 0];
            il[v][r] = in_c ? a : 0.0;
        }
    }
    __asm__ volatile("" ::: "memory");
#pragma unroll
    for (int u = 0; u < UQ; u++) {
        const int x = (int)threadIdx.x + u * kMgLineThreads;
        if (x < CP * CP) Q[(x / CP) * P + x % CP] = vq[u];
    }
#pragma unroll
    for (int u = 0; u < UL; u++) {
        const int x = (int)threadIdx.x + u * kMgLineThreads;
        if (x < 16 * CP) L[(x & 15) * P + (x >> 4)] = vl[u];
    }
    if (done) return;
    lds_barrier();
    // O[l][t] = ilam[t][j][k0 + l] sum_i L[l][i] Qf[t][i]
#pragma unroll
    for (int v = 0; v < TW; v++) {
        const int tt = wave + v * (kMgLineThreads / 64);
        if (tt >= nt) break;
        const int t0 = tt * 16;
        const mg_v4d acc = mfma_tile_f64(
            K4, [&](int l, int m) { return L[l * P + m]; },
            [&](int m, int t) { return Q[(t0 + t) * P + m]; });
        const int t = t0 + lr;
#pragma unroll
        for (int r = 0; r < 4; r++) O[(lq + 4 * r) * P + t] = acc[r] * il[v][r];
    }
    lds_barrier();
    // L[l][i] = sum_t O[l][t] Qf[t][i]
    for (int tt = wave; tt < nt; tt += kMgLineThreads / 64) {
        const int i0 = tt * 16;
        const mg_v4d acc = mfma_tile_f64(
            K4, [&](int l, int m) { return O[l * P + m]; },
            [&](int m, int i) { return Q[m * P + i0 + i]; });
#pragma unroll
        for (int r = 0; r < 4; r++) L[(lq + 4 * r) * P + i0 + lr] = acc[r];
    }
    lds_barrier();
    for (int x = threadIdx.x; x < 16 * c; x += kMgLineThreads) {
        const int l = x & 15, i = x >> 4;
        if (k0 + l < c) w[(long)i * c2 + j * c + k0 + l] = L[l * P + i];
    }
}

static void coarse_plane_solve(hipStream_t s, int c, const float* b, float* e, double* w0,
                               const double* Qf, const double* ilam, CGState* st) {
    MAZE_LAUNCH((k_mg_coarse_plane<true, float, double>), dim3(c), dim3(kMgPlaneThreads), 0, s, c,
                b, w0, Qf, st);
    MAZE_LAUNCH(k_mg_coarse_line, dim3(c * ((c + 15) / 16)), dim3(kMgLineThreads), 0, s, c, w0,
                Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_plane<false, double, float>), dim3(c), dim3(kMgPlaneThreads), 0, s,
                c, w0, e, Qf, st);
}

template <int R>
static void coarse_axis_passes(hipStream_t s, int c, const float* b, float* e, double* w0,
                               double* w1, const double* Qf, const double* ilam, CGState* st) {
    const dim3 grid(c * ((c + R - 1) / R)), block(kBlock);
    MAZE_LAUNCH((k_mg_coarse_axis<1, false, R>), grid, block, 0, s, c, b, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis<0, false, R>), grid, block, 0, s, c, w0, w1, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis<2, false, R>), grid, block, 0, s, c, w1, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis<0, true, R>), grid, block, 0, s, c, w0, w1, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis<0, true, R>), grid, block, 0, s, c, w1, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis<3, true, R>), grid, block, 0, s, c, w0, e, Qf, ilam, st);
}

static void coarse_axis_passes_mfma(hipStream_t s, int c, const float* b, float* e, double* w0,
                                    double* w1, const double* Qf, const double* ilam,
                                    CGState* st) {
    const dim3 grid(c * ((c + 15) / 16)), block(kBlock);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<1, false>), grid, block, 0, s, c, b, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<0, false>), grid, block, 0, s, c, w0, w1, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<2, false>), grid, block, 0, s, c, w1, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<0, true>), grid, block, 0, s, c, w0, w1, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<0, true>), grid, block, 0, s, c, w1, w0, Qf, ilam, st);
    MAZE_LAUNCH((k_mg_coarse_axis_mfma<3, true>), grid, block, 0, s, c, w0, e, Qf, ilam, st);
}

void launch_mg_coarse_big(hipStream_t s, int c, const float* b, float* e, double* w0, double* w1,
                          const double* Qf, const double* ilam, CGState* st) {
    if (c > kMgCoarseMax) {
        std::fprintf(stderr, "maze: coarsest multigrid level %d above %d\n", c, kMgCoarseMax);
        std::exit(1);
    }
    // The matrix-core passes by default; MAZE_COARSE_ROWS=4, 8 or 16 selects the
    // vector-ALU passes with that many rows per work-group (4 was their best:
    // passes on coarsest grids of 65^3, 51^3 and 41^3 points measured 7.3 us on
    // average against 9.6 us with 16 rows).
    // Up to kMgPlaneMax the three plane / line kernels (MAZE_COARSE_ROWS=1
    // selects the six matrix-core axis passes there too).
    static const int rows = [] {
        const char* v = std::getenv("MAZE_COARSE_ROWS");
        const int r = v ? std::atoi(v) : 0;
        return r == 1 || r == 4 || r == 8 || r == 16 ? r : 0;
    }();
    if (rows == 0 && c <= kMgPlaneMax) coarse_plane_solve(s, c, b, e, w0, Qf, ilam, st);
    else if (rows <= 1) coarse_axis_passes_mfma(s, c, b, e, w0, w1, Qf, ilam, st);
    else if (rows == 4) coarse_axis_passes<4>(s, c, b, e, w0, w1, Qf, ilam, st);
    else if (rows == 8) coarse_axis_passes<8>(s, c, b, e, w0, w1, Qf, ilam, st);
    else coarse_axis_passes<16>(s, c, b, e, w0, w1, Qf, ilam, st);
}

// ---- the coarse tail of the V-cycle in one work-group --------------------------
// Levels of at most 16^3 points are launch-latency bound when each pass is its
// own kernel.  k_mg_tail copies the right-hand side of the first such level
// into LDS and runs the rest of the cycle there -- the same pre-smoothing,
// residual and restriction down to the exact FP64 solve, prolongation and
// post-smoothing back up (FP32, as the fused level kernels) -- then writes
// that level's correction back; barriers separate the passes.  LDS holds b, d,
// t of every level (at most 3 * (16^3 + 8^3 + ...) floats); the exact solve's
// two double arrays reuse the first level's t.
//
// Thread (ti, tj, tk) = (tid >> 8, (tid >> 4) & 15, tid & 15) owns the (j, k)
// column (tj, tk) of every level (when tj, tk < m) at planes ti, ti + 4, ...,
// so neighbour offsets are set once per level and no index is divided.
constexpr int kTailThreads = 1024;
constexpr int kTailMaxLevels = 6;
constexpr int kTailMaxPts = 4096;  // 16^3
constexpr int kTailPool = 7168;    // doubles: 57344 bytes

struct TailLevels {
    int nlev;  // levels handled, the last one solved exactly
    int n[kTailMaxLevels];
};

// Periodic neighbours of column (j, k) on an m^3 level: row offsets jm, j0, jp
// (times m) and k indices km, k0, kp.
struct TailCol {
    int jm, j0, jp, km, k0, kp;
    __device__ TailCol(int j, int k, int m) {
        jm = (j == 0 ? m - 1 : j - 1) * m;
        j0 = j * m;
        jp = (j == m - 1 ? 0 : j + 1) * m;
        km = k == 0 ? m - 1 : k - 1;
        k0 = k;
        kp = k == m - 1 ? 0 : k + 1;
    }
    // A u at plane offsets (im, i0, ip) (times m^2) for this column.
    __device__ float lap(const float* u, int im, int i0, int ip) const {
        return u[im + j0 + k0] + u[ip + j0 + k0] + u[i0 + jm + k0] + u[i0 + jp + k0] +
               u[i0 + j0 + km] + u[i0 + j0 + kp] - 6.0f * u[i0 + j0 + k0];
    }
};

__global__ __launch_bounds__(kTailThreads) void k_mg_tail(TailLevels T, const float* __restrict__ bin,
                                                         float* __restrict__ dout, float c1,
                                                         float c2,
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
    const int tk = tid & 15, tj = (tid >> 4) & 15, ti = tid >> 8;
    {
        const int m = T.n[0], m2 = m * m;
        if (tj < m && tk < m)
            for (int i = ti; i < m; i += 4) {
                const int x = i * m2 + tj * m + tk;
                B[0][x] = bin[x];
            }
    }
    lds_barrier();
    const float cs = c1 + c2, cp = c1 * c2;
    // ---- down: d2 = -(c1 + c2) b - c1 c2 A b, t = b - A d2, restrict t ----
    for (int l = 0; l < T.nlev - 1; l++) {
        const int m = T.n[l], m2 = m * m;
        const float* b = B[l];
        float* d = D[l];
        float* t = R[l];
        const bool on = tj < m && tk < m;
        const TailCol col(on ? tj : 0, on ? tk : 0, m);
        if (on)
            for (int i = ti; i < m; i += 4) {
                const int i0 = i * m2, im = (i == 0 ? m - 1 : i - 1) * m2,
                          ip = (i == m - 1 ? 0 : i + 1) * m2;
                const int x = i0 + col.j0 + col.k0;
                d[x] = -cs * b[x] - cp * col.lap(b, im, i0, ip);
            }
        lds_barrier();
        if (on)
            for (int i = ti; i < m; i += 4) {
                const int i0 = i * m2, im = (i == 0 ? m - 1 : i - 1) * m2,
                          ip = (i == m - 1 ? 0 : i + 1) * m2;
                const int x = i0 + col.j0 + col.k0;
                t[x] = b[x] - col.lap(d, im, i0, ip);
            }
        lds_barrier();
        const int M = m / 2;
        float* bc = B[l + 1];
        if (tj < M && tk < M) {
            const int j = 2 * tj, k = 2 * tk;
            const int jj[3] = {(j == 0 ? m - 1 : j - 1) * m, j * m, (j + 1) * m};
            const int kk[3] = {k == 0 ? m - 1 : k - 1, k, k + 1};
            for (int I = ti; I < M; I += 4) {
                const int i = 2 * I;
                const int ii[3] = {(i == 0 ? m - 1 : i - 1) * m2, i * m2, (i + 1) * m2};
                float acc = 0.0f;
#pragma unroll
                for (int a = 0; a < 3; a++)
#pragma unroll
                    for (int bb = 0; bb < 3; bb++)
#pragma unroll
                        for (int cc = 0; cc < 3; cc++) {
                            const float w = (float)((a == 1 ? 2 : 1) * (bb == 1 ? 2 : 1) * (cc == 1 ? 2 : 1));
                            acc += w * t[ii[a] + jj[bb] + kk[cc]];
                        }
                bc[(I * M + tj) * M + tk] = acc * (4.0f / 64.0f);
            }
        }
        lds_barrier();
    }
    // ---- exact FP64 solve on the coarsest level ----
    {
        const int L = T.nlev - 1, cn = T.n[L], N = cn * cn * cn;
        double* A;
        if (T.nlev >= 2) {
            A = reinterpret_cast<double*>(R[0]);  // 4 m0^3 bytes >= 16 (m0/2)^3
        } else {
            A = pool + N;  // after b, d of the only level (2N floats = N doubles)
        }
        double* Bd = A + N;
        for (int x = tid; x < cn * cn; x += kTailThreads) Q[x] = Qf[x];
        for (int x = tid; x < N; x += kTailThreads) A[x] = (double)B[L][x];
        lds_barrier();
        coarse_exact(cn, Q, ilam, A, Bd, tid, kTailThreads);
        for (int x = tid; x < N; x += kTailThreads) D[L][x] = (float)A[x];
        lds_barrier();
    }
    // ---- up: d += trilinear(e), sweeps d <- d - c (b - A d) with c2 then c1 ----
    for (int l = T.nlev - 2; l >= 0; l--) {
        const int m = T.n[l], m2 = m * m, nc = T.n[l + 1];
        const float* b = B[l];
        float* d = D[l];
        float* t = R[l];
        const float* e = D[l + 1];
        const bool on = tj < m && tk < m;
        const TailCol col(on ? tj : 0, on ? tk : 0, m);
        if (on) {
            const int b0 = tj >> 1, b1 = (b0 + 1 == nc) ? 0 : b0 + 1;
            const int c0 = tk >> 1, c1 = (c0 + 1 == nc) ? 0 : c0 + 1;
            const float wj = (tj & 1) ? 0.5f : 0.0f, wk = (tk & 1) ? 0.5f : 0.0f;
            for (int i = ti; i < m; i += 4) {
                const int a0 = i >> 1, a1 = (a0 + 1 == nc) ? 0 : a0 + 1;
                const float wi = (i & 1) ? 0.5f : 0.0f;
                const float* e0 = e + a0 * nc * nc;
                const float* e1 = e + a1 * nc * nc;
                const float v00 = e0[b0 * nc + c0] + wk * (e0[b0 * nc + c1] - e0[b0 * nc + c0]);
                const float v01 = e0[b1 * nc + c0] + wk * (e0[b1 * nc + c1] - e0[b1 * nc + c0]);
                const float v10 = e1[b0 * nc + c0] + wk * (e1[b0 * nc + c1] - e1[b0 * nc + c0]);
                const float v11 = e1[b1 * nc + c0] + wk * (e1[b1 * nc + c1] - e1[b1 * nc + c0]);
                const float v0 = v00 + wj * (v01 - v00);
                const float v1 = v10 + wj * (v11 - v10);
                const int x = i * m2 + col.j0 + col.k0;
                d[x] = d[x] + (v0 + wi * (v1 - v0));
            }
        }
        lds_barrier();
        if (on)
            for (int i = ti; i < m; i += 4) {
                const int i0 = i * m2, im = (i == 0 ? m - 1 : i - 1) * m2,
                          ip = (i == m - 1 ? 0 : i + 1) * m2;
                const int x = i0 + col.j0 + col.k0;
                t[x] = d[x] - c2 * (b[x] - col.lap(d, im, i0, ip));
            }
        lds_barrier();
        if (on)
            for (int i = ti; i < m; i += 4) {
                const int i0 = i * m2, im = (i == 0 ? m - 1 : i - 1) * m2,
                          ip = (i == m - 1 ? 0 : i + 1) * m2;
                const int x = i0 + col.j0 + col.k0;
                d[x] = t[x] - c1 * (b[x] - col.lap(t, im, i0, ip));
            }
        lds_barrier();
    }
    {
        const int m = T.n[0], m2 = m * m;
        if (tj < m && tk < m)
            for (int i = ti; i < m; i += 4) {
                const int x = i * m2 + tj * m + tk;
                dout[x] = D[0][x];
            }
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

void launch_mg_tail(hipStream_t s, int nlev, const int* n, const float* b, float* d, double c1,
                    double c2, const double* Qf, const double* ilam, CGState* st) {
    TailLevels T;
    T.nlev = nlev;
    for (int l = 0; l < nlev; l++) T.n[l] = n[l];
    MAZE_LAUNCH(k_mg_tail, dim3(1), dim3(kTailThreads), 0, s, T, b, d, (float)c1, (float)c2, Qf,
                ilam, st);
}

// PCG direction p = -v (first) or beta p - v, with <p, A p>
// (library/src/plcg.c:130-155 form, v = M^-1 r).
// R: the storage type of p and v (float, or double for the FP64 multigrid-PCG);
// the value is rounded to it so that every thread sees the stored value.
template <class R, bool First>
struct SrcPcgDir {
    const R* p;
    const R* v;
    double beta;
    struct Raw { R p, v; };
    __device__ __forceinline__ Raw load(int i) const {
        if (First) return Raw{R(0), v[i]};
        return Raw{p[i], v[i]};
    }
    __device__ __forceinline__ double val(const Raw& x) const {
        const double w = First ? -(double)x.v : (double)x.p * beta - (double)x.v;
        return (double)(R)w;
    }
};

template <class R>
struct BodyPcgDir {
    R* p_new;
    double s = 0.0, m = 0.0;
    using Raw = NoRaw;
    __device__ Raw load(int) const { return Raw{}; }
    __device__ void apply(int c, const Raw&, double v, double lap) {
        p_new[c] = (R)v;
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

template <bool First, bool PB, class C = PBCoef>
__global__ __launch_bounds__(kBlock) void k_pcg_dir(GridDims d, C pb, const float* p_old,
                                                    const float* v, float* p_new, Reduce red,
                                                    CGState* st) {
    if (st->done) return;
    const SrcPcgDir<float, First> src{p_old, v, First ? 0.0 : st->beta};
    BodyPcgDir<float> bd{p_new};
    march<PB>(d, pb, src, bd);
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(bd.s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            BodyPcgDir<float>::finalize(st, T);
            reset_arrivals(red.counter);
        }
    }
}

void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red, int first, const PBCoef& pb) {
    DISPATCH_PB(pb, {
        if (first)
            MAZE_LAUNCH((k_pcg_dir<true, kPB>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb, p_old,
                        v, p_new, red, st);
        else
            MAZE_LAUNCH((k_pcg_dir<false, kPB>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb, p_old,
                        v, p_new, red, st);
    });
}

void launch_pcg_dir(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                    float* p_new, CGState* st, Reduce red, int first, const PBCoef32& pb) {
    if (first)
        MAZE_LAUNCH((k_pcg_dir<true, true, PBCoef32>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb,
                    p_old, v, p_new, red, st);
    else
        MAZE_LAUNCH((k_pcg_dir<false, true, PBCoef32>), dim3(d.blocks), dim3(kBlock), 0, s, d, pb,
                    p_old, v, p_new, red, st);
}

// Direction pass with deferred finalisation: beta = <r, z>_k / <r, z>_(k-1),
// both from partials (this iteration's from the finest up kernel, folded
// here; the previous one kept in rdv2 by parity), then p = -z (first) or
// beta p - z and the plain per-work-group partials of <p, A p>.  After the
// first iteration it also applies the previous update's stop test (its
// partials pus, pum, ucnt of them): a stop ends the solve here, before p
// changes (the V-cycle before it wrote only work arrays).  Work-group 0
// records the stop while later work-groups may be starting, so the flag of
// earlier kernels is thread 0's read at the start, handed to every wave by
// the fold's barrier; it is tested once the first plane's loads and the
// fold's loads are in flight, so no memory round trip precedes them.
template <bool First, class R = float>
__global__ __launch_bounds__(kBlock) void k_pcg_dir_df(GridDims d, const R* p_old,
                                                       const R* v, R* p_new,
                                                       const double* __restrict__ pz, int zcnt,
                                                       const double* __restrict__ pus,
                                                       const double* __restrict__ pum, int ucnt,
                                                       double* __restrict__ pp, int par,
                                                       CGState* st) {
    int done = vload_state(&st->done);  // thread 0's value counts
    const double rz_prev = First ? 1.0 : vload_state(&st->rdv2[par ^ 1]);
    __shared__ double sh[2 * kBlock / kWave];
    __shared__ double shf[3 * kBlock / kWave + 1];
    SrcPcgDir<R, First> src{p_old, v, 0.0};
    BodyPcgDir<R> bd{p_new};
    // The folds run while the first plane's loads are in flight.
    const bool go = march<false>(d, PBCoef{}, src, bd, [&] {
        double S, Su, Mu;
        fold_partials<kBlock, true, !First>(pz, zcnt, pus, pum, ucnt, shf, S, Su, Mu, done);
        if (done) return false;
        if (!First && df_decide(st, Su, Mu) != INT_MIN) return false;
        const double beta = First ? 0.0 : S / rz_prev;
        src.beta = beta;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->rdv2[par] = S;
            st->rdv = S;
            st->beta = beta;
            st->first = 0;
        }
        return true;
    });
    if (!go) return;
    double s = bd.s, m = 0.0;
    block_sum_max_lds<kBlock>(s, m, sh);
    if (threadIdx.x == 0) pp[blockIdx.x] = s;
}

// Update with deferred finalisation: alpha = <r, z>_k / <p, A p> (folded from
// the direction pass's partials while the first plane's loads are in flight),
// x += alpha p, r += alpha A p, and plain partials of <r, r> and max|r| for
// the consumer's stop test.  A vanishing <p, A p> stops the solve, as the
// breakdown guard of BodyDir does (alpha 0: x and r keep their values).
// Work-group 0 advances st->iter (no other work-group of this kernel reads it)
// and may record a breakdown while later work-groups start, so the stop flag
// is thread 0's read at the start, handed to every wave by the fold's barrier.
template <bool ZeroX, class R = float>
__global__ __launch_bounds__(kBlock) void k_cg_update_df(GridDims d, const R* p, R* x,
                                                         R* r, const double* __restrict__ pp,
                                                         double* __restrict__ pus,
                                                         double* __restrict__ pum, int par,
                                                         CGState* st) {
    int done = vload_state(&st->done);  // thread 0's value counts
    const double rz = vload_state(&st->rdv2[par]);
    __shared__ double sh[2 * kBlock / kWave];
    __shared__ double shf[3 * kBlock / kWave + 1];
    BodyUpdate<R, ZeroX> bd{x, r, 0.0};
    const bool go = march<false>(d, PBCoef{}, SrcArr<R>{p}, bd, [&] {
        double pAp, s0, m0;
        fold_partials<kBlock, true, false>(pp, d.blocks, nullptr, nullptr, 0, shf, pAp, s0, m0,
                                           done);
        if (done) return false;
        const bool breakdown = fabs(pAp) < 1e-300;
        const double alpha = breakdown ? 0.0 : rz / pAp;
        bd.alpha = alpha;
        if (blockIdx.x == 0 && threadIdx.x == 0) {
            st->pAp = pAp;
            st->alpha = alpha;
            if (breakdown) {
                st->res = st->iter;
                st->done = 1;
            } else {
                st->iter = st->iter + 1;
            }
        }
        return true;
    });
    if (!go) return;
    double s = bd.s, m = bd.m;
    block_sum_max_lds<kBlock>(s, m, sh);
    if (threadIdx.x == 0) {
        pus[blockIdx.x] = s;
        pum[blockIdx.x] = m;
    }
}

__global__ __launch_bounds__(kBlock) void k_df_finalize(CGState* st, const double* pus,
                                                        const double* pum, int cnt) {
    if (st->done) return;
    __shared__ double sh[2 * kBlock / kWave];
    (void)df_stop_test<kBlock>(st, pus, pum, cnt, sh);
}

template <class R>
static void pcg_dir_df_t(hipStream_t s, const GridDims& d, const R* p_old, const R* v, R* p_new,
                         CGState* st, const DeferRed& dr, int zcnt, int first, int par) {
    if (first)
        MAZE_LAUNCH((k_pcg_dir_df<true, R>), dim3(d.blocks), dim3(kBlock), 0, s, d, p_old, v,
                    p_new, dr.pz, zcnt, dr.pus, dr.pum, d.blocks, dr.pp, par, st);
    else
        MAZE_LAUNCH((k_pcg_dir_df<false, R>), dim3(d.blocks), dim3(kBlock), 0, s, d, p_old, v,
                    p_new, dr.pz, zcnt, dr.pus, dr.pum, d.blocks, dr.pp, par, st);
}

template <class R>
static void cg_update_df_t(hipStream_t s, const GridDims& d, const R* p, R* x, R* r, CGState* st,
                           const DeferRed& dr, int zero_x, int par) {
    if (zero_x)
        MAZE_LAUNCH((k_cg_update_df<true, R>), dim3(d.blocks), dim3(kBlock), 0, s, d, p, x, r,
                    dr.pp, dr.pus, dr.pum, par, st);
    else
        MAZE_LAUNCH((k_cg_update_df<false, R>), dim3(d.blocks), dim3(kBlock), 0, s, d, p, x, r,
                    dr.pp, dr.pus, dr.pum, par, st);
}

void launch_pcg_dir_df(hipStream_t s, const GridDims& d, const float* p_old, const float* v,
                       float* p_new, CGState* st, const DeferRed& dr, int zcnt, int first,
                       int par) {
    pcg_dir_df_t(s, d, p_old, v, p_new, st, dr, zcnt, first, par);
}

void launch_pcg_dir_df(hipStream_t s, const GridDims& d, const double* p_old, const double* v,
                       double* p_new, CGState* st, const DeferRed& dr, int zcnt, int first,
                       int par) {
    pcg_dir_df_t(s, d, p_old, v, p_new, st, dr, zcnt, first, par);
}

void launch_cg_update_df(hipStream_t s, const GridDims& d, const float* p, float* x, float* r,
                         CGState* st, const DeferRed& dr, int zero_x, int par) {
    cg_update_df_t(s, d, p, x, r, st, dr, zero_x, par);
}

void launch_cg_update_df(hipStream_t s, const GridDims& d, const double* p, double* x, double* r,
                         CGState* st, const DeferRed& dr, int zero_x, int par) {
    cg_update_df_t(s, d, p, x, r, st, dr, zero_x, par);
}

void launch_df_finalize(hipStream_t s, CGState* st, const DeferRed& dr, int cnt) {
    MAZE_LAUNCH(k_df_finalize, dim3(1), dim3(kBlock), 0, s, st, dr.pus, dr.pum, cnt);
}

// ---- FP32 multigrid for the linear Poisson-Boltzmann operator ----------------
// A u = sum_f eps_f (u_nb - u) - k2 u with FP32 face coefficients on every level
// (level 0: a copy of the maps) and an FP32 inverse diagonal g = 1 / (sum of the
// six face coefficients + k2).  Damped Jacobi: d <- d - w g (b - A d), with the
// Chebyshev weights of the constant-coefficient cycle (the Jacobi-scaled
// spectrum again lies in (0, 2)).  The passes are unfused march passes; the
// restriction (4 x full weighting) and the trilinear prolongation are the
// constant-coefficient kernels.  Coarse coefficients (Sim::pbmg_coefs): the
// two fine faces a coarse face spans combine in series (harmonic mean), the
// transverse fine faces by full weighting; k2 is fully weighted and scaled by
// (H/h)^2 = 4 like the residual.  The coarsest level (at most 16^3) runs a
// fixed number of Chebyshev iterations in one work-group, so the cycle is a
// fixed linear operator.  The PCG direction and update of the inner solve use
// the same FP32 coefficients (the FP64 refinement keeps the FP64 maps).

struct BodyPbSmooth {
    static constexpr bool kReduce = false, kMax = false, kSkipWhenDone = true;
    const float* b;
    const float* g;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    struct Raw { float b, g; };
    __device__ Raw load(int i) const { return Raw{b[i], g[i]}; }
    __device__ void apply(int i, const Raw& r, double v, double lap) {
        dout[i] = (float)(v - c * (double)r.g * ((double)r.b - lap));
    }
    __device__ static void finalize(CGState*, double, double) {}
};

// Last post-smoothing sweep of the finest level with <r, z> and the PCG beta.
struct BodyPbSmoothDot {
    static constexpr bool kReduce = true, kMax = false, kSkipWhenDone = true;
    const float* b;
    const float* g;
    float* dout;
    double c;
    double s = 0.0, m = 0.0;
    struct Raw { float b, g; };
    __device__ Raw load(int i) const { return Raw{b[i], g[i]}; }
    __device__ void apply(int i, const Raw& r, double v, double lap) {
        const float z = (float)(v - c * (double)r.g * ((double)r.b - lap));
        dout[i] = z;
        s += (double)r.b * (double)z;
    }
    __device__ static void finalize(CGState* st, double S, double) {
        st->beta = st->first ? 0.0 : S / st->rdv;
        st->rdv = S;
    }
};

void launch_pb_smooth(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* g,
                      const float* b, const float* din, float* dout, double c, CGState* st,
                      int dot, Reduce red) {
    if (dot)
        run_pass<true>(s, d, pb, SrcArr<float>{din}, BodyPbSmoothDot{b, g, dout, c}, red, st);
    else
        run_pass<true>(s, d, pb, SrcArr<float>{din}, BodyPbSmooth{b, g, dout, c}, Reduce{}, st);
}

// First sweep from zero evaluated on the fly: d1 = -c1 g b (rounded to FP32),
// so the two pre-smoothing sweeps are one pass: d2 = d1 - c2 g (b - A d1).
struct SrcPbFirst {
    const float* b;
    const float* g;
    double c;
    struct Raw { float b, g; };
    __device__ __forceinline__ Raw load(int i) const { return Raw{b[i], g[i]}; }
    __device__ __forceinline__ double val(const Raw& r) const {
        return (double)(float)(-c * (double)r.g * (double)r.b);
    }
};

void launch_pb_pre(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* g,
                   const float* b, float* dout, double c1, double c2, CGState* st) {
    run_pass<true>(s, d, pb, SrcPbFirst{b, g, c1}, BodyPbSmooth{b, g, dout, c2}, Reduce{}, st);
}

void launch_pb_resid(hipStream_t s, const GridDims& d, const PBCoef32& pb, const float* b,
                     const float* din, float* r, CGState* st) {
    run_pass<true>(s, d, pb, SrcArr<float>{din}, BodyMgResid{b, r}, Reduce{}, st);
}

// FP32 copies of the FP64 maps.
__global__ void k_pb_to32(long N, PBCoef in, float* __restrict__ ex, float* __restrict__ ey,
                          float* __restrict__ ez, float* __restrict__ k2) {
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        ex[t] = (float)in.ex[t];
        ey[t] = (float)in.ey[t];
        ez[t] = (float)in.ez[t];
        k2[t] = (float)in.k2[t];
    }
}

void launch_pb_to32(hipStream_t s, long count, const PBCoef& in, float* ex, float* ey, float* ez,
                    float* k2) {
    MAZE_LAUNCH(k_pb_to32, dim3(pw_blocks_mg(count)), dim3(kBlock), 0, s, count, in, ex, ey, ez, k2);
}

// g = 1 / (eps_x + eps_x(i-1) + eps_y + eps_y(j-1) + eps_z + eps_z(k-1) + k2).
__global__ void k_pb_idg(int n, PBCoef32 pb, float* __restrict__ g) {
    const long N = (long)n * n * n, n2 = (long)n * n;
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < N;
         t += (long)gridDim.x * blockDim.x) {
        const int k = (int)(t % n), j = (int)((t / n) % n), i = (int)(t / n2);
        const long im = (i == 0 ? n - 1 : i - 1) * n2 + (long)j * n + k;
        const long jm = (long)i * n2 + (j == 0 ? n - 1 : j - 1) * (long)n + k;
        const long km = (long)i * n2 + (long)j * n + (k == 0 ? n - 1 : k - 1);
        const double dg = (double)pb.ex[t] + pb.ex[im] + pb.ey[t] + pb.ey[jm] + pb.ez[t] +
                          pb.ez[km] + pb.k2[t];
        g[t] = (float)(1.0 / dg);
    }
}

void launch_pb_idg(hipStream_t s, int n, const PBCoef32& pb, float* g) {
    const long N = (long)n * n * n;
    MAZE_LAUNCH(k_pb_idg, dim3(pw_blocks_mg(N)), dim3(kBlock), 0, s, n, pb, g);
}

// Coarse coefficients of an (nf/2)^3 level from the nf^3 one.
__global__ void k_pb_coarsen(int nf, PBCoef32 f, float* __restrict__ cx, float* __restrict__ cy,
                             float* __restrict__ cz, float* __restrict__ ck2) {
    const int M = nf / 2;
    const long total = (long)M * M * M, N = nf, N2 = N * N;
    auto w = [nf](int i) { return i < 0 ? i + nf : (i >= nf ? i - nf : i); };
    auto harm = [](double a, double b) { return 2.0 * a * b / (a + b); };
    for (long t = blockIdx.x * (long)blockDim.x + threadIdx.x; t < total;
         t += (long)gridDim.x * blockDim.x) {
        const int K = (int)(t % M), J = (int)((t / M) % M), I = (int)(t / ((long)M * M));
        const int i = 2 * I, j = 2 * J, k = 2 * K;
        double sx = 0.0, sy = 0.0, sz = 0.0, s2 = 0.0;
        for (int a = -1; a <= 1; a++) {
            const double wa = a == 0 ? 2.0 : 1.0;
            for (int b = -1; b <= 1; b++) {
                const double wb = b == 0 ? 2.0 : 1.0;
                // x faces at (i, j + a, k + b) and (i + 1, j + a, k + b)
                const long xa = (long)i * N2 + w(j + a) * N + w(k + b);
                sx += wa * wb * harm(f.ex[xa], f.ex[xa + N2 * (i + 1 < nf ? 1 : 1 - nf)]);
                // y faces at (i + a, j, k + b) and (i + a, j + 1, k + b)
                const long ya = (long)w(i + a) * N2 + (long)j * N + w(k + b);
                sy += wa * wb * harm(f.ey[ya], f.ey[ya + N * (j + 1 < nf ? 1 : 1 - nf)]);
                // z faces at (i + a, j + b, k) and (i + a, j + b, k + 1)
                const long za = (long)w(i + a) * N2 + w(j + b) * N + k;
                sz += wa * wb * harm(f.ez[za], f.ez[za + (k + 1 < nf ? 1 : 1 - nf)]);
                for (int c = -1; c <= 1; c++) {
                    const double wc = c == 0 ? 2.0 : 1.0;
                    s2 += wa * wb * wc * f.k2[(long)w(i + a) * N2 + w(j + b) * N + w(k + c)];
                }
            }
        }
        cx[t] = (float)(sx / 16.0);
        cy[t] = (float)(sy / 16.0);
        cz[t] = (float)(sz / 16.0);
        ck2[t] = (float)(4.0 * s2 / 64.0);
    }
}

void launch_pb_coarsen(hipStream_t s, int nf, const PBCoef32& f, float* cx, float* cy, float* cz,
                       float* ck2) {
    const long M = nf / 2;
    MAZE_LAUNCH(k_pb_coarsen, dim3(pw_blocks_mg(M * M * M)), dim3(kBlock), 0, s, nf, f, cx, cy,
                cz, ck2);
}

// Coarsest level (c^3 <= kMgCoarseLds^3): `iters` Chebyshev iterations for
// A x = b from x = 0, Jacobi-preconditioned, for the spectrum [lo, hi] of
// g (-A) (Saad, Iterative Methods for Sparse Linear Systems, Alg. 12.1), in
// one work-group with the level in LDS; the update direction alternates
// between two arrays so that one barrier per iteration suffices.
constexpr int kPbCoarseThreads = 1024;

// The iterations on a c^3 level (c^3 <= CMAX) by NT threads: coefficients pb
// and g, right-hand side b (global or LDS), result x_out (global or LDS); dv
// (2 CMAX) and ex, ey, ez, k2 (CMAX each) are the caller's LDS.
template <int CMAX, int NT>
__device__ __forceinline__ void pb_cheb(int c, const PBCoef32& pb, const float* g, const float* b,
                                        float* x_out, int iters, double lo, double hi,
                                        float (*dv)[CMAX], float* ex, float* ey, float* ez,
                                        float* k2) {
    constexpr int PER = (CMAX + NT - 1) / NT;
    const int N = c * c * c, c2 = c * c, tid = threadIdx.x;
    float x[PER], r[PER], gg[PER];
    int nb[PER][6];
    const double theta = 0.5 * (hi + lo), delta = 0.5 * (hi - lo), sigma = theta / delta;
    double rho = 1.0 / sigma;
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + q * NT;
        if (t < N) {
            ex[t] = pb.ex[t];
            ey[t] = pb.ey[t];
            ez[t] = pb.ez[t];
            k2[t] = pb.k2[t];
            const int k = t % c, j = (t / c) % c, i = t / c2;
            nb[q][0] = (i == c - 1 ? 0 : i + 1) * c2 + j * c + k;
            nb[q][1] = (i == 0 ? c - 1 : i - 1) * c2 + j * c + k;
            nb[q][2] = i * c2 + (j == c - 1 ? 0 : j + 1) * c + k;
            nb[q][3] = i * c2 + (j == 0 ? c - 1 : j - 1) * c + k;
            nb[q][4] = i * c2 + j * c + (k == c - 1 ? 0 : k + 1);
            nb[q][5] = i * c2 + j * c + (k == 0 ? c - 1 : k - 1);
            x[q] = 0.0f;
            r[q] = -b[t];  // residual of (-A) x = -b at x = 0
            gg[q] = g[t];
            dv[0][t] = (float)(gg[q] * r[q] / theta);
        }
    }
    lds_barrier();
    int cur = 0;
    for (int it = 0; it < iters; it++) {
        const double rho1 = 1.0 / (2.0 * sigma - rho);
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + q * NT;
            if (t < N) {
                const float* d = dv[cur];
                const float dc = d[t];
                // (-A) d = sum_f eps_f (d - d_nb) + k2 d
                const float Ad = ex[t] * (dc - d[nb[q][0]]) + ex[nb[q][1]] * (dc - d[nb[q][1]]) +
                                 ey[t] * (dc - d[nb[q][2]]) + ey[nb[q][3]] * (dc - d[nb[q][3]]) +
                                 ez[t] * (dc - d[nb[q][4]]) + ez[nb[q][5]] * (dc - d[nb[q][5]]) +
                                 k2[t] * dc;
                x[q] += dc;
                r[q] -= Ad;
                dv[cur ^ 1][t] = (float)(rho1 * rho * dc + 2.0 * rho1 / delta * gg[q] * r[q]);
            }
        }
        rho = rho1;
        cur ^= 1;
        lds_barrier();
    }
#pragma unroll
    for (int q = 0; q < PER; q++) {
        const int t = tid + q * NT;
        if (t < N) x_out[t] = x[q];
    }
}

__global__ __launch_bounds__(kPbCoarseThreads) void k_pb_coarse(int c, PBCoef32 pb,
                                                             const float* __restrict__ g,
                                                             const float* __restrict__ b,
                                                             float* __restrict__ x_out, int iters,
                                                             double lo, double hi,
                                                             const CGState* st) {
    if (st->done) return;
    constexpr int CMAX = kMgCoarseLds * kMgCoarseLds * kMgCoarseLds;
    __shared__ float dv[2][CMAX];
    __shared__ float ex[CMAX], ey[CMAX], ez[CMAX], k2[CMAX];
    pb_cheb<CMAX, kPbCoarseThreads>(c, pb, g, b, x_out, iters, lo, hi, dv, ex, ey, ez, k2);
}

// ---- the two smallest PB levels in one work-group --------------------------
// The level above the coarsest (n <= kPbTailMaxN) and the coarsest (n / 2)
// with the level in LDS: the passes of Sim::vcycle_pb for that level --
// pre-smoothing (both sweeps, the first from zero), residual, restriction,
// the coarsest level's iterations, prolongation-correction and the two
// post-smoothing sweeps -- each one the same arithmetic, in the same order,
// as the separate kernel it replaces (k_pass with SrcPbFirst + BodyPbSmooth,
// BodyMgResid, k_mg_restrict, k_pb_coarse, k_mg_prolong_add, BodyPbSmooth),
// so the result is bit-identical to them, in one launch instead of seven.
constexpr int kPbTailF = kPbTailMaxN * kPbTailMaxN * kPbTailMaxN;
constexpr int kPbTailC = kPbTailF / 8;

// The march's PB operator at point (i, j, k) of an n^3 level for the values u
// (LDS): sum_f eps_f (u_nb - u) - k2 u, in its order and in FP64.
__device__ __forceinline__ double pb_lap_tail(const float* u, const PBCoef32& pb, int n, int i,
                                              int j, int k) {
    const int n2 = n * n, t = i * n2 + j * n + k;
    const int ip = i == n - 1 ? 0 : i + 1, im = i == 0 ? n - 1 : i - 1;
    const int jp = j == n - 1 ? 0 : j + 1, jm = j == 0 ? n - 1 : j - 1;
    const int kp = k == n - 1 ? 0 : k + 1, km = k == 0 ? n - 1 : k - 1;
    const double vp = u[ip * n2 + j * n + k], vm = u[im * n2 + j * n + k];
    const double vjp = u[i * n2 + jp * n + k], vjm = u[i * n2 + jm * n + k];
    const double vkp = u[i * n2 + j * n + kp], vkm = u[i * n2 + j * n + km];
    const double vc = u[t];
    const double exc = pb.ex[t], exm = pb.ex[im * n2 + j * n + k];
    const double eyc = pb.ey[t], eym = pb.ey[i * n2 + jm * n + k];
    const double ezc = pb.ez[t], ezm = pb.ez[i * n2 + j * n + km];
    const double k2c = pb.k2[t];
    return vp * exc + vm * exm + vjp * eyc + vjm * eym + vkp * ezc + vkm * ezm -
           vc * (exc + exm + eyc + eym + ezc + ezm + k2c);
}

__global__ __launch_bounds__(kPbCoarseThreads) void k_pb_tail(int n, PBCoef32 pf,
                                                           const float* __restrict__ gf,
                                                           const float* __restrict__ bin,
                                                           float* __restrict__ dout, PBCoef32 pc,
                                                           const float* __restrict__ gc, double c1,
                                                           double c2, int iters, double lo,
                                                           double hi, const CGState* st) {
    if (st->done) return;
    constexpr int NT = kPbCoarseThreads;
    __shared__ float B[kPbTailF], D[kPbTailF], T[kPbTailF];
    __shared__ float cb[kPbTailC], ce[kPbTailC];
    __shared__ float dv[2][kPbTailC];
    __shared__ float cex[kPbTailC], cey[kPbTailC], cez[kPbTailC], ck2[kPbTailC];
    const int tid = threadIdx.x, n2 = n * n, N = n2 * n, c = n / 2, C = c * c * c;
    // A thread's points t = tid + q NT, q < PER, unrolled so that the loads of
    // all of them are in flight together.
    constexpr int PER = (kPbTailF + NT - 1) / NT;
    auto each = [&](auto&& f) {
#pragma unroll
        for (int q = 0; q < PER; q++) {
            const int t = tid + q * NT;
            if (t < N) f(t, t / n2, (t / n) % n, t % n);
        }
    };
    // First pre-smoothing sweep from zero, as SrcPbFirst evaluates it.
    each([&](int t, int, int, int) {
        const float bv = bin[t];
        B[t] = bv;
        T[t] = (float)(-c1 * (double)gf[t] * (double)bv);
    });
    lds_barrier();
    each([&](int t, int i, int j, int k) {  // second sweep (BodyPbSmooth, weight c2)
        const double lap = pb_lap_tail(T, pf, n, i, j, k);
        D[t] = (float)((double)T[t] - c2 * (double)gf[t] * ((double)B[t] - lap));
    });
    lds_barrier();
    each([&](int t, int i, int j, int k) {  // residual (BodyMgResid)
        T[t] = (float)((double)B[t] - pb_lap_tail(D, pf, n, i, j, k));
    });
    lds_barrier();
    for (int u = tid; u < C; u += NT) {  // restriction (k_mg_restrict)
        const int K = u % c, J = (u / c) % c, I = u / (c * c);
        const int i = 2 * I, j = 2 * J, k = 2 * K;
        const int ii[3] = {(i == 0 ? n - 1 : i - 1), i, i + 1};
        const int jj[3] = {(j == 0 ? n - 1 : j - 1), j, j + 1};
        const int kk[3] = {(k == 0 ? n - 1 : k - 1), k, k + 1};
        double acc = 0.0;
#pragma unroll
        for (int a = 0; a < 3; a++)
#pragma unroll
            for (int bb = 0; bb < 3; bb++)
#pragma unroll
                for (int cc = 0; cc < 3; cc++) {
                    const double w =
                        (double)((a == 1 ? 2 : 1) * (bb == 1 ? 2 : 1) * (cc == 1 ? 2 : 1));
                    acc += w * (double)T[ii[a] * n2 + jj[bb] * n + kk[cc]];
                }
        cb[u] = (float)(acc * (4.0 / 64.0));
    }
    lds_barrier();
    pb_cheb<kPbTailC, NT>(c, pc, gc, cb, ce, iters, lo, hi, dv, cex, cey, cez, ck2);
    lds_barrier();
    each([&](int t, int i, int j, int k) {  // prolongation-correction (k_mg_prolong_add)
        const int a0 = i >> 1, a1 = (a0 + 1 == c) ? 0 : a0 + 1;
        const int b0 = j >> 1, b1 = (b0 + 1 == c) ? 0 : b0 + 1;
        const int c0 = k >> 1, cc1 = (c0 + 1 == c) ? 0 : c0 + 1;
        const double wi = (i & 1) ? 0.5 : 0.0, wj = (j & 1) ? 0.5 : 0.0, wk = (k & 1) ? 0.5 : 0.0;
        const int M = c, M2 = c * c;
        const float e000 = ce[a0 * M2 + b0 * M + c0], e001 = ce[a0 * M2 + b0 * M + cc1];
        const float e010 = ce[a0 * M2 + b1 * M + c0], e011 = ce[a0 * M2 + b1 * M + cc1];
        const float e100 = ce[a1 * M2 + b0 * M + c0], e101 = ce[a1 * M2 + b0 * M + cc1];
        const float e110 = ce[a1 * M2 + b1 * M + c0], e111 = ce[a1 * M2 + b1 * M + cc1];
        const float dt = D[t];
        const double v00 = e000 + wk * ((double)e001 - e000);
        const double v01 = e010 + wk * ((double)e011 - e010);
        const double v10 = e100 + wk * ((double)e101 - e100);
        const double v11 = e110 + wk * ((double)e111 - e110);
        const double v0 = v00 + wj * (v01 - v00);
        const double v1 = v10 + wj * (v11 - v10);
        D[t] = (float)((double)dt + v0 + wi * (v1 - v0));
    });
    lds_barrier();
    each([&](int t, int i, int j, int k) {  // post-smoothing, weight c2
        const double lap = pb_lap_tail(D, pf, n, i, j, k);
        T[t] = (float)((double)D[t] - c2 * (double)gf[t] * ((double)B[t] - lap));
    });
    lds_barrier();
    each([&](int t, int i, int j, int k) {  // post-smoothing, weight c1
        const double lap = pb_lap_tail(T, pf, n, i, j, k);
        dout[t] = (float)((double)T[t] - c1 * (double)gf[t] * ((double)B[t] - lap));
    });
}

void launch_pb_tail(hipStream_t s, int n, const PBCoef32& pf, const float* gf, const float* b,
                    float* d, const PBCoef32& pc, const float* gc, double c1, double c2, int iters,
                    double lo, double hi, CGState* st) {
    if (n > kPbTailMaxN || n % 2 != 0) {
        std::fprintf(stderr, "maze: PB tail level %d above %d or odd\n", n, kPbTailMaxN);
        std::exit(1);
    }
    MAZE_LAUNCH(k_pb_tail, dim3(1), dim3(kPbCoarseThreads), 0, s, n, pf, gf, b, d, pc, gc, c1, c2,
                iters, lo, hi, st);
}

void launch_pb_coarse(hipStream_t s, int c, const PBCoef32& pb, const float* g, const float* b,
                      float* x, int iters, double lo, double hi, CGState* st) {
    if (c > kMgCoarseLds) {
        std::fprintf(stderr, "maze: PB multigrid coarsest level %d above %d\n", c, kMgCoarseLds);
        std::exit(1);
    }
    MAZE_LAUNCH(k_pb_coarse, dim3(1), dim3(kPbCoarseThreads), 0, s, c, pb, g, b, x, iters, lo, hi,
                st);
}

}  // namespace maze

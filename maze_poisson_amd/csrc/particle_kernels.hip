// particle_kernels.hip -- charge assignment, forces and integrators for gfx950.
//
// Point-wise expressions follow the reference's operation order
// (library/src/charges.c:8-14, :123-141; library/src/forces.c:52-94,
// :132-183; library/src/integrator_verlet.c:37-70;
// library/src/integrator_ovrvo.c:45-123) so that per-ion results match the
// CPU oracle bit for bit wherever no transcendental function is involved.

#include "particles.hpp"

namespace maze {

// CIC weight with the periodic distance (library/src/charges.c:8-14).
__device__ __forceinline__ double cic_w(double x, double L, double h) {
    x = fabs(x - round(x / L) * L);
    if (x >= h) return 0.0;
    return 1.0 - x / h;
}

__constant__ int kCorner[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};

static int blocks_for(long n) {
    long b = (n + kBlock - 1) / kBlock;
    return (int)(b < 1 ? 1 : (b > 65535 ? 65535 : b));
}

// One thread per item, no grid-stride loop: every item gets its own thread.
static int blocks_exact(long n) {
    const long b = (n + kBlock - 1) / kBlock;
    if (b > 0x7fffffffL) {
        std::fprintf(stderr, "maze: %ld items exceed one launch\n", n);
        std::exit(1);
    }
    return (int)(b < 1 ? 1 : b);
}

// Base cell floor(pos/h), wrapped into [0, n) (library/src/particles.c:315-321).
__global__ void k_cic_base(int np, int n, double h, const double* __restrict__ pos,
                           int* __restrict__ base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
#pragma unroll
    for (int d = 0; d < 3; d++) {
        int c = (int)floor(pos[3 * i + d] / h);
        c %= n;
        if (c < 0) c += n;
        base[3 * i + d] = c;
    }
}

void launch_cic_base(hipStream_t s, const PDev& p) {
    if (p.np == 0) return;
    MAZE_LAUNCH(k_cic_base, dim3(blocks_for(p.np)), dim3(kBlock), 0, s, p.np, p.n, p.h, p.pos,
                p.base);
}

// Quadratic and cubic B-spline weights (library/src/charges.c:16-38).
__device__ __forceinline__ double spline_quadr_w(double x, double L, double h) {
    x = fabs(x - round(x / L) * L) / h;
    if (x < 0.5) return 0.75 - x * x;
    if (x < 1.5) {
        const double y = 1.5 - x;
        return 0.5 * (y * y);
    }
    return 0.0;
}

__device__ __forceinline__ double spline_cubic_w(double x, double L, double h) {
    x = fabs(x - round(x / L) * L) / h;
    if (x < 1) return (4.0 - 6.0 * (x * x) + 3.0 * cube(x)) / 6.0;
    if (x < 2) return cube(2 - x) / 6.0;
    return 0.0;
}

// Charge-assignment schemes (library/src/particles.c:38-66): CIC on the 8
// corners of the base cell in the reference's order (:297-346), splines on
// the 4x4x4 block base + {-1..2}^3 with the last offset fastest (:348-385).
template <int CAS>
struct Assign;

template <>
struct Assign<0> {
    static constexpr int kNodes = 8, kLo = 0, kSpan = 2;  // node - base in [kLo, kLo + kSpan)
    __device__ static int off(int c, int d) { return kCorner[c][d]; }
    __device__ static double w(double x, double L, double h) { return cic_w(x, L, h); }
};

template <int CAS>
struct AssignSpline {
    static constexpr int kNodes = 64, kLo = -1, kSpan = 4;
    __device__ static int off(int c, int d) {
        return (d == 0 ? c >> 4 : (d == 1 ? (c >> 2) & 3 : c & 3)) - 1;
    }
    __device__ static double w(double x, double L, double h) {
        return CAS == 1 ? spline_quadr_w(x, L, h) : spline_cubic_w(x, L, h);
    }
};
template <>
struct Assign<1> : AssignSpline<1> {};
template <>
struct Assign<2> : AssignSpline<2> {};

// Contribution of ion j to grid node (gi, gj, gk): q_j * wx * wy * wz
// (library/src/charges.c:136-137).
template <class A>
__device__ __forceinline__ double assign_contrib(const double* pos, const double* charges, int j,
                                                 int gi, int gj, int gk, double L, double h) {
    const double app = A::w(pos[3 * j] - gi * h, L, h) * A::w(pos[3 * j + 1] - gj * h, L, h) *
                       A::w(pos[3 * j + 2] - gk * h, L, h);
    return charges[j] * app;
}

// The charge spread without float atomics, in two passes over the (ion i,
// node c) contributions t = i nn + c -- the order of the reference's serial
// scatter, ion by ion and node by node (library/src/charges.c:123-141).
//
// link: thread t computes its contribution, clears the node q held from the
// previous spread (zero_old), and pushes itself onto its node's list with
// one 64-bit exchange (the tag is the spread's epoch, so lists of earlier
// spreads read as empty without clearing the head array).
// own:  the thread at the head of a node's list -- exactly one per touched
// node -- walks the list, sums the contributions in increasing t starting
// from 0.0 as the reference's += does, and writes q once.  The exchange
// order only decides which thread sums; the sum itself is in t order, so q
// equals the reference's serial result bit for bit.  Lists are short (a CIC
// node is reached from the 8 cells around it, about 0.02 ions at melt
// density); the selection below costs (length)^2 loads.
__device__ __forceinline__ unsigned long long spread_tag(const PScalars* ps) {
    return (unsigned long long)(ps->epoch + 1u) << 32;
}

template <int CAS>
__global__ void k_assign_link(int np, int n, double h, const double* __restrict__ pos,
                              const double* __restrict__ charges, const int* __restrict__ base,
                              const PScalars* __restrict__ ps, double* __restrict__ contrib,
                              int* __restrict__ node, int* __restrict__ link,
                              unsigned long long* __restrict__ head, double* __restrict__ q,
                              int zero_old) {
    using A = Assign<CAS>;
    constexpr int NN = A::kNodes;
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= (long)NN * np) return;
    const int i = (int)(t / NN), c = (int)(t % NN);
    const double L = n * h;
    auto wrap = [n](int v) { return v < 0 ? v + n : (v >= n ? v - n : v); };
    const int gi = wrap(base[3 * i] + A::off(c, 0));
    const int gj = wrap(base[3 * i + 1] + A::off(c, 1));
    const int gk = wrap(base[3 * i + 2] + A::off(c, 2));
    contrib[t] = assign_contrib<A>(pos, charges, i, gi, gj, gk, L, h);
    const int g = (gi * n + gj) * n + gk;
    if (zero_old) {
        const int old = node[t];
        q[old] = 0.0;
    }
    node[t] = g;
    const unsigned long long tag = spread_tag(ps);
    const unsigned long long prev = atomicExch(&head[g], tag | (unsigned int)t);
    link[t] = (prev & 0xffffffff00000000ull) == tag ? (int)(unsigned int)prev : -1;
}

__global__ void k_assign_own(long total, const PScalars* __restrict__ ps,
                             const double* __restrict__ contrib, const int* __restrict__ node,
                             const int* __restrict__ link,
                             const unsigned long long* __restrict__ head, double* __restrict__ q) {
    const long t = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (t >= total) return;
    const int g = node[t];
    if (head[g] != (spread_tag(ps) | (unsigned int)t)) return;
    double acc = 0.0;
    int last = -1;
    for (;;) {
        int next = 0x7fffffff;
        for (int u = (int)t; u >= 0; u = link[u])
            if (u > last && u < next) next = u;
        if (next == 0x7fffffff) break;
        acc += contrib[next];
        last = next;
    }
    q[g] = acc;
}

void launch_assign(hipStream_t s, const PDev& p, const PScalars* ps, double* q, int zero_old) {
    if (p.np == 0) return;
    const long total = (long)p.nn * p.np;
    const int nb = blocks_exact(total);
    if (p.cas == 0)
        MAZE_LAUNCH(k_assign_link<0>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, p.pos,
                    p.charges, p.base, ps, p.contrib, p.node, p.link, p.head, q, zero_old);
    else if (p.cas == 1)
        MAZE_LAUNCH(k_assign_link<1>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, p.pos,
                    p.charges, p.base, ps, p.contrib, p.node, p.link, p.head, q, zero_old);
    else
        MAZE_LAUNCH(k_assign_link<2>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, p.pos,
                    p.charges, p.base, ps, p.contrib, p.node, p.link, p.head, q, zero_old);
    MAZE_LAUNCH(k_assign_own, dim3(nb), dim3(kBlock), 0, s, total, ps, p.contrib, p.node, p.link,
                p.head, q);
}

// Sum of the contributions (the reference's charge-conservation total) and
// the end of the spread's epoch.
__global__ void k_contrib_sum(long total, const double* __restrict__ v, Reduce red,
                              PScalars* ps) {
    double s = 0.0;
    for (long i = blockIdx.x * (long)blockDim.x + threadIdx.x; i < total;
         i += (long)gridDim.x * blockDim.x)
        s += v[i];
    __shared__ double sh[kBlock / kWave];
    const double S = block_sum(s, sh);
    if (publish_arrive(red.psum, S, nullptr, 0.0, red.counter, gridDim.x)) {
        const double T = reduce_partials_sum(red.psum, gridDim.x, sh);
        if (threadIdx.x == 0) {
            ps->q_sum = T;
            ps->epoch += 1u;
            reset_arrivals(red.counter);
        }
    }
}

void launch_contrib_sum(hipStream_t s, const PDev& p, PScalars* ps, Reduce red) {
    const long total = (long)p.nn * p.np;
    long b = (total + 4L * kBlock - 1) / (4L * kBlock);
    b = b < 1 ? 1 : (b > 2048 ? 2048 : b);
    MAZE_LAUNCH(k_contrib_sum, dim3((int)b), dim3(kBlock), 0, s, total, p.contrib, red, ps);
}

// Field force -q w grad(phi) by central differences at the assignment nodes
// (library/src/forces.c:52-94).  One thread per (ion, node): each forms its
// node's three terms q w E, and the node-0 thread of the ion adds them in node
// order starting from 0.0 -- the reference's loop order, so the force matches
// its serial sum bit for bit.  A 256-thread work-group holds 256 / nn whole
// ions (nn = 8 or 64 divides 256).
// Kick: the MD step's path also forms the total force as k_ftot<false> does
// and the velocity half-kick v += cdt F_tot / m as k_kick does, value for
// value (one launch instead of three).
struct ForceKick {
    const double* fn;  // non-electrostatic force
    double* ft;        // total force
    double* vel;
    const double* mass;
    double cdt;
    const int* gate = nullptr;  // if set: the launch does nothing while *gate is 0
};

template <int CAS, bool Kick = false>
__global__ __launch_bounds__(kBlock) void k_force_fd(int np, int n, double h,
                                                     const double* __restrict__ phi,
                                                     const int* __restrict__ base,
                                                     const double* __restrict__ charges,
                                                     const double* __restrict__ pos,
                                                     double* __restrict__ forces,
                                                     ForceKick fk = ForceKick{}) {
    using A = Assign<CAS>;
    constexpr int NN = A::kNodes;
    static_assert(kBlock % NN == 0, "whole ions per work-group");
    if (Kick && fk.gate && *fk.gate == 0) return;
    __shared__ double term[3][kBlock];
    const long t = blockIdx.x * (long)kBlock + threadIdx.x;
    const int ip = (int)(t / NN), c = (int)(t % NN);
    if (ip < np) {
        const long N = n, N2 = N * N;
        const double h2 = 2.0 * h, L = N * h;
        const long i = (base[3 * ip] + A::off(c, 0) + n) % n;
        const long j = (base[3 * ip + 1] + A::off(c, 1) + n) % n;
        const long k = (base[3 * ip + 2] + A::off(c, 2) + n) % n;
        const long im = (i == 0 ? N - 1 : i - 1), ii = (i == N - 1 ? 0 : i + 1);
        const long jm = (j == 0 ? N - 1 : j - 1), jj = (j == N - 1 ? 0 : j + 1);
        const long km = (k == 0 ? N - 1 : k - 1), kk = (k == N - 1 ? 0 : k + 1);
        const double pxm = phi[im * N2 + j * N + k], pxp = phi[ii * N2 + j * N + k];
        const double pym = phi[i * N2 + jm * N + k], pyp = phi[i * N2 + jj * N + k];
        const double pzm = phi[i * N2 + j * N + km], pzp = phi[i * N2 + j * N + kk];
        const double qc = charges[ip] * A::w(pos[3 * ip] - i * h, L, h) *
                          A::w(pos[3 * ip + 1] - j * h, L, h) * A::w(pos[3 * ip + 2] - k * h, L, h);
        term[0][threadIdx.x] = qc * ((pxm - pxp) / h2);
        term[1][threadIdx.x] = qc * ((pym - pyp) / h2);
        term[2][threadIdx.x] = qc * ((pzm - pzp) / h2);
    }
    __syncthreads();
    if (ip < np && c == 0) {
        double f0 = 0.0, f1 = 0.0, f2 = 0.0;
        for (int u = 0; u < NN; u++) {
            f0 += term[0][threadIdx.x + u];
            f1 += term[1][threadIdx.x + u];
            f2 += term[2][threadIdx.x + u];
        }
        forces[3 * ip] = f0;
        forces[3 * ip + 1] = f1;
        forces[3 * ip + 2] = f2;
        if (Kick) {
            const double fe[3] = {f0, f1, f2};
#pragma unroll
            for (int d = 0; d < 3; d++) {
                double v = 0.0;
                v += 1.0 * fe[d];
                v += 1.0 * fk.fn[3 * ip + d];
                fk.ft[3 * ip + d] = v;
                fk.vel[3 * ip + d] += fk.cdt * v / fk.mass[ip];
            }
        }
    }
}

void launch_force_fd(hipStream_t s, const PDev& p, const double* phi) {
    if (p.np == 0) return;
    const int nb = blocks_exact((long)p.nn * p.np);
    if (p.cas == 0)
        MAZE_LAUNCH(k_force_fd<0>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi, p.base,
                    p.charges, p.pos, p.fe, ForceKick{});
    else if (p.cas == 1)
        MAZE_LAUNCH(k_force_fd<1>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi, p.base,
                    p.charges, p.pos, p.fe, ForceKick{});
    else
        MAZE_LAUNCH(k_force_fd<2>, dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi, p.base,
                    p.charges, p.pos, p.fe, ForceKick{});
}

void launch_force_fd_kick(hipStream_t s, const PDev& p, const double* phi, double cdt,
                          const int* gate) {
    if (p.np == 0) return;
    const int nb = blocks_exact((long)p.nn * p.np);
    const ForceKick fk{p.fn, p.ft, p.vel, p.mass, cdt, gate};
    if (p.cas == 0)
        MAZE_LAUNCH((k_force_fd<0, true>), dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi,
                    p.base, p.charges, p.pos, p.fe, fk);
    else if (p.cas == 1)
        MAZE_LAUNCH((k_force_fd<1, true>), dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi,
                    p.base, p.charges, p.pos, p.fe, fk);
    else
        MAZE_LAUNCH((k_force_fd<2, true>), dim3(nb), dim3(kBlock), 0, s, p.np, p.n, p.h, phi,
                    p.base, p.charges, p.pos, p.fe, fk);
}

// Shifted-force Tosi-Fumi over all pairs within r_cut (library/src/forces.c:
// 132-183), each unordered pair evaluated once (Newton's third law).
//
// Ions are split into blocks of 64; one work-group of 4 waves handles a block
// pair (I, J), I <= J.  Lane l owns ion i = 64 I + l and at step s meets
// j = 64 J + (l + s) mod 64, so the 64 lanes of a step meet 64 distinct j;
// wave w takes the steps s = w, w + 4, ...  The force on i stays in
// registers; the reaction on j goes to the wave's own LDS accumulator of j
// (no two lanes of a wave touch the same j in a step; a barrier separates
// steps).  The diagonal pair (I, I) takes steps 1..32 (step 32 on lanes < 32
// only) so that every pair inside the block is met once.  The waves' partials
// are summed in wave order and the work-group writes the i-role and j-role
// partials (force, energy, count) of its block pair; k_tf_fold sums each
// ion's partials in a fixed order (deterministic).  The pair energy goes to
// both ions, so the potential energy is half the sum over ions, as in the
// reference.
constexpr int kTfB = 64;
constexpr int kTfWaves = 4;

__device__ __forceinline__ int tf_item(int I, int J, int nb) { return I * nb - I * (I - 1) / 2 + (J - I); }

// The type-pair table sits in LDS for up to kTfLdsTypes types, else it is
// read from global memory.
constexpr int kTfLdsTypes = 4;

template <bool LdsTab>
__global__ __launch_bounds__(kTfB * kTfWaves) void k_tf_tiles(
    int np, int n_typ, double L, double invL, double r_cut, const double* __restrict__ pos,
    const int* __restrict__ types, const double* __restrict__ table, double* __restrict__ part,
    int i_off) {
    const int J = blockIdx.x, I = blockIdx.y + i_off, nb = gridDim.x;
    if (I > J) return;
    __shared__ double sx[kTfB], sy[kTfB], sz[kTfB];
    __shared__ int styp[kTfB];
    __shared__ double acc[kTfWaves][5][kTfB];  // j-role, per wave
    __shared__ double own[kTfWaves][5][kTfB];  // i-role, per wave
    __shared__ double tab_lds[kTfLdsTypes * kTfLdsTypes * 7];
    const int tid = threadIdx.x, l = tid & (kTfB - 1), w = tid / kTfB;
    if (LdsTab)
        for (int t = tid; t < n_typ * n_typ * 7; t += kTfB * kTfWaves) tab_lds[t] = table[t];
    const double* tab = LdsTab ? tab_lds : table;
    if (w == 0) {
        const int jl = J * kTfB + l;
        sx[l] = jl < np ? pos[3 * jl] : 0.0;
        sy[l] = jl < np ? pos[3 * jl + 1] : 0.0;
        sz[l] = jl < np ? pos[3 * jl + 2] : 0.0;
        styp[l] = jl < np ? types[jl] : 0;
    }
    for (int c = 0; c < 5; c++) acc[w][c][l] = 0.0;
    const int i = I * kTfB + l;
    const bool iv = i < np;
    const double xi = iv ? pos[3 * i] : 0.0, yi = iv ? pos[3 * i + 1] : 0.0,
                 zi = iv ? pos[3 * i + 2] : 0.0;
    const int ti = iv ? types[i] : 0;
    __syncthreads();
    const bool diag = I == J;
    const int s0 = diag ? 1 : 0, s1 = diag ? 32 : 63;
    double f0 = 0.0, f1 = 0.0, f2 = 0.0, e = 0.0, cnt = 0.0;
    // Waves share only the read-only j block; each accumulates reactions in its
    // own LDS rows, and within a wave the lanes of one step touch distinct j and
    // successive steps are in program order, so the steps need no barrier.
    const int nsteps = (s1 - s0 + kTfWaves) / kTfWaves;
    for (int k = 0; k < nsteps; k++) {
        const int s = s0 + w + k * kTfWaves;
        const int jj = (l + s) & (kTfB - 1);
        const int j = J * kTfB + jj;
        const bool ok = s <= s1 && iv && j < np && !(diag && s == 32 && l >= 32);
        if (ok) {
            double app = xi - sx[jj];
            app -= L * round(app * invL);
            double r2 = app * app;
            const double d0 = app;
            app = yi - sy[jj];
            app -= L * round(app * invL);
            const double d1 = app;
            r2 += app * app;
            app = zi - sz[jj];
            app -= L * round(app * invL);
            const double d2 = app;
            r2 += app * app;
            const double r = sqrt(r2);
            if (r <= r_cut) {
                const double ir = 1.0 / r;
                const double* t = tab + (ti * n_typ + styp[jj]) * 7;
                const double a = t[0], b = t[1], c = t[2], dd = t[3], sg = t[4], al = t[5],
                             be = t[6];
                const double ex = exp(b * (sg - r));
                const double ir2 = ir * ir;
                const double ir6 = ir2 * ir2 * ir2;
                const double ir8 = ir6 * ir2;
                const double fm = b * a * ex - 6 * c * (ir6 * ir) - 8 * dd * (ir8 * ir) - al;
                const double vm = a * ex - c * ir6 - dd * ir8 + al * r + be;
                const double g0 = fm * (d0 * ir), g1 = fm * (d1 * ir), g2 = fm * (d2 * ir);
                f0 += g0;
                f1 += g1;
                f2 += g2;
                e += vm;
                cnt += 1.0;
                acc[w][0][jj] -= g0;
                acc[w][1][jj] -= g1;
                acc[w][2][jj] -= g2;
                acc[w][3][jj] += vm;
                acc[w][4][jj] += 1.0;
            }
        }
    }
    own[w][0][l] = f0;
    own[w][1][l] = f1;
    own[w][2][l] = f2;
    own[w][3][l] = e;
    own[w][4][l] = cnt;
    __syncthreads();
    // part[item][component][slot]: slots 0..63 the i-role, 64..127 the j-role.
    double* o = part + (long)tf_item(I, J, nb) * 5 * 2 * kTfB;
    for (int x = tid; x < 5 * 2 * kTfB; x += kTfB * kTfWaves) {
        const int c = x / (2 * kTfB), slot = x % (2 * kTfB);
        const bool jr = slot >= kTfB;
        const int q = jr ? slot - kTfB : slot;
        double v = 0.0;
        for (int u = 0; u < kTfWaves; u++) v += jr ? acc[u][c][q] : own[u][c][q];
        o[x] = v;
    }
}

// Ion p = 64 B + l: its i-role partials of the pairs (B, J), J = B..nb-1, then
// its j-role partials of the pairs (I, B), I = 0..B -- nb + 1 items, split
// over kFoldGroups thread groups (group g takes items g, g + G, ...) and
// summed in group order.
constexpr int kFoldGroups = 8;

__global__ __launch_bounds__(kTfB * kFoldGroups) void k_tf_fold(
    int np, int nb, const double* __restrict__ part, double* __restrict__ fn,
    double* __restrict__ eion, int* __restrict__ counts) {
    __shared__ double sum[kFoldGroups][5][kTfB];
    const int B = blockIdx.x;
    const int l = threadIdx.x & (kTfB - 1), g = threadIdx.x / kTfB;
    double v[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int k = g; k <= nb; k += kFoldGroups) {
        // k < nb - B: i-role of (B, B + k); else j-role of (k - (nb - B), B).
        const int ni = nb - B;
        const double* o = k < ni ? part + (long)tf_item(B, B + k, nb) * 5 * 2 * kTfB + l
                                 : part + (long)tf_item(k - ni, B, nb) * 5 * 2 * kTfB + kTfB + l;
        double t[5];
#pragma unroll
        for (int c = 0; c < 5; c++) t[c] = o[c * 2 * kTfB];
#pragma unroll
        for (int c = 0; c < 5; c++) v[c] += t[c];
    }
#pragma unroll
    for (int c = 0; c < 5; c++) sum[g][c][l] = v[c];
    __syncthreads();
    const int p = B * kTfB + l;
    if (g != 0 || p >= np) return;
    double t[5] = {0.0, 0.0, 0.0, 0.0, 0.0};
    for (int u = 0; u < kFoldGroups; u++)
        for (int c = 0; c < 5; c++) t[c] += sum[u][c][l];
    fn[3 * p] = t[0];
    fn[3 * p + 1] = t[1];
    fn[3 * p + 2] = t[2];
    eion[p] = t[3];
    counts[p] = (int)t[4];
}

int tf_blocks(int np) { return (np + kTfB - 1) / kTfB; }
long tf_part_doubles(int np) {
    const long nb = tf_blocks(np);
    return nb * (nb + 1) / 2 * 5 * 2 * kTfB;
}

void launch_tf_rows(hipStream_t s, const PDev& p, int i_lo, int i_hi) {
    const int nb = tf_blocks(p.np);
    if (p.np == 0 || i_hi <= i_lo) return;
    const dim3 g(nb, i_hi - i_lo);
    if (p.n_typ <= kTfLdsTypes)
        MAZE_LAUNCH(k_tf_tiles<true>, g, dim3(kTfB * kTfWaves), 0, s, p.np, p.n_typ, p.L,
                    1.0 / p.L, p.r_cut, p.pos, p.types, p.table, p.tpart, i_lo);
    else
        MAZE_LAUNCH(k_tf_tiles<false>, g, dim3(kTfB * kTfWaves), 0, s, p.np, p.n_typ, p.L,
                    1.0 / p.L, p.r_cut, p.pos, p.types, p.table, p.tpart, i_lo);
}

void launch_tf_finish(hipStream_t s, const PDev& p, PScalars* ps, Reduce red) {
    if (p.np == 0) {
        launch_sum_scaled(s, 0, 0.5, p.eion, &ps->pe_tf, red);
        return;
    }
    const int nb = tf_blocks(p.np);
    MAZE_LAUNCH(k_tf_fold, dim3(nb), dim3(kTfB * kFoldGroups), 0, s, p.np, nb, p.tpart,
                       p.fn, p.eion, p.counts);
    launch_sum_scaled(s, p.np, 0.5, p.eion, &ps->pe_tf, red);
}

void launch_tf(hipStream_t s, const PDev& p, PScalars* ps, Reduce red) {
    launch_tf_rows(s, p, 0, tf_blocks(p.np));
    launch_tf_finish(s, p, ps, red);
}

// F_tot = 0 + 1 F_elec + 1 F_noel [+ 1 F_db + 1 F_ib + 1 F_np]
// (library/src/particles.c:777-795).
template <bool PB>
__global__ void k_ftot(long n, const double* fe, const double* fn, const double* fdb,
                       const double* fib, const double* fnp, double* ft) {
    const long i = blockIdx.x * (long)blockDim.x + threadIdx.x;
    if (i >= n) return;
    double v = 0.0;
    v += 1.0 * fe[i];
    v += 1.0 * fn[i];
    if (PB) {
        v += 1.0 * fdb[i];
        v += 1.0 * fib[i];
        v += 1.0 * fnp[i];
    }
    ft[i] = v;
}

void launch_force_total(hipStream_t s, const PDev& p, const double* fdb, const double* fib,
                        const double* fnp) {
    if (p.np == 0) return;
    if (fdb)
        MAZE_LAUNCH(k_ftot<true>, dim3(blocks_for(3L * p.np)), dim3(kBlock), 0, s, 3L * p.np,
                    p.fe, p.fn, fdb, fib, fnp, p.ft);
    else
        MAZE_LAUNCH(k_ftot<false>, dim3(blocks_for(3L * p.np)), dim3(kBlock), 0, s, 3L * p.np,
                    p.fe, p.fn, fdb, fib, fnp, p.ft);
}

// K = 1/2 sum m |v|^2 (library/src/particles.c:802-820); optionally the
// Verlet thermostat factor sqrt(T / T_now) (integrator_verlet.c:26-35).
__global__ void k_kinetic(int np, const double* vel, const double* mass, PScalars* ps,
                          double T_target, int want_scale) {
    __shared__ double sh[kBlock / kWave];
    double s = 0.0;
    for (int i = threadIdx.x; i < np; i += kBlock) {
        double app = 0.0;
        for (int d = 0; d < 3; d++) app += vel[3 * i + d] * vel[3 * i + d];
        s += mass[i] * app;
    }
    s = block_sum(s, sh);
    if (threadIdx.x == 0) {
        const double K = 0.5 * s;
        ps->kinetic = K;
        if (want_scale) {
            const double T_now = 2 * K / (3 * np * 3.1668e-6);
            ps->scale = sqrt(T_target / T_now);
        }
    }
}

void launch_kinetic(hipStream_t s, const PDev& p, PScalars* ps, Reduce, double T_target,
                    int want_scale) {
    MAZE_LAUNCH(k_kinetic, dim3(1), dim3(kBlock), 0, s, p.np, p.vel, p.mass, ps, T_target,
                       want_scale);
}

__device__ __forceinline__ double wrap_pos(double a, double L) {
    if (a < 0) return a + L;
    if (a >= L) return a - L;
    return a;
}

// Velocity Verlet part 1 (library/src/integrator_verlet.c:12-53); with base
// (the MD step's path) also the charge assignment's base cell of the new
// position, as k_cic_base forms it.
__global__ void k_verlet1(int np, double dt, double L, double* pos, double* vel, const double* ft,
                          const double* mass, const PScalars* ps, int rescale, int n, double h,
                          int* base) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    const double scale = rescale ? ps->scale : 1.0;
    for (int d = 0; d < 3; d++) {
        const int k = 3 * i + d;
        double v = vel[k];
        if (rescale) v *= scale;
        v += 0.5 * dt * ft[k] / mass[i];
        vel[k] = v;
        const double x = wrap_pos(pos[k] + dt * v, L);
        pos[k] = x;
        if (base) {
            int c = (int)floor(x / h);
            c %= n;
            if (c < 0) c += n;
            base[k] = c;
        }
    }
}

void launch_verlet_part1(hipStream_t s, const PDev& p, double dt, const PScalars* ps, int rescale,
                         bool with_base) {
    if (p.np == 0) return;
    MAZE_LAUNCH(k_verlet1, dim3(blocks_for(p.np)), dim3(kBlock), 0, s, p.np, dt, p.L, p.pos,
                       p.vel, p.ft, p.mass, ps, rescale, p.n, p.h, with_base ? p.base : nullptr);
}

// v += cdt F / m with cdt = 0.5 dt (Verlet) or 0.5 dt c2 (OVRVO V block).
__global__ void k_kick(int np, double cdt, double* vel, const double* ft, const double* mass) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    for (int d = 0; d < 3; d++) vel[3 * i + d] += cdt * ft[3 * i + d] / mass[i];
}

void launch_kick(hipStream_t s, const PDev& p, double cdt) {
    if (p.np == 0) return;
    MAZE_LAUNCH(k_kick, dim3(blocks_for(p.np)), dim3(kBlock), 0, s, p.np, cdt, p.vel, p.ft,
                       p.mass);
}

// x += c2 dt v with a single-image wrap (library/src/integrator_ovrvo.c:98-123).
__global__ void k_drift(int np, double c2dt, double L, double* pos, const double* vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    for (int d = 0; d < 3; d++) pos[3 * i + d] = wrap_pos(pos[3 * i + d] + c2dt * vel[3 * i + d], L);
}

void launch_drift(hipStream_t s, const PDev& p, double c2dt) {
    if (p.np == 0) return;
    MAZE_LAUNCH(k_drift, dim3(blocks_for(p.np)), dim3(kBlock), 0, s, p.np, c2dt, p.L, p.pos,
                       p.vel);
}

// Philox4x32-10 (Salmon et al., SC'11): 10 rounds of the published
// multiply-xor network with Weyl key increments.
__device__ __forceinline__ void philox4x32(uint32_t c[4], uint32_t k0, uint32_t k1) {
    for (int r = 0; r < 10; r++) {
        const uint64_t p0 = (uint64_t)0xD2511F53u * c[0];
        const uint64_t p1 = (uint64_t)0xCD9E8D57u * c[2];
        const uint32_t n0 = (uint32_t)(p1 >> 32) ^ c[1] ^ k0;
        const uint32_t n2 = (uint32_t)(p0 >> 32) ^ c[3] ^ k1;
        c[1] = (uint32_t)p1;
        c[3] = (uint32_t)p0;
        c[0] = n0;
        c[2] = n2;
        k0 += 0x9E3779B9u;
        k1 += 0xBB67AE85u;
    }
}

// O block: v <- sqrt(c1) v + sqrt((1 - c1) kB T / m) xi
// (library/src/integrator_ovrvo.c:45-78) with xi from (seed, call, ion).
__global__ void k_o_block(int np, uint64_t seed, uint64_t call, double c1s, double var2,
                          const double* mass, double* vel) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    uint32_t c[4] = {(uint32_t)i, (uint32_t)call, (uint32_t)(call >> 32), 0x4D415A45u};
    philox4x32(c, (uint32_t)seed, (uint32_t)(seed >> 32));
    const double sc = 2.3283064365386963e-10;
    const double u0 = ((double)c[0] + 0.5) * sc, u1 = ((double)c[1] + 0.5) * sc;
    const double u2 = ((double)c[2] + 0.5) * sc, u3 = ((double)c[3] + 0.5) * sc;
    const double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
    const double ta = 6.283185307179586 * u1, tb = 6.283185307179586 * u3;
    const double xi[3] = {ra * cos(ta), ra * sin(ta), rb * cos(tb)};
    const double var1 = sqrt(var2 / mass[i]);
    for (int d = 0; d < 3; d++) {
        double v = vel[3 * i + d];
        v *= c1s;
        v += var1 * xi[d];
        vel[3 * i + d] = v;
    }
}

void launch_o_block(hipStream_t s, const PDev& p, uint64_t seed, uint64_t call, double c1s,
                    double var2) {
    if (p.np == 0) return;
    MAZE_LAUNCH(k_o_block, dim3(blocks_for(p.np)), dim3(kBlock), 0, s, p.np, seed, call, c1s,
                       var2, p.mass, p.vel);
}

// sum m v per component (library/src/particles.c:822-839).
__global__ void k_momentum(int np, const double* vel, const double* mass, PScalars* ps) {
    __shared__ double sh[kBlock / kWave];
    for (int d = 0; d < 3; d++) {
        double s = 0.0;
        for (int i = threadIdx.x; i < np; i += kBlock) s += mass[i] * vel[3 * i + d];
        s = block_sum(s, sh);
        if (threadIdx.x == 0) ps->mom[d] = s;
        __syncthreads();
    }
}

void launch_momentum(hipStream_t s, const PDev& p, PScalars* ps) {
    MAZE_LAUNCH(k_momentum, dim3(1), dim3(kBlock), 0, s, p.np, p.vel, p.mass, ps);
}

// v -= 2 (sum of v over the ion's type) / n_p (library/src/particles.c:841-857).
__global__ void k_rescale(int np, int n_typ, const int* types, double* vel, PScalars* ps) {
    __shared__ double sh[kBlock / kWave];
    for (int t = 0; t < n_typ; t++)
        for (int d = 0; d < 3; d++) {
            double s = 0.0;
            for (int i = threadIdx.x; i < np; i += kBlock)
                if (types[i] == t) s += vel[3 * i + d];
            s = block_sum(s, sh);
            if (threadIdx.x == 0) ps->typesum[t * 3 + d] = s;
            __syncthreads();
        }
    __syncthreads();
    for (int i = threadIdx.x; i < np; i += kBlock)
        for (int d = 0; d < 3; d++) vel[3 * i + d] -= 2 * ps->typesum[types[i] * 3 + d] / np;
}

void launch_rescale_velocities(hipStream_t s, const PDev& p, PScalars* ps) {
    MAZE_LAUNCH(k_rescale, dim3(1), dim3(kBlock), 0, s, p.np, p.n_typ, p.types, p.vel, ps);
}

}  // namespace maze

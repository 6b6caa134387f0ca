// pb_kernels.hip -- Poisson-Boltzmann boundary from ion radii.
//
// Maps (library/src/grid_general.c:138-301).  With the smooth step
//   H(s) = -(1/(4 w^3)) s^3 + (3/(4 w^2)) s^2,  s = |x - x_ion| - r + w,
// each ion multiplies the solvent fraction of the points and faces of its box
// (+-(floor((r+w)/h)+1) points around its cell) by 0 inside r - w and by H in
// the shell up to r + w; eps_x/y/z = eps_int + (eps_s - eps_int) * product on
// the +x/+y/+z face of each point, k2 = kbar2 * product on the point.  The
// reference scatters ion by ion; here each point gathers: a work-group owns an
// 8^3 tile, collects the ions whose boxes can reach it from the bins around
// it, sorts them by ion id in LDS and applies their factors in that order, so
// every product is formed in the reference's order.
//
// Forces (library/src/particles.c:420-775).  One work-group per ion walks the
// ion's box (+-(floor((r+w)/h)+2) points) and sums the ionic-boundary term at
// the shell points and the dielectric-boundary and non-polar terms at the
// shell faces; the sums are reduced in a fixed tree (deterministic, ordered
// differently from the reference's serial loop).
#include <hipcub/hipcub.hpp>

#include "pb.hpp"

namespace maze {

namespace {

struct StepCoef {
    double c3, c2;   // H(s) = c3 s^3 + c2 s^2
    double d2, d1;   // H'(s) = d2 s^2 + d1 s
};

inline StepCoef step_coef(double w) {
    const double w2 = w * w, w3 = w2 * w;
    return {-(1 / (4 * w3)), 3 / (4 * w2), -(3 / (4 * w3)), 3 / (2 * w2)};
}

__device__ __forceinline__ double step_h(const StepCoef& c, double s) {
    return c.c3 * cube(s) + c.c2 * (s * s);
}

// H'(s) / H(s) (library/src/particles.c:407-417).
__device__ __forceinline__ double step_dlog(const StepCoef& c, double s) {
    return (c.d2 * (s * s) + c.d1 * s) / (c.c3 * cube(s) + c.c2 * (s * s));
}

__device__ __forceinline__ int cell_of(double x, double h) { return (int)floor(x / h); }

__device__ __forceinline__ int wrap1(int i, int n) {
    return i < 0 ? i + n : (i >= n ? i - n : i);
}

// i mod n in [0, n) for any i (cells of positions outside the box included).
__device__ __forceinline__ int pmod(int i, int n) {
    const int r = i % n;
    return r < 0 ? r + n : r;
}

}  // namespace

// ---- binning ---------------------------------------------------------------------
__global__ void k_pb_bin(int np, int n, double h, int nb, const double* __restrict__ pos,
                         unsigned int* __restrict__ keys, int* __restrict__ ids) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= np) return;
    int b[3];
    for (int a = 0; a < 3; a++) {
        b[a] = pmod(cell_of(pos[3 * i + a], h), n) / kPbTile;
    }
    keys[i] = (unsigned int)((b[0] * nb + b[1]) * nb + b[2]);
    ids[i] = i;
}

size_t pb_sort_bytes(int np, int key_bits) {
    size_t bytes = 0;
    MAZE_HIP(hipcub::DeviceRadixSort::SortPairs(nullptr, bytes, (unsigned int*)nullptr,
                                                (unsigned int*)nullptr, (int*)nullptr,
                                                (int*)nullptr, np, 0, key_bits));
    return bytes;
}

void launch_pb_bin(hipStream_t s, PBDev& b, const double* pos) {
    if (b.np == 0) return;
    MAZE_LAUNCH(k_pb_bin, dim3((b.np + kBlock - 1) / kBlock), dim3(kBlock), 0, s, b.np, b.n, b.h,
                b.nb, pos, b.keys, b.ids);
    size_t bytes = b.sort_bytes;
    MAZE_HIP(hipcub::DeviceRadixSort::SortPairs(b.sort_tmp, bytes, b.keys, b.keys_sorted, b.ids,
                                                b.ids_sorted, b.np, 0, b.key_bits, s));
}

// ---- maps --------------------------------------------------------------------------
// Bins along one axis whose ions can reach the tile's points [t0, t0 + kPbTile):
// those holding cells in [t0 - R, t0 + kPbTile - 1 + R] (periodic).
struct AxisBins {
    int first, count;
};

__device__ __forceinline__ AxisBins axis_bins(int t0, int R, int n, int nb) {
    const int len = kPbTile + 2 * R;
    if (len + kPbTile > n) return {0, nb};
    int v0 = (t0 - R) % n;
    if (v0 < 0) v0 += n;
    const int v1 = (v0 + len - 1) % n;
    const int b0 = v0 / kPbTile, b1 = v1 / kPbTile;
    return {b0, (b1 - b0 + nb) % nb + 1};
}

// Offset d = g - c (mod n) of grid index g from an ion's cell c, as the one
// value in [-R, R] (the ion's box), or a value beyond R when g is outside it.
__device__ __forceinline__ int box_offset(int g, int c, int n, int R) {
    int d = g - c;
    if (d < 0) d += n;
    else if (d >= n) d -= n;
    if (d > R) d = (d >= n - R) ? d - n : R + 1;
    return d;
}

// The reference's mask of one array entry: unchanged beyond r + w, times H in
// the shell, 0 inside r - w.
__device__ __forceinline__ void apply_mask(double& v, double r2, double rp2, double rm2, double r,
                                           double w, const StepCoef& c) {
    if (r2 >= rp2) return;
    if (r2 > rm2) v *= step_h(c, sqrt(r2) - r + w);
    else v = 0.0;
}

struct PBMapArgs {
    int n, nb, rmax, np;
    double h, w, eps_s, eps_int, kbar2;
    StepCoef c;
    const double* pos;
    const double* radii;
    const double* rp2;
    const double* rm2;
    const int* rmap;
    const unsigned int* keys_sorted;
    const int* ids_sorted;
    double *ex, *ey, *ez, *k2;
    int* flag;
};

// One work-group per 8^3 tile; thread t owns points (i, j, k) and (i + 4, j, k)
// of the tile with k = t % 8, j = t / 8 % 8, i = t / 64.
static_assert(kPbTile == 8 && kBlock == 256, "tile-to-thread map assumes 8^3 tiles, 256 threads");
__global__ __launch_bounds__(kBlock) void k_pb_maps(PBMapArgs a) {
    __shared__ int list[kPbCap];
    __shared__ int seg_lo[kPbMaxSeg], seg_cnt[kPbMaxSeg], seg_off[kPbMaxSeg];
    __shared__ int s_total, s_bad;
    const int nb = a.nb, n = a.n;
    const int b = blockIdx.x;
    const int tb[3] = {b / (nb * nb), (b / nb) % nb, b % nb};
    AxisBins ab[3];
    for (int d = 0; d < 3; d++) ab[d] = axis_bins(tb[d] * kPbTile, a.rmax, n, nb);
    // Segments: rows (bx, by) of the candidate bins, each split where its z
    // range wraps, as ranges of consecutive bin keys.
    const bool zwrap = ab[2].first + ab[2].count > nb;
    const int per_row = zwrap ? 2 : 1;
    const int nseg = ab[0].count * ab[1].count * per_row;
    if (threadIdx.x == 0) s_bad = nseg > kPbMaxSeg;
    __syncthreads();
    if (s_bad) {
        if (threadIdx.x == 0) a.flag[0] = 1;
        return;
    }
    for (int sidx = threadIdx.x; sidx < nseg; sidx += kBlock) {
        const int row = sidx / per_row, part = sidx % per_row;
        const int bx = (ab[0].first + row / ab[1].count) % nb;
        const int by = (ab[1].first + row % ab[1].count) % nb;
        int z0, z1;  // bins [z0, z1)
        if (!zwrap) {
            z0 = ab[2].first;
            z1 = ab[2].first + ab[2].count;
        } else if (part == 0) {
            z0 = ab[2].first;
            z1 = nb;
        } else {
            z0 = 0;
            z1 = ab[2].first + ab[2].count - nb;
        }
        const unsigned int base = (unsigned int)((bx * nb + by) * nb);
        const int lo = lower_bound_u32(a.keys_sorted, a.np, base + z0);
        const int hi = lower_bound_u32(a.keys_sorted, a.np, base + z1);
        seg_lo[sidx] = lo;
        seg_cnt[sidx] = hi - lo;
    }
    __syncthreads();
    if (threadIdx.x == 0) {
        int t = 0;
        for (int sidx = 0; sidx < nseg; sidx++) {
            seg_off[sidx] = t;
            t += seg_cnt[sidx];
        }
        s_total = t;
        s_bad = t > kPbCap;
    }
    __syncthreads();
    if (s_bad) {
        if (threadIdx.x == 0) a.flag[0] = 1;
        return;
    }
    const int total = s_total;
    for (int sidx = threadIdx.x; sidx < nseg; sidx += kBlock)
        for (int q = 0; q < seg_cnt[sidx]; q++)
            list[seg_off[sidx] + q] = a.ids_sorted[seg_lo[sidx] + q];
    int P = 1;
    while (P < total) P <<= 1;
    for (int q = total + threadIdx.x; q < P; q += kBlock) list[q] = 0x7fffffff;
    __syncthreads();
    // Bitonic sort of the candidate ids (ascending).
    for (int k = 2; k <= P; k <<= 1) {
        for (int j = k >> 1; j > 0; j >>= 1) {
            for (int q = threadIdx.x; q < P; q += kBlock) {
                const int o = q ^ j;
                if (o > q) {
                    const int x = list[q], y = list[o];
                    if ((x > y) == ((q & k) == 0)) {
                        list[q] = y;
                        list[o] = x;
                    }
                }
            }
            __syncthreads();
        }
    }
    const int tk = threadIdx.x & 7, tj = (threadIdx.x >> 3) & 7, ti = threadIdx.x >> 6;
    const int gk = tb[2] * kPbTile + tk, gj = tb[1] * kPbTile + tj;
    const int gi0 = tb[0] * kPbTile + ti, gi1 = gi0 + 4;
    const bool ok_jk = gj < n && gk < n;
    const bool ok0 = ok_jk && gi0 < n, ok1 = ok_jk && gi1 < n;
    const double h = a.h, hd2 = h / 2.0, w = a.w;
    const double e0 = a.eps_s - a.eps_int;
    double k2v[2] = {a.kbar2, a.kbar2}, exv[2] = {e0, e0}, eyv[2] = {e0, e0}, ezv[2] = {e0, e0};
    // Candidates in id order, a chunk at a time: those whose sphere r + w can reach a
    // point or face of the tile are kept (order preserved) with their data in LDS.
    __shared__ double c_p[3][kPbChunk], c_r[kPbChunk], c_rp2[kPbChunk], c_rm2[kPbChunk];
    __shared__ int c_cell[3][kPbChunk], c_R[kPbChunk];
    __shared__ int c_wsum[kBlock / kWave];
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    bool done = false;  // this wave's points are all 0 in every map (later factors keep 0)
    for (int c0 = 0; c0 < total; c0 += kBlock) {
        const int c = c0 + threadIdx.x;
        bool keep = false;
        double px = 0, py = 0, pz = 0, r = 0, rp2 = 0, rm2 = 0;
        int R = 0, cell[3] = {0, 0, 0};
        if (c < total) {
            const int id = list[c];
            px = a.pos[3 * id]; py = a.pos[3 * id + 1]; pz = a.pos[3 * id + 2];
            r = a.radii[id]; rp2 = a.rp2[id]; rm2 = a.rm2[id]; R = a.rmap[id];
            const double pp[3] = {px, py, pz};
            double d2 = 0.0;
            keep = true;
            for (int d = 0; d < 3 && keep; d++) {
                cell[d] = cell_of(pp[d], h);
                // Nearest point or +face of the tile along this axis inside the box.
                double best = -1.0;
                for (int q = 0; q < kPbTile; q++) {
                    const int g = tb[d] * kPbTile + q;
                    if (g >= n) break;
                    const int o = box_offset(g, cell[d], n, R);
                    if (o > R) continue;
                    double dd = fabs(pp[d] - (cell[d] + o) * h) - hd2;
                    dd = dd > 0.0 ? dd : 0.0;
                    if (best < 0.0 || dd < best) best = dd;
                }
                if (best < 0.0) keep = false;
                else d2 += best * best;
            }
            keep = keep && d2 < rp2 * (1.0 + 1e-9) + 1e-300;
        }
        // Stable compaction of the kept candidates of this chunk.
        const unsigned long long bal = __ballot(keep);
        const int before = __popcll(bal & ((1ull << lane) - 1ull));
        if (lane == 0) c_wsum[wave] = __popcll(bal);
        __syncthreads();
        int off = 0, kept = 0;
        for (int q = 0; q < kBlock / kWave; q++) {
            if (q < wave) off += c_wsum[q];
            kept += c_wsum[q];
        }
        if (keep) {
            const int slot = off + before;
            c_p[0][slot] = px; c_p[1][slot] = py; c_p[2][slot] = pz;
            c_r[slot] = r; c_rp2[slot] = rp2; c_rm2[slot] = rm2; c_R[slot] = R;
            c_cell[0][slot] = cell[0]; c_cell[1][slot] = cell[1]; c_cell[2][slot] = cell[2];
        }
        __syncthreads();
        for (int q = 0; q < kept && !done; q++) {
            const int Rq = c_R[q];
            const int oj = box_offset(gj, c_cell[1][q], n, Rq);
            const int ok = box_offset(gk, c_cell[2][q], n, Rq);
            if (oj <= Rq && ok <= Rq) {
                const double qx = c_p[0][q], qy = c_p[1][q], qz = c_p[2][q];
                const double rq = c_r[q], qp2 = c_rp2[q], qm2 = c_rm2[q];
                const double dy = qy - (c_cell[1][q] + oj) * h, dy2 = dy * dy;
                const double dz = qz - (c_cell[2][q] + ok) * h, dz2 = dz * dz;
                const double ty = dy - hd2, tz = dz - hd2;
#pragma unroll
                for (int u = 0; u < 2; u++) {
                    const int gi = u == 0 ? gi0 : gi1;
                    const int oi = box_offset(gi, c_cell[0][q], n, Rq);
                    if (oi > Rq) continue;
                    const double dx = qx - (c_cell[0][q] + oi) * h, dx2 = dx * dx;
                    const double tx = dx - hd2;
                    apply_mask(k2v[u], dx2 + dy2 + dz2, qp2, qm2, rq, w, a.c);
                    apply_mask(exv[u], tx * tx + dy2 + dz2, qp2, qm2, rq, w, a.c);
                    apply_mask(eyv[u], dx2 + ty * ty + dz2, qp2, qm2, rq, w, a.c);
                    apply_mask(ezv[u], dx2 + dy2 + tz * tz, qp2, qm2, rq, w, a.c);
                }
            }
            const bool zero = k2v[0] == 0.0 && exv[0] == 0.0 && eyv[0] == 0.0 && ezv[0] == 0.0 &&
                              k2v[1] == 0.0 && exv[1] == 0.0 && eyv[1] == 0.0 && ezv[1] == 0.0;
            done = __all(zero);
        }
        __syncthreads();
    }
    const long n2 = (long)n * n;
#pragma unroll
    for (int u = 0; u < 2; u++) {
        if (!(u == 0 ? ok0 : ok1)) continue;
        const long idx = (u == 0 ? gi0 : gi1) * n2 + (long)gj * n + gk;
        a.k2[idx] = k2v[u];
        a.ex[idx] = exv[u] + a.eps_int;
        a.ey[idx] = eyv[u] + a.eps_int;
        a.ez[idx] = ezv[u] + a.eps_int;
    }
}

void launch_pb_maps(hipStream_t s, const PBDev& b, const double* pos, double* ex, double* ey,
                    double* ez, double* k2) {
    PBMapArgs a{};
    a.n = b.n; a.nb = b.nb; a.rmax = b.rmax; a.np = b.np;
    a.h = b.h; a.w = b.w; a.eps_s = b.eps_s; a.eps_int = b.eps_int; a.kbar2 = b.kbar2;
    a.c = step_coef(b.w);
    a.pos = pos; a.radii = b.radii; a.rp2 = b.rp2; a.rm2 = b.rm2; a.rmap = b.rmap;
    a.keys_sorted = b.keys_sorted; a.ids_sorted = b.ids_sorted;
    a.ex = ex; a.ey = ey; a.ez = ez; a.k2 = k2; a.flag = b.flag;
    MAZE_LAUNCH(k_pb_maps, dim3(b.nb * b.nb * b.nb), dim3(kBlock), 0, s, a);
}

// ---- forces ------------------------------------------------------------------------
struct PBForceArgs {
    int n, nonpolar;
    double h, w, eps_int, sdb, snp;
    StepCoef c;
    const double* pos;
    const double* radii;
    const double* rp2;
    const double* rm2;
    const int* rfrc;
    const double *phi, *ex, *ey, *ez, *k2;
    double *fdb, *fib, *fnp;
};

__device__ __forceinline__ bool in_shell(double r2, double rp2, double rm2) {
    return r2 < rp2 && r2 > rm2;
}

// One face term (library/src/particles.c:585-764): the face at (fx, fy, fz)
// from the ion, eps_f its dielectric, dphi the potential step across it, g the
// normalised dielectric gradient component; sgn -1 below the point, +1 above.
__device__ __forceinline__ void face_term(double fx, double fy, double fz, double r2, double r,
                                          double w, const StepCoef& c, double eps_f,
                                          double eps_int, double dphi, double phic, double g,
                                          int sgn, int nonpolar, double* acc) {
    const double r1 = sqrt(r2) + 1E-12;
    const double ux = fx / r1, uy = fy / r1, uz = fz / r1;
    const double k = (eps_f - eps_int) * step_dlog(c, r1 - r + w);
    acc[0] -= k * dphi * phic * ux;
    acc[1] -= k * dphi * phic * uy;
    acc[2] -= k * dphi * phic * uz;
    if (nonpolar) {
        const double s = sgn < 0 ? -1.0 : 1.0;
        acc[6] += s * (k * g * ux);
        acc[7] += s * (k * g * uy);
        acc[8] += s * (k * g * uz);
    }
}

__global__ __launch_bounds__(kBlock) void k_pb_forces(PBForceArgs a) {
    __shared__ double sh[kBlock / kWave];
    const int ion = blockIdx.x;
    const int n = a.n;
    const long n2 = (long)n * n;
    const double h = a.h, hd2 = h / 2.0, w = a.w;
    const double px = a.pos[3 * ion], py = a.pos[3 * ion + 1], pz = a.pos[3 * ion + 2];
    const double r = a.radii[ion], rp2 = a.rp2[ion], rm2 = a.rm2[ion];
    const int R = a.rfrc[ion], S = 2 * R + 1;
    const int cx = cell_of(px, h), cy = cell_of(py, h), cz = cell_of(pz, h);
    double acc[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};  // db xyz, ib xyz, np xyz
    // Only points whose centre lies within h/2 of the shell (rm2, rp2) can have a
    // shell point or face: per box row (qi, qj) the k ranges of centre distances in
    // (sqrt(rm2) - h/2, sqrt(rp2) + h/2), widened by a point on each side, are listed
    // (two per row) and their points spread over the threads.
    __shared__ int e_start[2 * kBlock], e_off[2 * kBlock + 1];
    __shared__ int w_tot[kBlock / kWave];
    const double band_hi = (sqrt(rp2) + hd2) * (1.0 + 1e-12);
    double band_lo = (sqrt(rm2) - hd2) * (1.0 - 1e-12);
    band_lo = band_lo > 0.0 ? band_lo : 0.0;
    const double hi2 = band_hi * band_hi, lo2 = band_lo * band_lo;
    const int lane = threadIdx.x & (kWave - 1), wave = threadIdx.x / kWave;
    const int rows = S * S;
    for (int row0 = 0; row0 < rows; row0 += kBlock) {
        int st[2] = {0, 0}, cnt[2] = {0, 0};
        const int row = row0 + threadIdx.x;
        if (row < rows) {
            const int qi = row / S, qj = row % S;
            const double dx = px - (cx + qi - R) * h, dy = py - (cy + qj - R) * h;
            const double rho2 = dx * dx + dy * dy;
            if (rho2 < hi2) {
                const double b = sqrt(hi2 - rho2);
                const double c = rho2 < lo2 ? sqrt(lo2 - rho2) : 0.0;
                // dz = pz - (cz + qk - R) h in [c, b] or [-b, -c].
                const double base = pz / h - cz + R;
                int a0 = (int)floor(base - b / h) - 1, a1 = (int)ceil(base - c / h) + 1;
                int b0 = (int)floor(base + c / h) - 1, b1 = (int)ceil(base + b / h) + 1;
                a0 = max(a0, 0); a1 = min(a1, S - 1);
                b0 = max(b0, 0); b1 = min(b1, S - 1);
                if (a1 >= b0 - 1) {  // the two ranges touch: one range
                    a1 = max(a1, b1);
                    b0 = 1; b1 = 0;
                }
                if (a1 >= a0) { st[0] = row * S + a0; cnt[0] = a1 - a0 + 1; }
                if (b1 >= b0) { st[1] = row * S + b0; cnt[1] = b1 - b0 + 1; }
            }
        }
        // Exclusive scan of the 2 * kBlock range lengths.
        int mine = cnt[0] + cnt[1], incl = mine;
        for (int o = 1; o < kWave; o <<= 1) {
            const int v = __shfl_up(incl, o);
            if (lane >= o) incl += v;
        }
        if (lane == kWave - 1) w_tot[wave] = incl;
        __syncthreads();
        int wbase = 0, total = 0;
        for (int q = 0; q < kBlock / kWave; q++) {
            if (q < wave) wbase += w_tot[q];
            total += w_tot[q];
        }
        const int ex0 = wbase + incl - mine;
        e_start[2 * threadIdx.x] = st[0];
        e_start[2 * threadIdx.x + 1] = st[1];
        e_off[2 * threadIdx.x] = ex0;
        e_off[2 * threadIdx.x + 1] = ex0 + cnt[0];
        if (threadIdx.x == 0) e_off[2 * kBlock] = total;
        __syncthreads();
        for (int f = threadIdx.x; f < total; f += kBlock) {
            // Last range with e_off <= f (empty ranges share their successor's offset).
            int lo = 0, hi = 2 * kBlock - 1;
            while (lo < hi) {
                const int mid = (lo + hi + 1) >> 1;
                if (e_off[mid] <= f) lo = mid;
                else hi = mid - 1;
            }
            const int q = e_start[lo] + (f - e_off[lo]);
            const int qi = q / (S * S), qj = (q / S) % S, qk = q % S;
            const int ii = cx + qi - R, jj = cy + qj - R, kk = cz + qk - R;
            const double dx = px - ii * h, dy = py - jj * h, dz = pz - kk * h;
            const double dx2 = dx * dx, dy2 = dy * dy, dz2 = dz * dz;
            const double xm = dx + hd2, xp = dx - hd2, ym = dy + hd2, yp = dy - hd2;
            const double zm = dz + hd2, zp = dz - hd2;
            const double r2c = dx2 + dy2 + dz2;
            const double r2xm = xm * xm + dy2 + dz2, r2xp = xp * xp + dy2 + dz2;
            const double r2ym = dx2 + ym * ym + dz2, r2yp = dx2 + yp * yp + dz2;
            const double r2zm = dx2 + dy2 + zm * zm, r2zp = dx2 + dy2 + zp * zp;
            const bool bc = in_shell(r2c, rp2, rm2);
            const bool bxm = in_shell(r2xm, rp2, rm2), bxp = in_shell(r2xp, rp2, rm2);
            const bool bym = in_shell(r2ym, rp2, rm2), byp = in_shell(r2yp, rp2, rm2);
            const bool bzm = in_shell(r2zm, rp2, rm2), bzp = in_shell(r2zp, rp2, rm2);
            if (!(bc | bxm | bxp | bym | byp | bzm | bzp)) continue;
            const int i = pmod(ii, n), j = pmod(jj, n), k = pmod(kk, n);
            const long ci = i * n2, cj = (long)j * n;
            const long im = wrap1(i - 1, n) * n2, ip = wrap1(i + 1, n) * n2;
            const long jm = (long)wrap1(j - 1, n) * n, jp = (long)wrap1(j + 1, n) * n;
            const int km = wrap1(k - 1, n), kp = wrap1(k + 1, n);
            const long c = ci + cj + k;
            const double phic = a.phi[c];
            // A term whose coefficient is exactly 0 (k2 = 0, or eps_f = eps_int inside
            // another ion's core) adds +-0 (H > 0 in the shell, so its log-derivative
            // is finite) and is skipped.
            const double k2c = bc ? a.k2[c] : 0.0;
            if (k2c != 0.0) {
                const double r1 = sqrt(r2c) + 1E-12;
                const double f = k2c * phic * phic * step_dlog(a.c, r1 - r + w);
                acc[3] += f * (dx / r1);
                acc[4] += f * (dy / r1);
                acc[5] += f * (dz / r1);
            }
            if (bxm | bxp | bym | byp | bzm | bzp) {
                const double exc = a.ex[c], eyc = a.ey[c], ezc = a.ez[c];
                const double exb = a.ex[im + cj + k], eyb = a.ey[ci + jm + k];
                const double ezb = a.ez[ci + cj + km];
                const double gx = exc - exb, gy = eyc - eyb, gz = ezc - ezb;
                const double gn = sqrt(gx * gx + gy * gy + gz * gz) / h;
                const double ig = (gn > 0.0) ? 1.0 / gn : 0.0;
                const double gxn = ig * gx, gyn = ig * gy, gzn = ig * gz;
                const double ei = a.eps_int;
                auto face = [&](bool in, double fx, double fy, double fz, double r2f,
                                double eps_f, long nbr, double g, int sgn) {
                    if (in && eps_f != ei)
                        face_term(fx, fy, fz, r2f, r, w, a.c, eps_f, ei, a.phi[nbr] - phic, phic,
                                  g, sgn, a.nonpolar, acc);
                };
                face(bxm, xm, dy, dz, r2xm, exb, im + cj + k, gxn, -1);
                face(bxp, xp, dy, dz, r2xp, exc, ip + cj + k, gxn, +1);
                face(bym, dx, ym, dz, r2ym, eyb, ci + jm + k, gyn, -1);
                face(byp, dx, yp, dz, r2yp, eyc, ci + jp + k, gyn, +1);
                face(bzm, dx, dy, zm, r2zm, ezb, ci + cj + km, gzn, -1);
                face(bzp, dx, dy, zp, r2zp, ezc, ci + cj + kp, gzn, +1);
            }
        }
        __syncthreads();
    }
    double tot[9];
#pragma unroll
    for (int q = 0; q < 9; q++) tot[q] = block_sum(acc[q], sh);
    if (threadIdx.x == 0) {
        for (int d = 0; d < 3; d++) {
            a.fdb[3 * ion + d] = tot[d] * a.sdb;
            a.fib[3 * ion + d] = tot[3 + d] * a.sdb;
            a.fnp[3 * ion + d] = tot[6 + d] * a.snp;
        }
    }
}

void launch_pb_forces(hipStream_t s, const PBDev& b, const double* pos, const double* phi,
                      const double* ex, const double* ey, const double* ez, const double* k2) {
    if (b.np == 0) return;
    PBForceArgs a{};
    a.n = b.n; a.nonpolar = b.nonpolar;
    a.h = b.h; a.w = b.w; a.eps_int = b.eps_int;
    a.sdb = b.h / (8.0 * kPi);
    a.snp = -b.gamma_np * b.h / (b.eps_s - b.eps_int);
    a.c = step_coef(b.w);
    a.pos = pos; a.radii = b.radii; a.rp2 = b.rp2; a.rm2 = b.rm2; a.rfrc = b.rfrc;
    a.phi = phi; a.ex = ex; a.ey = ey; a.ez = ez; a.k2 = k2;
    a.fdb = b.fdb; a.fib = b.fib; a.fnp = b.fnp;
    MAZE_LAUNCH(k_pb_forces, dim3(b.np), dim3(kBlock), 0, s, a);
}

}  // namespace maze

/*
 * maze_oracle.c -- CPU oracle (test infrastructure, FP64, single thread).
 *
 * Restates the reference hot path (Crivella/MaZe_Poisson library/src) with
 * the same floating-point operation order so that it can be pinned bit for
 * bit against the reference C.  Only tests/, __graft_entry__.smoke() and
 * bench.py's cpu_baseline leg may load it; the product path never does.
 *
 * Citations are path:line in the reference tree.
 */
#include <math.h>
#include <stdint.h>
#include <stdlib.h>
#include <string.h>

#include "maze_oracle.h"

#define OR_PI 3.14159265358979323846

static inline long wrapi(long i, long n) { return i < 0 ? i + n : (i >= n ? i - n : i); }

/* ------------------------------------------------------------------------ */
/* 7-point periodic Laplacian, unscaled: sum of the 6 neighbours - 6 u.
 * library/src/laplace.c:24-75.  Summation order: i+1, i-1, j+1, j-1, k+1,
 * k-1, then minus 6u (laplace.c:301-309). */
void or_laplace(const double *u, double *out, int n) {
    const long N = n, N2 = N * N;
    for (long i = 0; i < N; i++) {
        const long ip = wrapi(i + 1, N) * N2, im = wrapi(i - 1, N) * N2, i0 = i * N2;
        for (long j = 0; j < N; j++) {
            const long jp = wrapi(j + 1, N) * N, jm = wrapi(j - 1, N) * N, j0 = j * N;
            for (long k = 0; k < N; k++) {
                const long kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                out[i0 + j0 + k] = u[ip + j0 + k] + u[im + j0 + k] + u[i0 + jp + k] +
                                   u[i0 + jm + k] + u[i0 + j0 + kp] + u[i0 + j0 + km] -
                                   u[i0 + j0 + k] * 6.0;
            }
        }
    }
}

/* Variable-coefficient (linear Poisson-Boltzmann) operator.  eps_x[c] is the
 * dielectric value on the face between c and its +x neighbour.
 * library/src/laplace.c:77-145 (term order :366-379). */
void or_laplace_pb(const double *u, double *out, int n, const double *ex,
                   const double *ey, const double *ez, const double *k2) {
    const long N = n, N2 = N * N;
    for (long i = 0; i < N; i++) {
        const long ip = wrapi(i + 1, N) * N2, im = wrapi(i - 1, N) * N2, i0 = i * N2;
        for (long j = 0; j < N; j++) {
            const long jp = wrapi(j + 1, N) * N, jm = wrapi(j - 1, N) * N, j0 = j * N;
            for (long k = 0; k < N; k++) {
                const long kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                const long c = i0 + j0 + k, cx = im + j0 + k, cy = i0 + jm + k, cz = i0 + j0 + km;
                out[c] = u[ip + j0 + k] * ex[c] + u[cx] * ex[cx] + u[i0 + jp + k] * ey[c] +
                         u[cy] * ey[cy] + u[i0 + j0 + kp] * ez[c] + u[cz] * ez[cz] -
                         u[c] * (ex[c] + ex[cx] + ey[c] + ey[cy] + ez[c] + ez[cz] + k2[c]);
            }
        }
    }
}

/* ---- BLAS-1 with the reference's serial order (library/src/linalg.c:105-212) */
static double dot(const double *a, const double *b, long n) {
    double s = 0.0;
    for (long i = 0; i < n; i++) s += a[i] * b[i];
    return s;
}
static void axpy(const double *v, double *u, double alpha, long n) { /* u += alpha v */
    for (long i = 0; i < n; i++) u[i] += alpha * v[i];
}
static void scal(double *x, double alpha, long n) {
    for (long i = 0; i < n; i++) x[i] *= alpha;
}
static double amax(const double *u, long n) {
    double m = 0.0;
    for (long i = 0; i < n; i++) {
        double a = fabs(u[i]);
        if (a > m) m = a;
    }
    return m;
}

typedef void (*op_fn)(const double *, double *, int, const void *);
struct pb_coef { const double *ex, *ey, *ez, *k2; };
static void op_lap(const double *u, double *o, int n, const void *c) { (void)c; or_laplace(u, o, n); }
static void op_pb(const double *u, double *o, int n, const void *c) {
    const struct pb_coef *p = (const struct pb_coef *)c;
    or_laplace_pb(u, o, n, p->ex, p->ey, p->ez, p->k2);
}

/* CG with the implicit Jacobi preconditioner P = -6 I, so v = -r/6 and
 * p = -v = r/6.  x holds the initial guess on entry (the reference's
 * in-place form x == x0).  Stops when max|r| <= tol, checked after each
 * update; returns the 0-based iteration index, -1 after n*n iterations.
 * library/src/plcg.c:28-90 (Poisson) and :169-238 (PB, same recurrence). */
static int cg_implicit(op_fn op, const void *ctx, const double *b, double *x, double tol, int n) {
    const long G = (long)n * n * n, limit = (long)n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    int res = -1;
    op(x, r, n, ctx);               /* r = A x */
    axpy(b, r, -1.0, G);            /* r = A x - b */
    memcpy(p, r, G * sizeof(double));
    scal(p, 1.0 / 6.0, G);          /* p = r / 6 */
    double r_dot_v = -dot(r, p, G);
    for (long it = 0; it < limit; it++) {
        op(p, Ap, n, ctx);
        double alpha = r_dot_v / dot(p, Ap, G);
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        double rr = dot(r, r, G);
        if (amax(r, G) <= tol) { res = (int)it; break; }
        double rn_dot_vn = -rr / 6.0;
        double beta = rn_dot_vn / r_dot_v;
        r_dot_v = rn_dot_vn;
        scal(p, beta, G);
        axpy(r, p, 1.0 / 6.0, G);
    }
    free(r); free(Ap); free(p);
    return res;
}

int or_conj_grad(const double *b, double *x, double tol, int n) {
    return cg_implicit(op_lap, NULL, b, x, tol, n);
}

int or_conj_grad_pb(const double *b, double *x, double tol, int n, const double *ex,
                    const double *ey, const double *ez, const double *k2) {
    struct pb_coef c = {ex, ey, ez, k2};
    return cg_implicit(op_pb, &c, b, x, tol, n);
}

/* CG with an explicit Jacobi preconditioner v = -r/6 (precond_jacobi.c:8-14),
 * library/src/plcg.c:104-166. */
int or_conj_grad_jacobi(const double *b, double *x, double tol, int n) {
    const long G = (long)n * n * n, limit = (long)n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    double *v = (double *)malloc(G * sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    int res = -1;
    or_laplace(x, r, n);
    axpy(b, r, -1.0, G);
    for (long i = 0; i < G; i++) v[i] = -r[i] / 6.0;
    memcpy(p, v, G * sizeof(double));
    scal(p, -1.0, G);
    double r_dot_v = dot(r, v, G);
    for (long it = 0; it < limit; it++) {
        or_laplace(p, Ap, n);
        double alpha = r_dot_v / dot(p, Ap, G);
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        (void)dot(r, r, G);
        if (amax(r, G) <= tol) { res = (int)it; break; }
        for (long i = 0; i < G; i++) v[i] = -r[i] / 6.0;
        double rn_dot_vn = dot(r, v, G);
        double beta = rn_dot_vn / r_dot_v;
        r_dot_v = rn_dot_vn;
        scal(p, beta, G);
        axpy(v, p, -1.0, G);
    }
    free(r); free(Ap); free(v); free(p);
    return res;
}

/* Block-Jacobi preconditioner (library/src/precond_blockjacobi.c:14-69 and
 * :170-206): the grid is cut into 10^3 blocks (n % 10 == 0) and v = A_B^-1 r
 * on every block, A_B the 7-point Laplacian of one block with zero values
 * outside it (diagonal -6, +1 for each in-block neighbour).  The reference
 * inverts A_B with LAPACK's dgetrf / dgetri and applies it as one row-major
 * dgemm over the reordered blocks; here the inverse comes from Gauss-Jordan
 * elimination with partial pivoting, and each block's product sums over the
 * block's points in increasing order, as the loop form of the reference's
 * dgemm does (library/src/linalg.c:170-180).  Rows and columns follow the
 * reference's in-block order ii * 100 + jj * 10 + kk. */
#define OR_BJ 10
#define OR_BJ3 (OR_BJ * OR_BJ * OR_BJ)
static double *bj_inv = NULL;

static void bj_build(void) {
    const int B = OR_BJ, M = OR_BJ3;
    double *a = (double *)calloc((size_t)M * 2 * M, sizeof(double)); /* [A | I] */
    for (int i = 0; i < B; i++)
        for (int j = 0; j < B; j++)
            for (int k = 0; k < B; k++) {
                const int r = (i * B + j) * B + k;
                double *row = a + (size_t)r * 2 * M;
                row[r] = -6.0;
                if (i > 0) row[r - B * B] = 1.0;
                if (i < B - 1) row[r + B * B] = 1.0;
                if (j > 0) row[r - B] = 1.0;
                if (j < B - 1) row[r + B] = 1.0;
                if (k > 0) row[r - 1] = 1.0;
                if (k < B - 1) row[r + 1] = 1.0;
                row[M + r] = 1.0;
            }
    double *tmp = (double *)malloc((size_t)2 * M * sizeof(double));
    for (int c = 0; c < M; c++) {
        int piv = c;
        for (int r = c + 1; r < M; r++)
            if (fabs(a[(size_t)r * 2 * M + c]) > fabs(a[(size_t)piv * 2 * M + c])) piv = r;
        if (piv != c) {
            memcpy(tmp, a + (size_t)c * 2 * M, (size_t)2 * M * sizeof(double));
            memcpy(a + (size_t)c * 2 * M, a + (size_t)piv * 2 * M, (size_t)2 * M * sizeof(double));
            memcpy(a + (size_t)piv * 2 * M, tmp, (size_t)2 * M * sizeof(double));
        }
        double *pr = a + (size_t)c * 2 * M;
        const double d = 1.0 / pr[c];
        for (int q = 0; q < 2 * M; q++) pr[q] *= d;
        for (int r = 0; r < M; r++) {
            if (r == c) continue;
            double *rr = a + (size_t)r * 2 * M;
            const double f = rr[c];
            if (f == 0.0) continue;
            for (int q = c; q < 2 * M; q++) rr[q] -= f * pr[q];
        }
    }
    bj_inv = (double *)malloc((size_t)M * M * sizeof(double));
    for (int r = 0; r < M; r++) memcpy(bj_inv + (size_t)r * M, a + (size_t)r * 2 * M + M, (size_t)M * sizeof(double));
    free(tmp);
    free(a);
}

void or_blockjacobi_apply(const double *in, double *out, int n) {
    const int B = OR_BJ, M = OR_BJ3, nb = n / B;
    const long N = n, N2 = N * N;
    if (!bj_inv) bj_build();
    double blk[OR_BJ3];
    for (int bi = 0; bi < nb; bi++)
        for (int bj = 0; bj < nb; bj++)
            for (int bk = 0; bk < nb; bk++) {
                for (int l = 0; l < M; l++) {
                    const int ii = l / (B * B), jj = (l / B) % B, kk = l % B;
                    blk[l] = in[(bi * B + ii) * N2 + (bj * B + jj) * N + bk * B + kk];
                }
                for (int c = 0; c < M; c++) {
                    double s = 0.0;
                    for (int l = 0; l < M; l++) s += blk[l] * bj_inv[(size_t)l * M + c];
                    const int ii = c / (B * B), jj = (c / B) % B, kk = c % B;
                    out[(bi * B + ii) * N2 + (bj * B + jj) * N + bk * B + kk] = s;
                }
            }
}

/* CG with the block-Jacobi preconditioner, library/src/plcg.c:104-166. */
int or_conj_grad_bj(const double *b, double *x, double tol, int n) {
    const long G = (long)n * n * n, limit = (long)n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    double *v = (double *)malloc(G * sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    int res = -1;
    or_laplace(x, r, n);
    axpy(b, r, -1.0, G);
    or_blockjacobi_apply(r, v, n);
    memcpy(p, v, G * sizeof(double));
    scal(p, -1.0, G);
    double r_dot_v = dot(r, v, G);
    for (long it = 0; it < limit; it++) {
        or_laplace(p, Ap, n);
        double alpha = r_dot_v / dot(p, Ap, G);
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        (void)dot(r, r, G);
        if (amax(r, G) <= tol) { res = (int)it; break; }
        or_blockjacobi_apply(r, v, n);
        double rn_dot_vn = dot(r, v, G);
        double beta = rn_dot_vn / r_dot_v;
        r_dot_v = rn_dot_vn;
        scal(p, beta, G);
        axpy(v, p, -1.0, G);
    }
    free(r); free(Ap); free(v); free(p);
    return res;
}

/* CG with the reference's multigrid preconditioner: v = one two-grid V-cycle
 * with 5 RBGS sweeps per side started from the previous v
 * (library/src/precond_mg.c:3-5, multigrid.c:99-168 via multigrid_apply_recursive
 * :170-173, whose zeroing of `out` is commented out), inside the explicit
 * recurrence of library/src/plcg.c:104-166.  v holds the V-cycle's start and
 * keeps the previous v from one iteration to the next.  The reference takes
 * its first start from freshly malloc'd memory (mpi_grid_allocate,
 * library/src/mpi_base.c:223-233); here it starts from zero, which is what
 * the reference gets whenever malloc hands out fresh pages (the golden
 * vectors are made that way). */
int or_conj_grad_mg(const double *b, double *x, double tol, int n) {
    const long G = (long)n * n * n, limit = (long)n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    double *v = (double *)calloc(G, sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    int res = -1;
    or_laplace(x, r, n);
    axpy(b, r, -1.0, G);
    or_mg_vcycle(r, v, n, 5);
    memcpy(p, v, G * sizeof(double));
    scal(p, -1.0, G);
    double r_dot_v = dot(r, v, G);
    for (long it = 0; it < limit; it++) {
        or_laplace(p, Ap, n);
        double alpha = r_dot_v / dot(p, Ap, G);
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        (void)dot(r, r, G);
        if (amax(r, G) <= tol) { res = (int)it; break; }
        or_mg_vcycle(r, v, n, 5);
        double rn_dot_vn = dot(r, v, G);
        double beta = rn_dot_vn / r_dot_v;
        r_dot_v = rn_dot_vn;
        scal(p, beta, G);
        axpy(v, p, -1.0, G);
    }
    free(r); free(Ap); free(v); free(p);
    return res;
}

/* MaZe step: phi* = 2 phi - phi_prev, phi_prev <- phi; sigma = A phi* +
 * (4 pi / h) q; solve A y = sigma warm-started from y; phi = phi* - y.
 * library/src/verlet.c:31-71 (h already multiplied by eps_s by the caller,
 * library/src/grid_maze_lcg.c:158-164).  precond: 0 the implicit-Jacobi CG
 * (conj_grad), 1 the explicit Jacobi preconditioner, 2 the multigrid one,
 * 4 the block-Jacobi one (conj_grad_precond, library/src/grid_maze_lcg.c:130-145). */
int or_verlet_poisson(double tol, double h, double *phi, double *phi_prev, const double *q,
                      double *y, int n, int precond) {
    const long G = (long)n * n * n;
    double *sigma = (double *)malloc(G * sizeof(double));
    for (long i = 0; i < G; i++) {
        double a = phi[i];
        phi[i] = 2 * a - phi_prev[i];
        phi_prev[i] = a;
    }
    or_laplace(phi, sigma, n);
    axpy(q, sigma, (4 * OR_PI) / h, G);
    int it = precond == 4   ? or_conj_grad_bj(sigma, y, tol, n)
             : precond == 2 ? or_conj_grad_mg(sigma, y, tol, n)
             : precond == 1 ? or_conj_grad_jacobi(sigma, y, tol, n)
                            : or_conj_grad(sigma, y, tol, n);
    axpy(y, phi, -1.0, G);
    free(sigma);
    return it;
}

/* library/src/verlet.c:152-194 */
int or_verlet_poisson_pb(double tol, double h, double *phi, double *phi_prev, const double *q,
                         double *y, int n, const double *ex, const double *ey, const double *ez,
                         const double *k2) {
    const long G = (long)n * n * n;
    double *sigma = (double *)malloc(G * sizeof(double));
    for (long i = 0; i < G; i++) {
        double a = phi[i];
        phi[i] = 2 * a - phi_prev[i];
        phi_prev[i] = a;
    }
    or_laplace_pb(phi, sigma, n, ex, ey, ez, k2);
    axpy(q, sigma, (4 * OR_PI) / h, G);
    int it = or_conj_grad_pb(sigma, y, tol, n, ex, ey, ez, k2);
    axpy(y, phi, -1.0, G);
    free(sigma);
    return it;
}

/* ------------------------------------------------------------------------ */
/* Reference multigrid (library/src/multigrid.c), serial ranks, n even.     */

/* Red-black Gauss-Seidel, `iters` red+black pairs; red = (i+j+k) even.
 * library/src/multigrid.c:636-729.  Neighbour order i-1, i+1, j-1, j+1,
 * k-1, k+1 (:686-689). */
void or_mg_smooth_rbgs(const double *in, double *out, int n, int iters) {
    const long N = n, N2 = N * N;
    const double inv_diag = -1.0 / 6.0;
    for (int it = 0; it < iters; it++) {
        for (int color = 0; color < 2; color++) {
            for (long i = 0; i < N; i++) {
                const long i0 = i * N2, ip = wrapi(i + 1, N) * N2, im = wrapi(i - 1, N) * N2;
                for (long j = 0; j < N; j++) {
                    const long j0 = j * N, jp = wrapi(j + 1, N) * N, jm = wrapi(j - 1, N) * N;
                    for (long k = (i + j + color) & 1; k < N; k += 2) {
                        const long kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                        const long c = i0 + j0 + k;
                        double s = out[im + j0 + k] + out[ip + j0 + k] + out[i0 + jm + k] +
                                   out[i0 + jp + k] + out[i0 + j0 + km] + out[i0 + j0 + kp];
                        out[c] = (in[c] - s) * inv_diag;
                    }
                }
            }
        }
    }
}

/* 27-point nodal full weighting onto the (n/2)^3 grid, coarse (I,J,K) at
 * fine (2I,2J,2K); weights 8/4/2/1 over 64.  library/src/multigrid.c:523-594
 * (sum order :571-590). */
void or_mg_restrict27(const double *in, double *out, int n) {
    const long N = n, N2 = N * N, M = N / 2;
#define F(a, b, c) in[(a) * N2 + (b) * N + (c)]
    for (long I = 0; I < M; I++) {
        const long i = 2 * I, ip = wrapi(i + 1, N), im = wrapi(i - 1, N);
        for (long J = 0; J < M; J++) {
            const long j = 2 * J, jp = wrapi(j + 1, N), jm = wrapi(j - 1, N);
            for (long K = 0; K < M; K++) {
                const long k = 2 * K, kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                double f = F(im, j, k) + F(ip, j, k) + F(i, jm, k) + F(i, jp, k) + F(i, j, km) + F(i, j, kp);
                double e = F(ip, j, km) + F(ip, j, kp) + F(im, j, km) + F(im, j, kp) + F(ip, jm, k) +
                           F(ip, jp, k) + F(im, jm, k) + F(im, jp, k) + F(i, jm, km) + F(i, jm, kp) +
                           F(i, jp, km) + F(i, jp, kp);
                double c = F(ip, jp, km) + F(ip, jp, kp) + F(ip, jm, km) + F(ip, jm, kp) + F(im, jp, km) +
                           F(im, jp, kp) + F(im, jm, km) + F(im, jm, kp);
                out[(I * M + J) * M + K] = (8.0 * F(i, j, k) + 4.0 * f + 2.0 * e + c) * (1.0 / 64.0);
            }
        }
    }
#undef F
}

/* Trilinear prolongation from an nc^3 grid to a (2 nc)^3 grid (overwrites
 * out).  library/src/multigrid.c:393-463 (term order :439-447). */
void or_mg_prolong_trilinear(const double *in, double *out, int nc) {
    const long M = nc, M2 = M * M, N = 2 * M;
    for (long i = 0; i < N; i++) {
        const long a0 = (i / 2) % M, a1 = (a0 + 1) % M;
        const double wi0 = (i & 1) ? 0.5 : 1.0, wi1 = (i & 1) ? 0.5 : 0.0;
        for (long j = 0; j < N; j++) {
            const long b0 = (j / 2) % M, b1 = (b0 + 1) % M;
            const double wj0 = (j & 1) ? 0.5 : 1.0, wj1 = (j & 1) ? 0.5 : 0.0;
            for (long k = 0; k < N; k++) {
                const long c0 = (k / 2) % M, c1 = (c0 + 1) % M;
                const double wk0 = (k & 1) ? 0.5 : 1.0, wk1 = (k & 1) ? 0.5 : 0.0;
#define C(a, b, c) in[(a) * M2 + (b) * M + (c)]
                out[(i * N + j) * N + k] =
                    C(a0, b0, c0) * wi0 * wj0 * wk0 + C(a1, b0, c0) * wi1 * wj0 * wk0 +
                    C(a0, b1, c0) * wi0 * wj1 * wk0 + C(a1, b1, c0) * wi1 * wj1 * wk0 +
                    C(a0, b0, c1) * wi0 * wj0 * wk1 + C(a1, b0, c1) * wi1 * wj0 * wk1 +
                    C(a0, b1, c1) * wi0 * wj1 * wk1 + C(a1, b1, c1) * wi1 * wj1 * wk1;
#undef C
            }
        }
    }
}

/* Coarse-grid CG with a relative infinity-norm stop and a breakdown guard.
 * Returns the number of iterations performed.  library/src/multigrid.c:13-96 */
int or_mg_cg_coarse(const double *b, double *x, int n, int maxit, double rtol) {
    const long G = (long)n * n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    int k;
    or_laplace(x, r, n);
    axpy(b, r, -1.0, G);
    memcpy(p, r, G * sizeof(double));
    scal(p, 1.0 / 6.0, G);
    double r0 = amax(r, G);
    if (r0 == 0.0) { free(r); free(p); free(Ap); return 0; }
    double r_dot_v = -dot(r, p, G);
    for (k = 0; k < maxit; ++k) {
        or_laplace(p, Ap, n);
        double denom = dot(p, Ap, G);
        if (fabs(denom) < 1e-300) break;
        double alpha = r_dot_v / denom;
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        if (amax(r, G) <= rtol * r0) { ++k; break; }
        double rn = dot(r, r, G);
        double rn_dot_vn = -rn / 6.0;
        double beta = rn_dot_vn / r_dot_v;
        r_dot_v = rn_dot_vn;
        scal(p, beta, G);
        axpy(r, p, 1.0 / 6.0, G);
    }
    free(r); free(p); free(Ap);
    return k;
}

/* One reference V-cycle: the recursion stops at depth 1, so it is a
 * two-grid cycle with `sm` RBGS pre/post sweeps and a coarse CG (<= 50
 * iterations, rtol 1e-5) on the unscaled coarse stencil.
 * library/src/multigrid.c:99-168, selectors :743-765. */
void or_mg_vcycle(const double *in, double *out, int n, int sm) {
    const long G = (long)n * n * n;
    const int nc = n / 2;
    const long Gc = (long)nc * nc * nc;
    double *r = (double *)malloc(G * sizeof(double));
    double *rhs = (double *)malloc(Gc * sizeof(double));
    double *eps = (double *)calloc(Gc, sizeof(double));
    or_mg_smooth_rbgs(in, out, n, sm);
    or_laplace(out, r, n);
    scal(r, -1.0, G);
    axpy(in, r, 1.0, G);
    or_mg_restrict27(r, rhs, n);
    or_mg_cg_coarse(rhs, eps, nc, 50, 1e-5);
    or_mg_prolong_trilinear(eps, r, nc);
    axpy(r, out, 1.0, G);
    or_mg_smooth_rbgs(in, out, n, sm);
    free(r); free(rhs); free(eps);
}

/* MaZe step solved with repeated V-cycles, stop on the true residual
 * max|A y - sigma| <= tol; returns the 1-based cycle count or -1 after 1000.
 * library/src/verlet.c:86-149. */
int or_verlet_poisson_mg(double tol, double h, double *phi, double *phi_prev, const double *q,
                         double *y, int n) {
    const long G = (long)n * n * n;
    double *tmp = (double *)malloc(G * sizeof(double));
    double *sigma = (double *)malloc(G * sizeof(double));
    int res = -1, it = 0;
    for (long i = 0; i < G; i++) {
        double a = phi[i];
        phi[i] = 2 * a - phi_prev[i];
        phi_prev[i] = a;
    }
    or_laplace(phi, sigma, n);
    axpy(q, sigma, (4 * OR_PI) / h, G);
    while (it < 1000) {
        or_mg_vcycle(sigma, y, n, 3);
        or_laplace(y, tmp, n);
        axpy(sigma, tmp, -1.0, G);
        double a = amax(tmp, G);
        it++;
        if (a <= tol) { res = it; break; }
    }
    axpy(y, phi, -1.0, G);
    free(tmp); free(sigma);
    return res;
}

/* ---- Poisson-Boltzmann multigrid ---------------------------------------- */

/* RBGS for A_pb u = in: each colour sets u = (sum of eps_f u_nb - in) / diag,
 * diag = sum of the six face coefficients + k2 (floored at 1e-14), red
 * ((i+j+k) even) then black, `iters` times, in place.
 * library/src/multigrid_pb.c:596-690 (term order :637-645). */
void or_mg_smooth_pb_rbgs(const double *in, double *out, int n, int iters, const double *ex,
                          const double *ey, const double *ez, const double *k2) {
    const long N = n, N2 = N * N;
    for (int it = 0; it < iters; it++) {
        for (int color = 0; color < 2; color++) {
            for (long i = 0; i < N; i++) {
                const long i0 = i * N2, ip = wrapi(i + 1, N) * N2, im = wrapi(i - 1, N) * N2;
                for (long j = 0; j < N; j++) {
                    const long j0 = j * N, jp = wrapi(j + 1, N) * N, jm = wrapi(j - 1, N) * N;
                    for (long k = (i + j + color) & 1; k < N; k += 2) {
                        const long kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                        const long c = i0 + j0 + k, cx = im + j0 + k, cy = i0 + jm + k,
                                   cz = i0 + j0 + km;
                        const double ex_p = ex[c], ex_m = ex[cx], ey_p = ey[c], ey_m = ey[cy];
                        const double ez_p = ez[c], ez_m = ez[cz];
                        double diag = ex_p + ex_m + ey_p + ey_m + ez_p + ez_m + k2[c];
                        if (diag < 1e-14) diag = 1e-14;
                        const double rhs = ex_p * out[ip + j0 + k] + ex_m * out[cx] +
                                           ey_p * out[i0 + jp + k] + ey_m * out[cy] +
                                           ez_p * out[i0 + j0 + kp] + ez_m * out[cz] - in[c];
                        out[c] = rhs / diag;
                    }
                }
            }
        }
    }
}

/* Face-centred coefficient onto the (n/2)^3 grid: 2x2 average over the plane
 * normal to `axis` (0: x faces on fine plane 2I; 1, 2: fine planes 2I and
 * 2I+1).  library/src/multigrid_pb.c:375-500 (sum order per axis). */
void or_mg_restrict_eps(const double *in, double *out, int n, int axis) {
    const long N = n, N2 = N * N, M = N / 2;
    for (long I = 0; I < M; I++) {
        const long i0 = 2 * I * N2, i1 = i0 + N2;
        for (long J = 0; J < M; J++) {
            const long j0 = 2 * J * N, j1 = wrapi(2 * J + 1, N) * N;
            for (long K = 0; K < M; K++) {
                const long k0 = 2 * K, k1 = wrapi(2 * K + 1, N);
                double s;
                if (axis == 0)
                    s = in[i0 + j0 + k0] + in[i0 + j0 + k1] + in[i0 + j1 + k0] + in[i0 + j1 + k1];
                else if (axis == 1)
                    s = in[i0 + j0 + k0] + in[i0 + j0 + k1] + in[i1 + j0 + k0] + in[i1 + j0 + k1];
                else
                    s = in[i0 + j0 + k0] + in[i1 + j0 + k0] + in[i0 + j1 + k0] + in[i1 + j1 + k0];
                out[(I * M + J) * M + K] = s * 0.25;
            }
        }
    }
}

/* Cell-centred screening term onto the (n/2)^3 grid: 8-point average of the
 * fine cells (2I+1, 2I+2) x (2J, 2J+1) x (2K, 2K+1) (the serial i offset is
 * 1).  library/src/multigrid_pb.c:503-568. */
void or_mg_restrict_k2(const double *in, double *out, int n) {
    const long N = n, N2 = N * N, M = N / 2;
    for (long i = 1; i < N; i += 2) {
        const long i0 = i * N2, i1 = wrapi(i + 1, N) * N2, a = (i / 2) * M * M;
        for (long j = 0; j < N; j += 2) {
            const long j0 = j * N, j1 = wrapi(j + 1, N) * N, b = (j / 2) * M;
            for (long k = 0; k < N; k += 2) {
                const long k1 = wrapi(k + 1, N);
                const double sum = in[i0 + j0 + k] + in[i0 + j0 + k1] + in[i0 + j1 + k] +
                                   in[i0 + j1 + k1] + in[i1 + j0 + k] + in[i1 + j0 + k1] +
                                   in[i1 + j1 + k] + in[i1 + j1 + k1];
                out[a + b + k / 2] = (1.0 / 8.0) * sum;
            }
        }
    }
}

/* Unpreconditioned coarse CG for A_pb x = b (p = -r, relative
 * infinity-norm stop, breakdown guard); returns the iterations performed.
 * library/src/multigrid_pb.c:16-98 */
int or_mg_cg_coarse_pb(const double *b, double *x, int n, int maxit, double rtol,
                       const double *ex, const double *ey, const double *ez, const double *k2) {
    const long G = (long)n * n * n;
    double *r = (double *)malloc(G * sizeof(double));
    double *p = (double *)malloc(G * sizeof(double));
    double *Ap = (double *)malloc(G * sizeof(double));
    int k;
    or_laplace_pb(x, r, n, ex, ey, ez, k2);
    axpy(b, r, -1.0, G);
    memcpy(p, r, G * sizeof(double));
    scal(p, -1.0, G);
    double r0 = amax(r, G);
    if (r0 == 0.0) { free(r); free(p); free(Ap); return 0; }
    double rr = dot(r, r, G);
    for (k = 0; k < maxit; ++k) {
        or_laplace_pb(p, Ap, n, ex, ey, ez, k2);
        double denom = dot(p, Ap, G);
        if (fabs(denom) < 1e-300) break;
        double alpha = rr / denom;
        axpy(p, x, alpha, G);
        axpy(Ap, r, alpha, G);
        if (amax(r, G) <= rtol * r0) { ++k; break; }
        double rn = dot(r, r, G);
        double beta = rn / rr;
        rr = rn;
        scal(p, beta, G);
        axpy(r, p, -1.0, G);
    }
    free(r); free(p); free(Ap);
    return k;
}

/* One reference PB V-cycle: two-grid, `sm` RBGS-PB pre/post sweeps, 27-point
 * restriction of the residual, coarse coefficients by restriction_eps /
 * restriction_k2screen, coarse CG-PB (<= 50 iterations, rtol 1e-5), trilinear
 * prolongation.  library/src/multigrid_pb.c:275-355. */
void or_mg_vcycle_pb(const double *in, double *out, int n, int sm, const double *ex,
                     const double *ey, const double *ez, const double *k2) {
    const long G = (long)n * n * n;
    const int nc = n / 2;
    const long Gc = (long)nc * nc * nc;
    double *r = (double *)malloc(G * sizeof(double));
    double *rhs = (double *)malloc(Gc * sizeof(double));
    double *e = (double *)calloc(Gc, sizeof(double));
    double *cx = (double *)malloc(Gc * sizeof(double));
    double *cy = (double *)malloc(Gc * sizeof(double));
    double *cz = (double *)malloc(Gc * sizeof(double));
    double *ck = (double *)malloc(Gc * sizeof(double));
    or_mg_smooth_pb_rbgs(in, out, n, sm, ex, ey, ez, k2);
    or_laplace_pb(out, r, n, ex, ey, ez, k2);
    scal(r, -1.0, G);
    axpy(in, r, 1.0, G);
    or_mg_restrict27(r, rhs, n);
    or_mg_restrict_eps(ex, cx, n, 0);
    or_mg_restrict_eps(ey, cy, n, 1);
    or_mg_restrict_eps(ez, cz, n, 2);
    or_mg_restrict_k2(k2, ck, n);
    or_mg_cg_coarse_pb(rhs, e, nc, 50, 1e-5, cx, cy, cz, ck);
    or_mg_prolong_trilinear(e, r, nc);
    axpy(r, out, 1.0, G);
    or_mg_smooth_pb_rbgs(in, out, n, sm, ex, ey, ez, k2);
    free(r); free(rhs); free(e); free(cx); free(cy); free(cz); free(ck);
}

/* MaZe PB step with repeated PB V-cycles (4 sweeps), stop on max|A_pb y -
 * sigma| <= tol; returns the 1-based cycle count or -1 after 1000.
 * library/src/verlet.c:211-271. */
int or_verlet_pb_multigrid(double tol, double h, double *phi, double *phi_prev, const double *q,
                           double *y, int n, const double *ex, const double *ey,
                           const double *ez, const double *k2) {
    const long G = (long)n * n * n;
    double *tmp = (double *)malloc(G * sizeof(double));
    double *sigma = (double *)malloc(G * sizeof(double));
    int res = -1, it = 0;
    for (long i = 0; i < G; i++) {
        double a = phi[i];
        phi[i] = 2 * a - phi_prev[i];
        phi_prev[i] = a;
    }
    or_laplace_pb(phi, sigma, n, ex, ey, ez, k2);
    axpy(q, sigma, (4 * OR_PI) / h, G);
    while (it < 1000) {
        or_mg_vcycle_pb(sigma, y, n, 4, ex, ey, ez, k2);
        or_laplace_pb(y, tmp, n, ex, ey, ez, k2);
        axpy(sigma, tmp, -1.0, G);
        double a = amax(tmp, G);
        it++;
        if (a <= tol) { res = it; break; }
    }
    axpy(y, phi, -1.0, G);
    free(tmp); free(sigma);
    return res;
}

/* Non-MaZe PB multigrid update: repeated PB V-cycles on A_pb phi = b from the
 * previous phi, stop on max|A_pb phi - b| <= tol.
 * library/src/grid_multigrid.c:149-175. */
int or_mg_pb_update(double tol, const double *b, double *phi, int n, const double *ex,
                    const double *ey, const double *ez, const double *k2) {
    const long G = (long)n * n * n;
    double *tmp = (double *)malloc(G * sizeof(double));
    int res = -1, it = 0;
    while (it < 1000) {
        or_mg_vcycle_pb(b, phi, n, 4, ex, ey, ez, k2);
        or_laplace_pb(phi, tmp, n, ex, ey, ez, k2);
        axpy(b, tmp, -1.0, G);
        double a = amax(tmp, G);
        it++;
        if (a <= tol) { res = it; break; }
    }
    free(tmp);
    return res;
}

/* ------------------------------------------------------------------------ */
/* Particles                                                                */

/* CIC weight with the periodic distance, Im et al. (1998) eq. 24.
 * library/src/charges.c:8-14 */
double or_cic_weight(double x, double L, double h) {
    x = fabs(x - round(x / L) * L);
    if (x >= h) return 0.0;
    return 1.0 - x / h;
}

/* Quadratic and cubic B-spline weights.  library/src/charges.c:16-38 */
double or_spline_quadr_weight(double x, double L, double h) {
    x = fabs(x - round(x / L) * L) / h;
    if (x < 0.5) return 0.75 - x * x;
    if (x < 1.5) return 0.5 * pow((1.5 - x), 2);
    return 0.0;
}

double or_spline_cubic_weight(double x, double L, double h) {
    x = fabs(x - round(x / L) * L) / h;
    if (x < 1) return (4.0 - 6.0 * pow(x, 2) + 3.0 * pow(x, 3)) / 6.0;
    if (x < 2) return pow(2 - x, 3) / 6.0;
    return 0.0;
}

/* Charge-assignment scheme: 0 CIC (8 nodes), 1 quadratic, 2 cubic spline (64
 * nodes).  library/src/particles.c:38-66 */
int or_cas_nodes(int cas) { return cas == 0 ? 8 : 64; }

static double cas_weight(int cas, double x, double L, double h) {
    if (cas == 1) return or_spline_quadr_weight(x, L, h);
    if (cas == 2) return or_spline_cubic_weight(x, L, h);
    return or_cic_weight(x, L, h);
}

/* Base cell floor(pos/h) and the assignment nodes, stored (n_p, nodes, 3):
 * CIC in the reference's corner order 000,100,010,001,110,101,011,111
 * (library/src/particles.c:297-346); splines the 4x4x4 block base + {-1..2}^3,
 * last offset fastest (:348-385).  The base index is wrapped into [0,n): the
 * reference leaves it unwrapped (:315-317, :367-369), which only matters for a
 * position that rounds to pos/h >= n, where the reference would index past
 * the grid. */
void or_neighbors(int cas, const double *pos, int np, int n, double h, int64_t *nb) {
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const int nn = or_cas_nodes(cas);
    for (int i = 0; i < np; i++) {
        int base[3];
        for (int d = 0; d < 3; d++) {
            int c = (int)floor(pos[3 * i + d] / h);
            c %= n;
            if (c < 0) c += n;
            base[d] = c;
        }
        int64_t *o = nb + (long)i * nn * 3;
        if (cas == 0) {
            for (int c = 0; c < 8; c++)
                for (int d = 0; d < 3; d++) o[c * 3 + d] = (base[d] + off[c][d]) % n;
        } else {
            for (int a = -1; a <= 2; a++)
                for (int b = -1; b <= 2; b++)
                    for (int c = -1; c <= 2; c++) {
                        o[0] = (base[0] + a + n) % n;
                        o[1] = (base[1] + b + n) % n;
                        o[2] = (base[2] + c + n) % n;
                        o += 3;
                    }
        }
    }
}

void or_cic_neighbors(const double *pos, int np, int n, double h, int64_t *nb) {
    or_neighbors(0, pos, np, n, h, nb);
}

/* Scatter q_ion * wx * wy * wz to the assignment nodes, serial ion order;
 * returns the sum of all contributions.  The period used inside the weight is
 * n*h.  library/src/charges.c:100-144 (:113, :130-140). */
double or_spread(int cas, int n, int np, double h, const double *pos, const int64_t *nb,
                 const double *charges, double *q) {
    const long N = n, N2 = N * N, G = N2 * N;
    const double L = n * h;
    const int nn = or_cas_nodes(cas);
    double q_tot = 0.0;
    for (long i = 0; i < G; i++) q[i] = 0.0;
    for (long i = 0; i < np; i++) {
        const double chg = charges[i], px = pos[3 * i], py = pos[3 * i + 1], pz = pos[3 * i + 2];
        for (int c = 0; c < nn; c++) {
            const int64_t *t = nb + (i * nn + c) * 3;
            double app = cas_weight(cas, px - t[0] * h, L, h) *
                         cas_weight(cas, py - t[1] * h, L, h) *
                         cas_weight(cas, pz - t[2] * h, L, h);
            double upd = chg * app;
            q_tot += upd;
            q[t[0] * N2 + t[1] * N + t[2]] += upd;
        }
    }
    return q_tot;
}

double or_spread_cic(int n, int np, double h, const double *pos, const int64_t *nb,
                     const double *charges, double *q) {
    return or_spread(0, n, np, h, pos, nb, charges, q);
}

/* Field force by weighted central differences at the assignment nodes;
 * returns the sum of the weighted charges.  library/src/forces.c:26-100. */
double or_force_fd_cas(int cas, int n, int np, double h, const double *phi, const int64_t *nb,
                       const double *charges, const double *pos, double *forces) {
    const long N = n, N2 = N * N;
    const double h2 = 2.0 * h, L = N * h;
    const int nn = or_cas_nodes(cas);
    double sum_q = 0.0;
    for (long ip = 0; ip < np; ip++) {
        double *f = forces + 3 * ip;
        const double px = pos[3 * ip], py = pos[3 * ip + 1], pz = pos[3 * ip + 2], chg = charges[ip];
        f[0] = f[1] = f[2] = 0.0;
        for (int c = 0; c < nn; c++) {
            const int64_t *t = nb + (ip * nn + c) * 3;
            const long i = t[0], j = t[1], k = t[2];
            double qc = chg * cas_weight(cas, px - i * h, L, h) * cas_weight(cas, py - j * h, L, h) *
                        cas_weight(cas, pz - k * h, L, h);
            sum_q += qc;
            double E = (phi[wrapi(i - 1, N) * N2 + j * N + k] - phi[wrapi(i + 1, N) * N2 + j * N + k]) / h2;
            f[0] += qc * E;
            E = (phi[i * N2 + wrapi(j - 1, N) * N + k] - phi[i * N2 + wrapi(j + 1, N) * N + k]) / h2;
            f[1] += qc * E;
            E = (phi[i * N2 + j * N + wrapi(k - 1, N)] - phi[i * N2 + j * N + wrapi(k + 1, N)]) / h2;
            f[2] += qc * E;
        }
    }
    return sum_q;
}

double or_force_fd(int n, int np, double h, const double *phi, const int64_t *nb,
                   const double *charges, const double *pos, double *forces) {
    return or_force_fd_cas(0, n, np, h, phi, nb, charges, pos, forces);
}

/* Shifted-force Tosi-Fumi constants per type pair: table[(t1*n_typ+t2)*7 +
 * {A,B,C,D,sigma,alpha,beta}], alpha = -V'(rc), beta = -V(rc) - alpha rc.
 * library/src/particles.c:181-228 (expressions :191-215). */
void or_tf_pair_table(int n_typ, const double *pp, double r_cut, double *table) {
    double rc6 = pow(r_cut, 6), rc7 = rc6 * r_cut, rc8 = rc7 * r_cut, rc9 = rc8 * r_cut;
    for (int t = 0; t < n_typ * n_typ; t++) {
        const double A = pp[t * 5], B = pp[t * 5 + 1], C = pp[t * 5 + 2], D = pp[t * 5 + 3],
                     s = pp[t * 5 + 4];
        double v_shift = A * exp(B * (s - r_cut)) - C / rc6 - D / rc8;
        double alpha = A * B * exp(B * (s - r_cut)) - 6 * C / rc7 - 8 * D / rc9;
        double beta = -v_shift - alpha * r_cut;
        double *o = table + t * 7;
        o[0] = A; o[1] = B; o[2] = C; o[3] = D; o[4] = s; o[5] = alpha; o[6] = beta;
    }
}

/* Minimum-image separation, accumulated x, y, z as in forces.c:145-157. */
static inline double min_image_r(const double *a, const double *b, double L, double d[3]) {
    double app = a[0] - b[0];
    app -= L * round(app / L);
    double r2 = app * app;
    d[0] = app;
    app = a[1] - b[1];
    app -= L * round(app / L);
    d[1] = app;
    r2 += app * app;
    app = a[2] - b[2];
    app -= L * round(app / L);
    d[2] = app;
    r2 += app * app;
    return sqrt(r2);
}

/* All-pairs shifted-force Tosi-Fumi, r <= r_cut, full j loop per i; returns
 * sum V / 2.  library/src/forces.c:112-186. */
double or_tf_forces(int np, double L, const double *pos, const int *types, int n_typ,
                    const double *table, double r_cut, double *forces) {
    double pe = 0.0;
    for (int i = 0; i < np; i++) {
        double *f = forces + 3 * i;
        f[0] = f[1] = f[2] = 0.0;
        for (int j = 0; j < np; j++) {
            if (i == j) continue;
            double d[3];
            double r = min_image_r(pos + 3 * i, pos + 3 * j, L, d);
            if (r > r_cut) continue;
            d[0] /= r; d[1] /= r; d[2] /= r;
            const double *t = table + (types[i] * n_typ + types[j]) * 7;
            const double a = t[0], b = t[1], c = t[2], dd = t[3], s = t[4], al = t[5], be = t[6];
            double fm = b * a * exp(b * (s - r)) - 6 * c / pow(r, 7) - 8 * dd / pow(r, 9) - al;
            double vm = a * exp(b * (s - r)) - c / pow(r, 6) - dd / pow(r, 8) + al * r + be;
            f[0] += fm * d[0];
            f[1] += fm * d[1];
            f[2] += fm * d[2];
            pe += vm;
        }
    }
    return pe / 2;
}

/* Number of j != i with r <= r_cut under the same predicate as
 * or_tf_forces (the pair list the reference evaluates, forces.c:158). */
void or_tf_neighbor_counts(int np, double L, const double *pos, double r_cut, int *counts) {
    for (int i = 0; i < np; i++) {
        int c = 0;
        for (int j = 0; j < np; j++) {
            if (i == j) continue;
            double d[3];
            if (min_image_r(pos + 3 * i, pos + 3 * j, L, d) <= r_cut) c++;
        }
        counts[i] = c;
    }
}

/* K = 1/2 sum m |v|^2, serial order.  library/src/particles.c:802-820 */
double or_kinetic(int np, const double *vel, const double *mass) {
    double kin = 0.0;
    for (int i = 0; i < np; i++) {
        double app = 0.0;
        for (int d = 0; d < 3; d++) app += pow(vel[3 * i + d], 2);
        kin += mass[i] * app;
    }
    return 0.5 * kin;
}

/* ------------------------------------------------------------------------ */
/* Philox4x32-10 (Salmon, Moraes, Dror, Shaw, SC'11) and Box-Muller.  The
 * reference draws its Langevin noise from rand() with a racy static cache
 * (library/src/integrator_ovrvo.c:12-36), which no two runs reproduce; this
 * build replaces it by a counter-based stream keyed by (seed, call, ion) so
 * that CPU and GPU draw the same deviates. */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]) {
    uint32_t c0 = ctr[0], c1 = ctr[1], c2 = ctr[2], c3 = ctr[3];
    uint32_t k0 = key[0], k1 = key[1];
    for (int r = 0; r < 10; r++) {
        uint64_t p0 = (uint64_t)0xD2511F53u * c0;
        uint64_t p1 = (uint64_t)0xCD9E8D57u * c2;
        uint32_t hi0 = (uint32_t)(p0 >> 32), lo0 = (uint32_t)p0;
        uint32_t hi1 = (uint32_t)(p1 >> 32), lo1 = (uint32_t)p1;
        uint32_t n0 = hi1 ^ c1 ^ k0, n2 = hi0 ^ c3 ^ k1;
        c0 = n0; c1 = lo1; c2 = n2; c3 = lo0;
        k0 += 0x9E3779B9u; k1 += 0xBB67AE85u;
    }
    out[0] = c0; out[1] = c1; out[2] = c2; out[3] = c3;
}

/* Three standard normal deviates for (seed, call, ion).  Uniforms are
 * (x + 0.5) 2^-32 in (0,1); pairs (u0,u1) and (u2,u3) go through
 * Box-Muller; the fourth deviate is unused. */
void or_normal3(uint64_t seed, uint64_t call, uint32_t ion, double out[3]) {
    uint32_t ctr[4] = {ion, (uint32_t)call, (uint32_t)(call >> 32), 0x4D415A45u};
    uint32_t key[2] = {(uint32_t)seed, (uint32_t)(seed >> 32)};
    uint32_t o[4];
    or_philox4x32(ctr, key, o);
    const double s = 2.3283064365386963e-10; /* 2^-32 */
    double u0 = ((double)o[0] + 0.5) * s, u1 = ((double)o[1] + 0.5) * s;
    double u2 = ((double)o[2] + 0.5) * s, u3 = ((double)o[3] + 0.5) * s;
    double ra = sqrt(-2.0 * log(u0)), rb = sqrt(-2.0 * log(u2));
    double ta = 6.283185307179586 * u1, tb = 6.283185307179586 * u3;
    out[0] = ra * cos(ta);
    out[1] = ra * sin(ta);
    out[2] = rb * cos(tb);
}

/* ---- Poisson-Boltzmann boundary from ion radii ----------------------------
 * Each ion carries a solvation radius r (plus the probe radius, added by the
 * driver) and a transition half-width w.  The smooth step
 *   H(s) = -(1/(4 w^3)) s^3 + (3/(4 w^2)) s^2,   s = |x - x_ion| - r + w,
 * rises from 0 at distance r - w to 1 at r + w; inside r - w the solvent is
 * excluded (factor 0) and beyond r + w the ion leaves the point unchanged.
 */
static inline double pb_step(double s, double w2, double w3) {
    return -(1 / (4 * w3)) * pow(s, 3) + (3 / (4 * w2)) * pow(s, 2);
}

/* d ln H / ds, library/src/particles.c:407-417 */
static inline double pb_step_dlog(double s, double w2, double w3) {
    return (-(3 / (4 * w3)) * pow(s, 2) + (3 / (2 * w2)) * s) /
           (-(1 / (4 * w3)) * pow(s, 3) + (3 / (4 * w2)) * pow(s, 2));
}

/* Multiply v by the ion's mask at squared distance r2 (a point outside the
 * shell keeps v, one inside the core becomes 0). */
static inline void pb_mask(double *v, double r2, double rp2, double rm2, double r, double w,
                           double w2, double w3) {
    if (r2 >= rp2) return;
    if (r2 > rm2) *v *= pb_step(sqrt(r2) - r + w, w2, w3);
    else *v = 0.0;
}

/* library/src/grid_general.c:138-301: the face dielectric eps_x/y/z (face
 * i+1/2 of point i along each axis) and the screening map k2 (at the point)
 * are eps_int + (eps_s - eps_int) * prod_ions H and kbar2 * prod_ions H, the
 * products taken in ion order over each ion's box of +-(floor((r+w)/h)+1)
 * points around its cell. */
void or_pb_maps(int n, double h, double w, double eps_s, double eps_int, double kbar2, int np,
                const double *pos, const double *radii, double *ex, double *ey, double *ez,
                double *k2) {
    const long N = n, N2 = N * N, G = N2 * N;
    const double w2 = w * w, w3 = w2 * w, hd2 = h / 2.0;
    for (long i = 0; i < G; i++) {
        ex[i] = eps_s - eps_int;
        ey[i] = eps_s - eps_int;
        ez[i] = eps_s - eps_int;
        k2[i] = kbar2;
    }
    for (int a = 0; a < np; a++) {
        const double r = radii[a];
        const double px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
        const double rp2 = pow(r + w, 2), rm2 = pow(r - w, 2);
        const int R = (int)floor((r + w) / h) + 1;
        const int cx = (int)floor(px / h), cy = (int)floor(py / h), cz = (int)floor(pz / h);
        for (int di = -R; di <= R; di++) {
            const int ii = cx + di;
            const double dx = px - ii * h, dx2 = dx * dx;
            const long bi = ((ii + n) % n) * N2;
            for (int dj = -R; dj <= R; dj++) {
                const int jj = cy + dj;
                const double dy = py - jj * h, dy2 = dy * dy;
                const long bj = bi + ((jj + n) % n) * N;
                for (int dk = -R; dk <= R; dk++) {
                    const int kk = cz + dk;
                    const double dz = pz - kk * h, dz2 = dz * dz;
                    const long c = bj + (kk + n) % n;
                    double t;
                    pb_mask(&k2[c], dx2 + dy2 + dz2, rp2, rm2, r, w, w2, w3);
                    t = dx - hd2;
                    pb_mask(&ex[c], t * t + dy2 + dz2, rp2, rm2, r, w, w2, w3);
                    t = dy - hd2;
                    pb_mask(&ey[c], dx2 + t * t + dz2, rp2, rm2, r, w, w2, w3);
                    t = dz - hd2;
                    pb_mask(&ez[c], dx2 + dy2 + t * t, rp2, rm2, r, w, w2, w3);
                }
            }
        }
    }
    for (long i = 0; i < G; i++) {
        ex[i] += eps_int;
        ey[i] += eps_int;
        ez[i] += eps_int;
    }
}

/* One face term of the boundary forces: the face at offset (fx, fy, fz) from
 * the ion, coefficient eps_f - eps_int of that face, the potential step
 * across it and, for the non-polar term, the normalised dielectric gradient
 * component g; `sgn` is -1 for the face below the point, +1 above. */
static inline void pb_face(double fx, double fy, double fz, double rp2, double rm2, double r,
                           double w, double w2, double w3, double eps_f, double eps_int,
                           double dphi, double phic, double g, int sgn, int nonpolar,
                           double *db, double *npf) {
    const double r2 = fx * fx + fy * fy + fz * fz;
    if (r2 >= rp2 || r2 <= rm2) return;
    const double r1 = sqrt(r2) + 1E-12;
    const double ux = fx / r1, uy = fy / r1, uz = fz / r1;
    const double c = (eps_f - eps_int) * pb_step_dlog(r1 - r + w, w2, w3);
    db[0] -= c * dphi * phic * ux;
    db[1] -= c * dphi * phic * uy;
    db[2] -= c * dphi * phic * uz;
    if (nonpolar) {
        if (sgn < 0) {
            npf[0] -= c * g * ux;
            npf[1] -= c * g * uy;
            npf[2] -= c * g * uz;
        } else {
            npf[0] += c * g * ux;
            npf[1] += c * g * uy;
            npf[2] += c * g * uz;
        }
    }
}

/* library/src/particles.c:420-775: dielectric-boundary (db), ionic-boundary
 * (ib) and non-polar (np) forces from the field phi and the maps, summed per
 * ion over its box of +-(floor((r+w)/h)+2) points in (i, j, k) order, then
 * scaled by h/(8 pi), h/(8 pi) and -gamma_np h/(eps_s - eps_int).  Returns
 * the non-polar energy, which the reference leaves at 0. */
double or_pb_forces(int n, double h, double w, double eps_s, double eps_int, int nonpolar,
                    double gamma_np, int np, const double *pos, const double *radii,
                    const double *phi, const double *ex, const double *ey, const double *ez,
                    const double *k2, double *fdb, double *fib, double *fnp) {
    const long N = n, N2 = N * N;
    const double w2 = w * w, w3 = w2 * w, hd2 = h / 2.0;
    for (long i = 0; i < 3L * np; i++) fdb[i] = fib[i] = fnp[i] = 0.0;
    for (int a = 0; a < np; a++) {
        const double r = radii[a];
        const double px = pos[3 * a], py = pos[3 * a + 1], pz = pos[3 * a + 2];
        const double rp2 = pow(r + w, 2), rm2 = pow(r - w, 2);
        const int R = (int)floor((r + w) / h) + 2;
        const int cx = (int)floor(px / h), cy = (int)floor(py / h), cz = (int)floor(pz / h);
        double *db = fdb + 3 * a, *ib = fib + 3 * a, *nf = fnp + 3 * a;
        for (int di = -R; di <= R; di++) {
            const int ii = cx + di;
            const double dx = px - ii * h, dx2 = dx * dx;
            const long i0 = ((ii + n) % n) * N2;
            const long ip = wrapi(i0 / N2 + 1, N) * N2, im = wrapi(i0 / N2 - 1, N) * N2;
            for (int dj = -R; dj <= R; dj++) {
                const int jj = cy + dj;
                const double dy = py - jj * h, dy2 = dy * dy;
                const long j = (jj + n) % n;
                const long j0 = j * N, jp = wrapi(j + 1, N) * N, jm = wrapi(j - 1, N) * N;
                for (int dk = -R; dk <= R; dk++) {
                    const int kk = cz + dk;
                    const double dz = pz - kk * h, dz2 = dz * dz;
                    const long k = (kk + n) % n, kp = wrapi(k + 1, N), km = wrapi(k - 1, N);
                    const long c = i0 + j0 + k;
                    const double exc = ex[c], eyc = ey[c], ezc = ez[c];
                    const double exb = ex[im + j0 + k], eyb = ey[i0 + jm + k], ezb = ez[i0 + j0 + km];
                    const double gx = exc - exb, gy = eyc - eyb, gz = ezc - ezb;
                    const double gn = sqrt(gx * gx + gy * gy + gz * gz) / h;
                    const double ig = (gn > 0.0) ? 1.0 / gn : 0.0;
                    const double phic = phi[c];
                    const double r2 = dx2 + dy2 + dz2;
                    if (r2 < rp2 && r2 > rm2) {
                        const double r1 = sqrt(r2) + 1E-12;
                        const double f = k2[c] * phic * phic * pb_step_dlog(r1 - r + w, w2, w3);
                        ib[0] += f * (dx / r1);
                        ib[1] += f * (dy / r1);
                        ib[2] += f * (dz / r1);
                    }
                    const double gxn = ig * gx, gyn = ig * gy, gzn = ig * gz;
                    pb_face(dx + hd2, dy, dz, rp2, rm2, r, w, w2, w3, exb, eps_int,
                            phi[im + j0 + k] - phic, phic, gxn, -1, nonpolar, db, nf);
                    pb_face(dx - hd2, dy, dz, rp2, rm2, r, w, w2, w3, exc, eps_int,
                            phi[ip + j0 + k] - phic, phic, gxn, +1, nonpolar, db, nf);
                    pb_face(dx, dy + hd2, dz, rp2, rm2, r, w, w2, w3, eyb, eps_int,
                            phi[i0 + jm + k] - phic, phic, gyn, -1, nonpolar, db, nf);
                    pb_face(dx, dy - hd2, dz, rp2, rm2, r, w, w2, w3, eyc, eps_int,
                            phi[i0 + jp + k] - phic, phic, gyn, +1, nonpolar, db, nf);
                    pb_face(dx, dy, dz + hd2, rp2, rm2, r, w, w2, w3, ezb, eps_int,
                            phi[i0 + j0 + km] - phic, phic, gzn, -1, nonpolar, db, nf);
                    pb_face(dx, dy, dz - hd2, rp2, rm2, r, w, w2, w3, ezc, eps_int,
                            phi[i0 + j0 + kp] - phic, phic, gzn, +1, nonpolar, db, nf);
                }
            }
        }
    }
    const double s = h / (8.0 * M_PI), snp = -gamma_np * h / (eps_s - eps_int);
    for (long i = 0; i < 3L * np; i++) {
        fdb[i] *= s;
        fib[i] *= s;
        fnp[i] *= snp;
    }
    return 0.0;
}

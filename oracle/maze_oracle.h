/*
 * maze_oracle.h -- CPU oracle for the MaZe electrostatics + MD hot path.
 *
 * TEST INFRASTRUCTURE ONLY.  This is a plain-C, single-threaded FP64
 * restatement of the reference algorithm (Crivella/MaZe_Poisson,
 * library/src/*.c).  It exists to check the HIP library, never to be
 * shipped or measured as the product: only tests/, __graft_entry__.smoke()
 * and bench.py's cpu_baseline leg may load liboracle.so.
 *
 * Every function mirrors the reference's floating-point operation order
 * (left-to-right sums, the same multiply/divide grouping, serial reduction
 * order), so that, compiled with -ffp-contract=off and no fast-math, it
 * reproduces the reference C compiled the same way bit for bit.  That is
 * how it is pinned (tests/test_oracle_golden.py); the remaining checks are
 * closed forms (tests/test_oracle_closed_forms.py).
 *
 * Grid arrays are n*n*n doubles, C order [i][j][k] with k fastest, and
 * periodic in all three directions (the reference keeps one ghost plane on
 * each side along i, library/src/mpi_base.c:223-233; here the wrap is done
 * with index arithmetic, which reads the same values).
 */
#ifndef MAZE_ORACLE_H
#define MAZE_ORACLE_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ---- grid operators ---------------------------------------------------- */
void or_laplace(const double *u, double *out, int n);
void or_laplace_pb(const double *u, double *out, int n, const double *ex,
                   const double *ey, const double *ez, const double *k2);

/* ---- Krylov solvers (return the 0-based iteration index, -1 if capped) -- */
int or_conj_grad(const double *b, double *x, double tol, int n);
int or_conj_grad_jacobi(const double *b, double *x, double tol, int n);
int or_conj_grad_mg(const double *b, double *x, double tol, int n);
int or_conj_grad_bj(const double *b, double *x, double tol, int n);
/* v = A_B^-1 r on every 10^3 block (n % 10 == 0). */
void or_blockjacobi_apply(const double *in, double *out, int n);
int or_conj_grad_pb(const double *b, double *x, double tol, int n,
                    const double *ex, const double *ey, const double *ez,
                    const double *k2);

/* ---- MaZe constraint step ---------------------------------------------- */
int or_verlet_poisson(double tol, double h, double *phi, double *phi_prev,
                      const double *q, double *y, int n, int precond);
int or_verlet_poisson_pb(double tol, double h, double *phi, double *phi_prev,
                         const double *q, double *y, int n, const double *ex,
                         const double *ey, const double *ez, const double *k2);
int or_verlet_poisson_mg(double tol, double h, double *phi, double *phi_prev,
                         const double *q, double *y, int n);

/* ---- reference multigrid pieces ---------------------------------------- */
void or_mg_smooth_rbgs(const double *in, double *out, int n, int iters);
void or_mg_restrict27(const double *in, double *out, int n);
void or_mg_prolong_trilinear(const double *in, double *out, int nc);
int or_mg_cg_coarse(const double *b, double *x, int n, int maxit, double rtol);
void or_mg_vcycle(const double *in, double *out, int n, int sm);

/* ---- reference Poisson-Boltzmann multigrid ------------------------------ */
void or_mg_smooth_pb_rbgs(const double *in, double *out, int n, int iters, const double *ex,
                          const double *ey, const double *ez, const double *k2);
void or_mg_restrict_eps(const double *in, double *out, int n, int axis);
void or_mg_restrict_k2(const double *in, double *out, int n);
int or_mg_cg_coarse_pb(const double *b, double *x, int n, int maxit, double rtol,
                       const double *ex, const double *ey, const double *ez, const double *k2);
void or_mg_vcycle_pb(const double *in, double *out, int n, int sm, const double *ex,
                     const double *ey, const double *ez, const double *k2);
int or_verlet_pb_multigrid(double tol, double h, double *phi, double *phi_prev, const double *q,
                           double *y, int n, const double *ex, const double *ey,
                           const double *ez, const double *k2);
int or_mg_pb_update(double tol, const double *b, double *phi, int n, const double *ex,
                    const double *ey, const double *ez, const double *k2);

/* ---- particles --------------------------------------------------------- */
double or_cic_weight(double x, double L, double h);
double or_spline_quadr_weight(double x, double L, double h);
double or_spline_cubic_weight(double x, double L, double h);
int or_cas_nodes(int cas);
void or_neighbors(int cas, const double *pos, int np, int n, double h, int64_t *nb);
double or_spread(int cas, int n, int np, double h, const double *pos, const int64_t *nb,
                 const double *charges, double *q);
double or_force_fd_cas(int cas, int n, int np, double h, const double *phi, const int64_t *nb,
                       const double *charges, const double *pos, double *forces);
void or_cic_neighbors(const double *pos, int np, int n, double h, int64_t *nb);
double or_spread_cic(int n, int np, double h, const double *pos,
                     const int64_t *nb, const double *charges, double *q);
double or_force_fd(int n, int np, double h, const double *phi,
                   const int64_t *nb, const double *charges, const double *pos,
                   double *forces);
void or_tf_pair_table(int n_typ, const double *pot_params, double r_cut,
                      double *table);
double or_tf_forces(int np, double L, const double *pos, const int *types,
                    int n_typ, const double *table, double r_cut,
                    double *forces);
void or_tf_neighbor_counts(int np, double L, const double *pos, double r_cut,
                           int *counts);

/* ---- Poisson-Boltzmann boundary from ion radii -------------------------- */
void or_pb_maps(int n, double h, double w, double eps_s, double eps_int, double kbar2, int np,
                const double *pos, const double *radii, double *ex, double *ey, double *ez,
                double *k2);
double or_pb_forces(int n, double h, double w, double eps_s, double eps_int, int nonpolar,
                    double gamma_np, int np, const double *pos, const double *radii,
                    const double *phi, const double *ex, const double *ey, const double *ez,
                    const double *k2, double *fdb, double *fib, double *fnp);

/* ---- counter-based normal deviates (Philox4x32-10 + Box-Muller) --------- */
void or_philox4x32(const uint32_t ctr[4], const uint32_t key[2], uint32_t out[4]);
void or_normal3(uint64_t seed, uint64_t call, uint32_t ion, double out[3]);

/* ---- reductions -------------------------------------------------------- */
double or_kinetic(int np, const double *vel, const double *mass);

#ifdef __cplusplus
}
#endif
#endif

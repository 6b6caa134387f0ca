/*
 * oracle_abi.c -- the reference's process-global C-ABI, restated on top of
 * the CPU oracle (test infrastructure; see maze_oracle.h).
 *
 * Same symbol names, argument meaning and error behaviour as the reference
 * facade library/src/solver_md.c:12-354 and the grid/particle/integrator
 * dispatch it reaches (grid_general.c:24-83, grid_lcg.c, grid_maze_lcg.c,
 * grid_multigrid.c, grid_maze_multigrid.c, particles.c, integrator_*.c), so
 * that one ctypes driver can run the compiled reference, this oracle and the
 * HIP library with identical calls.
 *
 * Deliberate differences, also made by the HIP library:
 *  - OVRVO noise comes from or_normal3 (Philox) instead of rand();
 *  - the Tosi-Fumi constants are held per type pair, not per particle pair
 *    (the values are identical: they only depend on the two types);
 *  - only TF pair potentials are provided (CIC and both spline assignments are).
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>

#include "maze_oracle.h"

#define KB_HA 3.1668e-6 /* library/include/constants.h:10 */
#define OR_PI 3.14159265358979323846

enum { G_LCG = 0, G_FFT = 1, G_MG = 2, G_MAZE_LCG = 3, G_MAZE_MG = 4 };
enum { P_NONE = 0, P_JACOBI = 1, P_MG = 2, P_SSOR = 3, P_BJ = 4 };
enum { I_OVRVO = 0, I_VERLET = 1 };

static const char *grid_names[5] = {"LCG", "FFT", "MULTIGRID", "MAZE-LCG", "MAZE-MULTIGRID"};
static const char *precond_names[5] = {"NONE", "JACOBI", "MG", "SSOR", "BLOCKJACOBI"};
static const char *pot_names[3] = {"TF", "LJ", "SC"};
static const char *cas_names[3] = {"CIC", "SPL_QUADR", "SPL_CUBIC"};
static const char *itg_names[2] = {"OVRVO", "VERLET"};

static struct {
    int ready;
    int n, type, precond;
    double L, h, tol, eps_s, eps_int;
    long G;
    double *q, *y, *phi_p, *phi_n;
    int pb, nonpolar;
    double w, kbar2;
    double *ex, *ey, *ez, *k2;
} g;

static struct {
    int ready;
    int n, np, n_typ;
    double L, h, r_cut;
    int *types;
    double *pos, *vel, *mass, *charges, *fe, *fn, *ft, *radii;
    double *fdb, *fib, *fnp;  /* PB boundary forces (with radii) */
    double gamma_np, beta_np;
    int64_t *nb;
    int cas, nn;  /* charge-assignment scheme and its nodes per ion */
    double *table;
} p;

static struct {
    int ready, type, enabled;
    double dt, T, c1, c2;
    uint64_t seed, calls;
} it = {0, 0, 0, 0.0, 0.0, 1.0, 1.0, 42u, 0u};

static double q_tot = 0.0; /* library/src/solver_md.c:16 */

static void die(const char *msg) {
    fprintf(stderr, "%s\n", msg);
    exit(1);
}

int init_mpi(void) { return 0; }
int get_rank(void) { return 0; }

int get_grid_type_num(void) { return 5; }
const char *get_grid_type_str(int i) { return grid_names[i]; }
int get_precond_type_num(void) { return 5; }
const char *get_precond_type_str(int i) { return precond_names[i]; }
int get_potential_type_num(void) { return 3; }
const char *get_potential_type_str(int i) { return pot_names[i]; }
int get_ca_scheme_type_num(void) { return 3; }
const char *get_ca_scheme_type_str(int i) { return cas_names[i]; }
int get_integrator_type_num(void) { return 2; }
const char *get_integrator_type_str(int i) { return itg_names[i]; }

void solver_initialize(void) {}

/* library/src/solver_md.c:40-44 -> grid_general.c:24-83 */
void solver_initialize_grid(int n, double L, double h, double tol, double eps, double eps_int,
                            int grid_type, int precond_type) {
    if (grid_type == G_FFT) die("oracle: FFT grid not provided");
    if (grid_type < 0 || grid_type > 4) die("oracle: invalid grid type");
    if ((grid_type == G_MG || grid_type == G_MAZE_MG) && (n % 2 || n < 32))
        die("oracle: multigrid needs an even n >= 32");
    if (grid_type == G_MAZE_LCG && precond_type == P_SSOR)
        die("oracle: MAZE-LCG provides the NONE, JACOBI, MG and BLOCKJACOBI preconditioners");
    if (grid_type == G_MAZE_LCG && precond_type == P_BJ && n % 10)
        die("oracle: the block-Jacobi preconditioner needs n to be a multiple of 10");
    if (grid_type == G_MAZE_LCG && precond_type == P_MG && (n % 2 || n < 32))
        die("oracle: the multigrid preconditioner needs an even n >= 32");
    memset(&g, 0, sizeof(g));
    g.n = n; g.L = L; g.h = h; g.tol = tol; g.eps_s = eps; g.eps_int = eps_int;
    g.type = grid_type; g.precond = precond_type;
    g.G = (long)n * n * n;
    g.q = (double *)calloc(g.G, sizeof(double));
    g.y = (double *)calloc(g.G, sizeof(double));
    g.phi_p = (double *)calloc(g.G, sizeof(double));
    g.phi_n = (double *)calloc(g.G, sizeof(double));
    g.ready = 1;
}

/* library/src/grid_general.c:85-100 */
void solver_initialize_grid_pois_boltz(double w, double kbar2, int nonpolar_enabled) {
    g.pb = 1; g.w = w; g.kbar2 = kbar2; g.nonpolar = nonpolar_enabled;
    g.ex = (double *)malloc(g.G * sizeof(double));
    g.ey = (double *)malloc(g.G * sizeof(double));
    g.ez = (double *)malloc(g.G * sizeof(double));
    g.k2 = (double *)malloc(g.G * sizeof(double));
    for (long i = 0; i < g.G; i++) { g.ex[i] = g.ey[i] = g.ez[i] = 1.0; g.k2[i] = 0.0; }
}

/* Extension: install face-centred dielectric maps and the screening map
 * directly (the reference derives them from ion radii). */
void solver_set_eps_k2(const double *ex, const double *ey, const double *ez, const double *k2) {
    memcpy(g.ex, ex, g.G * sizeof(double));
    memcpy(g.ey, ey, g.G * sizeof(double));
    memcpy(g.ez, ez, g.G * sizeof(double));
    memcpy(g.k2, k2, g.G * sizeof(double));
}

/* library/src/solver_md.c:135-138; a no-op until radii are given (the
 * reference needs them), so that maps set directly are kept. */
void get_eps_k2(double *ex, double *ey, double *ez, double *k2) {
    if (!g.pb) die("oracle: get_eps_k2 needs solver_initialize_grid_pois_boltz first");
    memcpy(ex, g.ex, g.G * sizeof(double));
    memcpy(ey, g.ey, g.G * sizeof(double));
    memcpy(ez, g.ez, g.G * sizeof(double));
    memcpy(k2, g.k2, g.G * sizeof(double));
}

void solver_update_eps_k2(void) {
    if (!g.pb || !p.radii) return;
    or_pb_maps(g.n, g.h, g.w, g.eps_s, g.eps_int, g.kbar2, p.np, p.pos, p.radii, g.ex, g.ey,
               g.ez, g.k2);
}

/* library/src/solver_md.c:51-65, particles.c:68-109, :181-228 */
void solver_initialize_particles(int n, int n_typ, double L, double h, int n_p, int pot_type,
                                 int cas_type, int *types, double *pos, double *vel, double *mass,
                                 double *charges, double *pot_params) {
    if (pot_type != 0) die("oracle: only the TF potential is provided");
    if (cas_type < 0 || cas_type > 2) die("Invalid charge assignment scheme type");
    memset(&p, 0, sizeof(p));
    p.cas = cas_type;
    p.nn = or_cas_nodes(cas_type);
    p.n = n; p.np = n_p; p.n_typ = n_typ; p.L = L; p.h = h;
    p.types = (int *)malloc(n_p * sizeof(int));
    p.pos = (double *)malloc(n_p * 3 * sizeof(double));
    p.vel = (double *)malloc(n_p * 3 * sizeof(double));
    p.mass = (double *)malloc(n_p * sizeof(double));
    p.charges = (double *)malloc(n_p * sizeof(double));
    p.fe = (double *)calloc(n_p * 3, sizeof(double));
    p.fn = (double *)calloc(n_p * 3, sizeof(double));
    p.ft = (double *)calloc(n_p * 3, sizeof(double));
    p.nb = (int64_t *)malloc((size_t)n_p * p.nn * 3 * sizeof(int64_t));
    memcpy(p.types, types, n_p * sizeof(int));
    memcpy(p.pos, pos, n_p * 3 * sizeof(double));
    memcpy(p.vel, vel, n_p * 3 * sizeof(double));
    memcpy(p.mass, mass, n_p * sizeof(double));
    memcpy(p.charges, charges, n_p * sizeof(double));
    p.r_cut = L / 2.0;
    p.table = (double *)malloc(n_typ * n_typ * 7 * sizeof(double));
    or_tf_pair_table(n_typ, pot_params, p.r_cut, p.table);
    p.ready = 1;
}

/* library/src/particles.c:111-125 */
void solver_initialize_particles_pois_boltz(double gamma_np, double beta_np, double *radii) {
    p.gamma_np = gamma_np; p.beta_np = beta_np;
    p.radii = (double *)malloc(p.np * sizeof(double));
    memcpy(p.radii, radii, p.np * sizeof(double));
    p.fdb = (double *)calloc(p.np * 3, sizeof(double));
    p.fib = (double *)calloc(p.np * 3, sizeof(double));
    p.fnp = (double *)calloc(p.np * 3, sizeof(double));
}

/* library/src/solver_md.c:71-89, integrator_general.c:19-41,
 * integrator_ovrvo.c:140-148, integrator_verlet.c:74-77 */
void solver_initialize_integrator(int n_p, double dt, double T, double gamma, int itg_type,
                                  int itg_enabled) {
    (void)n_p;
    it.type = itg_type; it.dt = dt; it.T = T; it.enabled = 0; it.c1 = 1.0; it.c2 = 1.0;
    it.calls = 0;
    if (itg_enabled == 1) {
        it.enabled = 1;
        if (itg_type == I_OVRVO) {
            it.c1 = exp(-gamma * dt);
            it.c2 = sqrt(2 / (gamma * dt) * tanh(0.5 * gamma * dt));
        }
    }
    it.ready = 1;
}

void solver_set_rng_seed(uint64_t seed) { it.seed = seed; it.calls = 0; }

/* library/src/solver_md.c:91-107 */
int solver_update_charges(void) {
    or_neighbors(p.cas, p.pos, p.np, p.n, p.h, p.nb);
    double loc = or_spread(p.cas, g.n, p.np, g.h, p.pos, p.nb, p.charges, g.q);
    if (fabs(q_tot - loc) > 1e-6) {
        printf("Charge conservation error: q_tot = %.6f, q_tot_loc = %.6f\n", q_tot, loc);
        exit(1);
    }
    q_tot = loc;
    return 0;
}

/* library/src/grid_lcg.c:100-124, grid_maze_lcg.c:100-124,
 * grid_multigrid.c:86-108, grid_maze_multigrid.c:86-112 */
void solver_init_field(void) {
    double c = -4 * OR_PI / g.h;
    if (!g.pb) c /= g.eps_s;
    memset(g.y, 0, g.G * sizeof(double));
    memcpy(g.phi_p, g.phi_n, g.G * sizeof(double));
    double *b = (double *)malloc(g.G * sizeof(double));
    for (long i = 0; i < g.G; i++) b[i] = g.q[i];
    for (long i = 0; i < g.G; i++) b[i] *= c;
    memset(g.phi_n, 0, g.G * sizeof(double)); /* x starts from y == 0 */
    if (g.pb)
        or_conj_grad_pb(b, g.phi_n, g.tol, g.n, g.ex, g.ey, g.ez, g.k2);
    else
        or_conj_grad(b, g.phi_n, g.tol, g.n);
    free(b);
}

void solver_set_field(double *phi) { memcpy(g.phi_n, phi, g.G * sizeof(double)); }
void solver_set_field_prev(double *phi) { memcpy(g.phi_p, phi, g.G * sizeof(double)); }

/* library/src/grid_lcg.c:126-149 */
static int lcg_update(void) {
    double c = -4 * OR_PI / g.h;
    if (!g.pb) c /= g.eps_s;
    memcpy(g.phi_p, g.phi_n, g.G * sizeof(double));
    for (long i = 0; i < g.G; i++) g.y[i] = g.q[i];
    for (long i = 0; i < g.G; i++) g.y[i] *= c;
    if (g.pb) return or_conj_grad_pb(g.y, g.phi_n, g.tol, g.n, g.ex, g.ey, g.ez, g.k2);
    return or_conj_grad(g.y, g.phi_n, g.tol, g.n); /* precond ignored here, as in grid_lcg.c */
}

/* library/src/grid_multigrid.c:110-212 (Poisson branch; PB branch :149-175) */
static int mg_update(void) {
    const long G = g.G;
    double c = -4 * OR_PI / g.h;
    if (!g.pb) c /= g.eps_s;
    for (long i = 0; i < G; i++) {
        double a = g.phi_n[i];
        g.phi_n[i] = 2 * a - g.phi_p[i];
        g.phi_p[i] = a;
    }
    double *b = (double *)malloc(G * sizeof(double));
    double *t = (double *)malloc(G * sizeof(double));
    for (long i = 0; i < G; i++) b[i] = g.q[i];
    for (long i = 0; i < G; i++) b[i] *= c;
    if (g.pb) {
        int r = or_mg_pb_update(g.tol, b, g.phi_n, g.n, g.ex, g.ey, g.ez, g.k2);
        free(b); free(t);
        return r;
    }
    int res = -1, k = 0;
    while (k < 1000) {
        or_mg_vcycle(b, g.phi_n, g.n, 3);
        or_laplace(g.phi_n, t, g.n);
        for (long i = 0; i < G; i++) t[i] += -1.0 * b[i];
        double m = 0.0;
        for (long i = 0; i < G; i++) if (fabs(t[i]) > m) m = fabs(t[i]);
        k++;
        if (m <= g.tol) { res = k; break; }
    }
    free(b); free(t);
    return res;
}

/* library/src/solver_md.c:131-133 and the per-type update_field */
int solver_update_field(void) {
    switch (g.type) {
    case G_LCG:
        return lcg_update();
    case G_MAZE_LCG:
        if (g.pb)
            return or_verlet_poisson_pb(g.tol, g.h, g.phi_n, g.phi_p, g.q, g.y, g.n, g.ex, g.ey,
                                        g.ez, g.k2);
        return or_verlet_poisson(g.tol, g.h * g.eps_s, g.phi_n, g.phi_p, g.q, g.y, g.n,
                                 g.precond);
    case G_MG:
        return mg_update();
    case G_MAZE_MG: {
        int r = g.pb ? or_verlet_pb_multigrid(g.tol, g.h, g.phi_n, g.phi_p, g.q, g.y, g.n, g.ex,
                                              g.ey, g.ez, g.k2)
                     : or_verlet_poisson_mg(g.tol, g.h * g.eps_s, g.phi_n, g.phi_p, g.q, g.y, g.n);
        if (g.precond != P_NONE) die("Maze Multigrid with preconditioner not implemented yet.");
        return r;
    }
    }
    return -1;
}

void solver_compute_forces_elec(void) {
    or_force_fd_cas(p.cas, p.n, p.np, p.h, g.phi_n, p.nb, p.charges, p.pos, p.fe);
}

double solver_compute_forces_noel(void) {
    return or_tf_forces(p.np, p.L, p.pos, p.types, p.n_typ, p.table, p.r_cut, p.fn);
}

/* library/src/solver_md.c:148-150, particles.c:420-775 */
double solver_compute_forces_pb(void) {
    if (!g.pb) die("Poisson-Boltzmann forces are not enabled in the grid.");
    if (!p.radii) die("Poisson-Boltzmann forces need the ion radii "
                      "(solver_initialize_particles_pois_boltz).");
    return or_pb_forces(g.n, g.h, g.w, g.eps_s, g.eps_int, g.nonpolar, p.gamma_np, p.np, p.pos,
                        p.radii, g.phi_n, g.ex, g.ey, g.ez, g.k2, p.fdb, p.fib, p.fnp);
}

/* library/src/particles.c:777-795 */
void solver_compute_forces_tot(void) {
    const int s = p.np * 3;
    for (int i = 0; i < s; i++) p.ft[i] = 0.0;
    for (int i = 0; i < s; i++) p.ft[i] += 1.0 * p.fe[i];
    for (int i = 0; i < s; i++) p.ft[i] += 1.0 * p.fn[i];
    if (p.fdb) {
        for (int i = 0; i < s; i++) p.ft[i] += 1.0 * p.fdb[i];
        for (int i = 0; i < s; i++) p.ft[i] += 1.0 * p.fib[i];
        for (int i = 0; i < s; i++) p.ft[i] += 1.0 * p.fnp[i];
    }
}

double get_kinetic_energy(void) { return or_kinetic(p.np, p.vel, p.mass); }
double get_temperature(void) { return 2 * get_kinetic_energy() / (3 * p.np * KB_HA); }

/* Single-image wrap after a drift, integrator_verlet.c:45-51 */
static inline double wrap_pos(double a, double L) {
    if (a < 0) return a + L;
    if (a >= L) return a - L;
    return a;
}

static void kick(double cdt) { /* v += cdt F / m, cdt = 0.5 dt [c2] */
    for (int i = 0; i < p.np; i++)
        for (int d = 0; d < 3; d++) p.vel[3 * i + d] += cdt * p.ft[3 * i + d] / p.mass[i];
}

/* O block: v <- sqrt(c1) v + sqrt((1-c1) kB T / m) xi.  integrator_ovrvo.c:45-78 */
static void o_block(void) {
    double c1s = sqrt(it.c1), var2 = (1 - it.c1) * KB_HA * it.T;
    for (int i = 0; i < p.np; i++) {
        double xi[3];
        or_normal3(it.seed, it.calls, (uint32_t)i, xi);
        double var1 = sqrt(var2 / p.mass[i]);
        for (int d = 0; d < 3; d++) {
            p.vel[3 * i + d] *= c1s;
            p.vel[3 * i + d] += var1 * xi[d];
        }
    }
    it.calls++;
}

/* integrator_verlet.c:12-53, integrator_ovrvo.c:80-132 */
void integrator_part_1(void) {
    if (it.type == I_VERLET) {
        if (it.enabled) {
            double scale = sqrt(it.T / get_temperature());
            for (int i = 0; i < 3 * p.np; i++) p.vel[i] *= scale;
        }
        for (int i = 0; i < p.np; i++)
            for (int d = 0; d < 3; d++) {
                const int k = 3 * i + d;
                p.vel[k] += 0.5 * it.dt * p.ft[k] / p.mass[i];
                p.pos[k] = wrap_pos(p.pos[k] + it.dt * p.vel[k], p.L);
            }
    } else {
        if (it.enabled) o_block();
        kick(0.5 * it.dt * it.c2);
        for (int k = 0; k < 3 * p.np; k++) p.pos[k] = wrap_pos(p.pos[k] + it.c2 * it.dt * p.vel[k], p.L);
    }
}

void integrator_part_2(void) {
    if (it.type == I_VERLET) {
        kick(0.5 * it.dt);
    } else {
        kick(0.5 * it.dt * it.c2);
        if (it.enabled) o_block();
    }
}

/* library/src/particles.c:841-857 */
void solver_rescale_velocities(void) {
    double *s = (double *)calloc(p.n_typ * 3, sizeof(double));
    for (int i = 0; i < p.np; i++)
        for (int d = 0; d < 3; d++) s[p.types[i] * 3 + d] += p.vel[i * 3 + d];
    for (int i = 0; i < p.np; i++)
        for (int d = 0; d < 3; d++) p.vel[i * 3 + d] -= 2 * s[p.types[i] * 3 + d] / p.np;
    free(s);
}

/* library/src/solver_md.c:210-222 */
int solver_check_thermostat(void) {
    if (it.enabled && fabs(get_temperature() - it.T) < 100) {
        it.enabled = 0;
        if (it.type == I_OVRVO) { it.c1 = 1.0; it.c2 = 1.0; }
        return 1;
    }
    return 0;
}

/* One MD step in the order of maze_poisson/solver.py:483-492 (elec and
 * not_elec enabled): reference-commented solver_md_loop_iter,
 * library/src/solver_md.c:202-208. */
void solver_md_loop_iter(void) {
    integrator_part_1();
    solver_update_charges();
    solver_update_field();
    solver_compute_forces_elec();
    solver_compute_forces_noel();
    solver_compute_forces_tot();
    integrator_part_2();
}

void solver_run_n_steps(int n_steps) {
    for (int s = 0; s < n_steps; s++) solver_md_loop_iter();
}

void solver_finalize(void) {
    free(g.q); free(g.y); free(g.phi_p); free(g.phi_n);
    free(g.ex); free(g.ey); free(g.ez); free(g.k2);
    memset(&g, 0, sizeof(g));
    free(p.types); free(p.pos); free(p.vel); free(p.mass); free(p.charges);
    free(p.fe); free(p.fn); free(p.ft); free(p.nb); free(p.table); free(p.radii);
    free(p.fdb); free(p.fib); free(p.fnp);
    memset(&p, 0, sizeof(p));
    q_tot = 0.0;
}

void get_pos(double *o) { memcpy(o, p.pos, p.np * 3 * sizeof(double)); }
void get_vel(double *o) { memcpy(o, p.vel, p.np * 3 * sizeof(double)); }
void get_fcs_elec(double *o) { memcpy(o, p.fe, p.np * 3 * sizeof(double)); }
void get_fcs_noel(double *o) { memcpy(o, p.fn, p.np * 3 * sizeof(double)); }
void get_fcs_tot(double *o) { memcpy(o, p.ft, p.np * 3 * sizeof(double)); }
static void get_or_zero(double *o, const double *src) {
    if (src) memcpy(o, src, p.np * 3 * sizeof(double));
    else memset(o, 0, p.np * 3 * sizeof(double));
}
void get_fcs_db(double *o) { get_or_zero(o, p.fdb); }
void get_fcs_ib(double *o) { get_or_zero(o, p.fib); }
void get_fcs_np(double *o) { get_or_zero(o, p.fnp); }
void get_types(int *o) { memcpy(o, p.types, p.np * sizeof(int)); }
void get_charges(double *o) { memcpy(o, p.charges, p.np * sizeof(double)); }
void get_masses(double *o) { memcpy(o, p.mass, p.np * sizeof(double)); }
void get_radii(double *o) {
    if (p.radii) memcpy(o, p.radii, p.np * sizeof(double));
    else memset(o, 0, p.np * sizeof(double));
}
void get_field(double *o) { memcpy(o, g.phi_n, g.G * sizeof(double)); }
void get_field_prev(double *o) { memcpy(o, g.phi_p, g.G * sizeof(double)); }
void get_q(double *o) { memcpy(o, g.q, g.G * sizeof(double)); }
void get_y(double *o) { memcpy(o, g.y, g.G * sizeof(double)); }
void get_neighbors(int64_t *o) { memcpy(o, p.nb, (size_t)p.np * p.nn * 3 * sizeof(int64_t)); }
void get_tf_neighbor_counts(int *o) { or_tf_neighbor_counts(p.np, p.L, p.pos, p.r_cut, o); }

/* library/src/grid_general.c:304-316 (half the sum of phi, no charge) */
double get_energy_elec(void) {
    double e = 0.0;
    for (long i = 0; i < g.G; i++) e += 0.5 * g.phi_n[i];
    return e;
}

/* Extension: 1/2 sum_g q_g phi_g, the electrostatic energy of the grid charges. */
double get_energy_elec_qphi(void) {
    double e = 0.0;
    for (long i = 0; i < g.G; i++) e += g.q[i] * g.phi_n[i];
    return 0.5 * e;
}

/* library/src/particles.c:822-839 */
void get_momentum(double *o) {
    double px = 0.0, py = 0.0, pz = 0.0;
    for (int i = 0; i < p.np; i++) {
        px += p.mass[i] * p.vel[3 * i];
        py += p.mass[i] * p.vel[3 * i + 1];
        pz += p.mass[i] * p.vel[3 * i + 2];
    }
    o[0] = px; o[1] = py; o[2] = pz;
}

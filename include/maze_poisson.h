/*
 * maze_poisson.h -- C-ABI of libmaze_poisson.so, the MI355X (gfx950) HIP
 * implementation of the MaZe electrostatics + MD hot path.
 *
 * Part 1 is the reference's ABI: same names, argument meaning and error
 * behaviour as the reference library that maze_poisson/c_api binds through
 * ctypes (maze_poisson/c_api/solver.py, get_functions.py, mpi_base.py).
 * Each declaration cites the reference definition it replaces.  State is one
 * process-global replica resident in GPU memory; array arguments are
 * caller-owned host buffers (C-contiguous, float64 unless stated), copied in
 * at initialisation and copied out by the getters.
 *
 * Part 2 lists this library's extensions (no reference counterpart, or a
 * counterpart that exists only as commented-out code).
 *
 * Part 3 is the replica engine: many independent MD replicas on one GPU,
 * each with its own state, advanced together (no reference counterpart; the
 * reference runs one replica per process).
 *
 * Units are atomic units throughout (Bohr, Hartree, electron masses, a.u.
 * time).  Grid arrays are n*n*n, C order [i][j][k], k fastest.
 */
#ifndef MAZE_POISSON_H
#define MAZE_POISSON_H

#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

/* ======================= Part 1: the reference ABI ========================= */

/* library/src/mpi_base.c:169-177 (serial build): returns 0. */
int init_mpi(void);
/* library/src/mpi_base.c (get_rank): returns 0 (one GPU, one rank). */
int get_rank(void);

/* Enum discovery, library/src/grid_general.c:8-22: grid LCG 0, FFT 1,
 * MULTIGRID 2, MAZE-LCG 3, MAZE-MULTIGRID 4; precond NONE 0, JACOBI 1, MG 2,
 * SSOR 3, BLOCKJACOBI 4. */
int get_grid_type_num(void);
char *get_grid_type_str(int n);
int get_precond_type_num(void);
char *get_precond_type_str(int n);
/* library/src/particles.c:17-36: potential TF 0, LJ 1, SC 2; charge
 * assignment CIC 0, SPL_QUADR 1, SPL_CUBIC 2. */
int get_potential_type_num(void);
char *get_potential_type_str(int n);
int get_ca_scheme_type_num(void);
char *get_ca_scheme_type_str(int n);
/* library/src/integrator_general.c:7-15: OVRVO 0, VERLET 1. */
int get_integrator_type_num(void);
char *get_integrator_type_str(int n);

/* library/src/solver_md.c:19-38: prints a banner, selects the GPU. */
void solver_initialize(void);

/* library/src/solver_md.c:40-44 -> grid_general.c:24-83.  MAZE-LCG takes the
 * preconditioners NONE, JACOBI and MG (library/src/grid_maze_lcg.c:130-145;
 * MG: one reference two-grid V-cycle per application, FP64 path).  Exits(1)
 * with a message for FFT (not provided), multigrid or the MG preconditioner
 * on an odd n or n < 32, and the SSOR and BLOCKJACOBI preconditioners (not
 * provided). */
void solver_initialize_grid(int n, double L, double h, double tol, double eps_s, double eps_int,
                            int grid_type, int precond_type);

/* library/src/solver_md.c:46-49 -> grid_general.c:85-100.  Enables the
 * linear Poisson-Boltzmann operator with transition width w (Bohr) and grid
 * screening kbar2; nonpolar_enabled switches on the non-polar boundary
 * force.  Dielectric maps start at 1 and the screening map at 0 until
 * solver_update_eps_k2 (with radii) or solver_set_eps_k2 sets them. */
void solver_initialize_grid_pois_boltz(double w, double kbar2, int nonpolar_enabled);

/* library/src/solver_md.c:51-65 -> particles.c:68-109, :181-228.
 * types int32[n_p]; pos, vel double[n_p][3]; mass, charges double[n_p];
 * pot_params double[n_typ][n_typ][5] = (A, B, C, D, sigma).  TF with CIC,
 * SPL_QUADR or SPL_CUBIC assignment (other potentials exit(1));
 * 1 <= n_typ <= 16. */
void solver_initialize_particles(int n, int n_typ, double L, double h, int n_p, int pot_type,
                                 int cas_type, int *types, double *pos, double *vel, double *mass,
                                 double *charges, double *pot_params);

/* library/src/solver_md.c:67-69 -> particles.c:111-125.  Per-ion PB radius
 * radii double[n_p] (solvation + probe radius, Bohr), the non-polar surface
 * tension gamma_np and offset beta_np (a.u.).  Needs
 * solver_initialize_grid_pois_boltz first; exits(1) when 2 (floor((r + w) /
 * h) + 2) + 1 exceeds n (the reference would then visit points twice). */
void solver_initialize_particles_pois_boltz(double gamma_np, double beta_np, double *radii);

/* library/src/solver_md.c:71-89, integrator_ovrvo.c:140-148. */
void solver_initialize_integrator(int n_p, double dt, double T, double gamma, int itg_type,
                                  int itg_enabled);

/* library/src/solver_md.c:91-107: CIC spreading; returns 0, or prints the
 * charge-conservation error and exit(1)s (see solver_set_charge_check). */
int solver_update_charges(void);

/* library/src/solver_md.c:109-111 -> grid_maze_lcg.c:100-124: cold solve. */
void solver_init_field(void);

/* library/src/solver_md.c:113-129. phi double[n][n][n]. */
void solver_set_field(double *phi);
void solver_set_field_prev(double *phi);

/* library/src/solver_md.c:131-133 -> grid_*.c update_field: the CG
 * iteration index at convergence (MAZE-LCG, LCG), the V-cycle count
 * (MULTIGRID, MAZE-MULTIGRID), or -1 if not converged.  In the mixed
 * precision mode (solver_set_precision(1)) it is the total number of FP32
 * inner iterations. */
int solver_update_field(void);

/* library/src/solver_md.c:135-138 -> grid_general.c:138-301: rebuilds the
 * dielectric and screening maps from the ion positions and radii (exits(1)
 * if more than 2048 ions can reach one 8^3 grid tile).  Without radii it
 * keeps the maps of solver_set_eps_k2. */
void solver_update_eps_k2(void);

/* library/src/solver_md.c:140-154 -> forces.c:26-100, :112-186,
 * particles.c:420-795.  solver_compute_forces_noel returns the Tosi-Fumi
 * potential energy; solver_compute_forces_pb computes the dielectric-
 * boundary, ionic-boundary and non-polar forces (needs the radii) and
 * returns the non-polar energy, 0 as in the reference; the total adds them
 * once the radii are set. */
void solver_compute_forces_elec(void);
double solver_compute_forces_noel(void);
double solver_compute_forces_pb(void);
void solver_compute_forces_tot(void);

/* library/src/solver_md.c:162-172 -> integrator_verlet.c:12-71,
 * integrator_ovrvo.c:45-138, particles.c:841-857. */
void integrator_part_1(void);
void integrator_part_2(void);
void solver_rescale_velocities(void);

/* library/src/solver_md.c:210-222. */
int solver_check_thermostat(void);

/* library/src/solver_md.c:230-245: frees the replica. */
void solver_finalize(void);

/* Getters, library/src/solver_md.c:247-354: copy device state into the
 * caller's arrays.  (n_p,3): pos, vel, fcs_*; (n_p): types (int32), charges,
 * masses, radii; (n,n,n): field, field_prev, q; (3): momentum.  fcs_db,
 * fcs_ib, fcs_np and radii are zero (radius-based PB forces not provided). */
void get_pos(double *recv);
void get_vel(double *recv);
void get_fcs_elec(double *recv);
void get_fcs_noel(double *recv);
void get_fcs_db(double *recv);
void get_fcs_ib(double *recv);
void get_fcs_np(double *recv);
void get_fcs_tot(double *recv);
void get_types(int *recv);
void get_charges(double *recv);
void get_masses(double *recv);
void get_radii(double *recv);
void get_field(double *recv);
void get_field_prev(double *recv);
void get_q(double *recv);
double get_kinetic_energy(void);
/* library/src/grid_general.c:304-316: half the sum of phi (no charge). */
double get_energy_elec(void);
void get_momentum(double *recv);
double get_temperature(void);

/* ========================= Part 2: extensions ============================= */

/* One MD step (integrator part 1, charges, field, forces, part 2) without
 * host round trips between the phases; the reference has it only as
 * commented-out code, library/src/solver_md.c:202-208 and :224-228. */
void solver_md_loop_iter(void);
void solver_run_n_steps(int n_steps);

/* Blocks until all work of the replica has finished. */
void solver_sync(void);

/* Seed of the counter-based (Philox4x32-10) Langevin noise that replaces
 * the reference's rand() stream (integrator_ovrvo.c:12-36); resets the
 * stream position. */
void solver_set_rng_seed(uint64_t seed);

/* Installs face-centred dielectric maps and the screening map of the linear
 * PB operator (each double[n][n][n]); needs
 * solver_initialize_grid_pois_boltz first. */
void solver_set_eps_k2(double *eps_x, double *eps_y, double *eps_z, double *k2);

/* Copies the current dielectric and screening maps out (each
 * double[n][n][n]); no counterpart in the reference, where the maps are
 * internal. */
void get_eps_k2(double *eps_x, double *eps_y, double *eps_z, double *k2);

/* Field-solve arithmetic: 0 = FP64 CG as in the reference (default);
 * 1 = FP32-storage CG inside FP64 iterative refinement, stopping on the
 * reference's criterion max|A y - sigma| <= tol evaluated in FP64;
 * 2 = as 1 with the multigrid-preconditioned FP32 CG below;
 * 3 = FP64 CG preconditioned by mode 2's V-cycle carried out in FP64, to the
 * same stop test (MAZE-LCG on the Poisson operator where mode 2 runs the
 * one-work-group coarsest solve; otherwise the MaZe solve runs mode 0).
 * Returns 0, or -1 for an unknown mode. */
int solver_set_precision(int mode);
/* Bandwidth in GB/s (bytes read + written) of the library's copy kernel over
 * n_doubles doubles, averaged over reps launches: the on-box copy rate that
 * memory-bound kernels are compared with (no reference counterpart). */
double maze_copy_bandwidth(long n_doubles, int reps);
/* 1: the cold solves of solver_init_field use the mixed path of the current
 * precision mode (refinement to the same FP64 stop test) instead of FP64 CG;
 * 0 (default): FP64 CG as in the reference.  MAZE-LCG only. */
void solver_set_mixed_init(int on);
int solver_get_precision(void);
/* The refinement passes of the last mixed-precision solve: returns their
 * number and writes each pass's inner iterations (negative when the pass
 * stopped on its iteration budget) to iters[0 .. min(count, max)). */
int solver_get_last_passes(int *iters, int max);
/* Relative residual reduction of each FP32 inner solve; <= 0 restores the
 * default of the mode (1e-4 for mode 1, 1e-6 for mode 2). */
void solver_set_inner_rtol(double rtol);
/* Mixed-precision MaZe solve: start from the linear extrapolation
 * 2 y_n-1 - y_n-2 of the correction field (1) instead of y_n-1 (0, the
 * reference's start).  The stop test is unchanged. */
void solver_set_y_extrapolation(int on);
/* 1/2 sum_g q_g phi_g (Hartree): the electrostatic energy of the grid charges.
 * get_energy_elec keeps the reference's 1/2 sum_g phi_g. */
double get_energy_elec_qphi(void);
/* Mode 2 = FP32 multigrid-preconditioned CG inside FP64 iterative
 * refinement (V(2,2) with two Jacobi-type sweeps per side, exact
 * coarsest solve); falls back to mode 1 when the grid has no multigrid
 * hierarchy (halving reaches an odd size above 128).  With the PB operator
 * the cycle is the variable-coefficient one (coefficients coarsened from the
 * maps after every map change, Chebyshev iterations on a coarsest level of at
 * most 16^3; otherwise mode 1).  One damping for both sweeps (plain damped
 * Jacobi): */
void solver_set_mg_omega(double omega);
/* Different weights for the two smoothing sweeps of each level (pre-smoothing
 * omega1 then omega2, post-smoothing omega2 then omega1).  Default: the
 * degree-2 Chebyshev weights 0.5695 and 1.7319 for the Jacobi-scaled spectrum
 * [1/3, 2]. */
void solver_set_mg_omegas(double omega1, double omega2);
int solver_mg_levels(void);

/* 1 (default): a charge-conservation failure exits like the reference;
 * 0: solver_update_charges returns 1 instead. */
void solver_set_charge_check(int strict);

/* The MaZe constraint field y, double[n][n][n]. */
void get_y(double *recv);
/* Assignment-node table as the reference stores it: int64[n_p][8][3] for CIC
 * (particles.c:297-346), int64[n_p][64][3] for the splines (:348-385). */
void get_neighbors(int64_t *recv);
/* Number of partners within the Tosi-Fumi cutoff per ion, int32[n_p]. */
void get_tf_neighbor_counts(int *recv);

/* Per-kernel timing with HIP events on the replica's stream (only solver
 * iterations that did work are counted).  Names: "cg_dir", "cg_update",
 * "cg_dir_f32", "cg_update_f32". */
void maze_prof_enable(int on);
void maze_prof_reset(void);
int maze_prof_get(const char *name, double *total_ms, long *count);
int maze_grid_blocks(void);
/* One launch of the plain copy kernel over n doubles (reads and writes 8n
 * bytes): the calibration reference for hardware byte counters. */
void maze_calib_copy(long n_doubles);

/* Kernel-level entry points on host buffers, run on a private replica (for
 * parity tests against the CPU oracle). */
void maze_test_laplace(const double *u, double *out, int n);
void maze_test_laplace_pb(const double *u, double *out, int n, const double *eps_x,
                          const double *eps_y, const double *eps_z, const double *k2);
int maze_test_conj_grad(const double *b, double *x, double tol, int n);
int maze_test_conj_grad_pb(const double *b, double *x, double tol, int n, const double *eps_x,
                           const double *eps_y, const double *eps_z, const double *k2);
void maze_test_rbgs(const double *in, double *out, int n, int iters);
int maze_test_restrict27(const double *in, double *out, int n);
int maze_test_prolong_add(const double *in, double *out, int nc);
int maze_test_mg_vcycle(const double *in, double *out, int n);
/* The block-Jacobi preconditioner v = A_B^-1 r on every 10^3 block (n a
 * multiple of 10; library/src/precond_blockjacobi.c:170-206). */
int maze_test_blockjacobi(const double *in, double *out, int n);
/* Reference Poisson-Boltzmann multigrid pieces (library/src/multigrid_pb.c):
 * RBGS-PB sweeps in place, the coarse coefficient restrictions
 * ((n/2)^3 outputs), one two-grid PB V-cycle (n even, >= 32). */
void maze_test_rbgs_pb(const double *in, double *out, int n, int iters, const double *eps_x,
                       const double *eps_y, const double *eps_z, const double *k2);
int maze_test_restrict_pb_coef(const double *eps_x, const double *eps_y, const double *eps_z,
                               const double *k2, double *c_eps_x, double *c_eps_y,
                               double *c_eps_z, double *c_k2, int n);
int maze_test_mg_vcycle_pb(const double *in, double *out, int n, const double *eps_x,
                           const double *eps_y, const double *eps_z, const double *k2);
int maze_test_mg_pcg(const double *b, double *x, double rtol, int n);
/* The fused FP32 multigrid level kernels (n even, c = damping / 6, tile = the
 * work-group tile shape 0..4, 5..6 for the plane-marching kernels, 64 + 0..4
 * for shape 0..4 with the scalar-load kernels, or -1 for the level's default):
 * down: d2 = two damped-Jacobi sweeps from zero on A d = b, bc = 4 x
 *       full-weighting restriction of b - A d2 ((n/2)^3 values);
 * up:   d = two damped-Jacobi sweeps on A d = b from d2 + trilinear(e). */
int maze_test_mg_down(const float *b, float *d2, float *bc, int n, double c, int tile);
int maze_test_mg_down2(const float *b, float *d2, float *bc, int n, double c1, double c2,
                       int tile);
int maze_test_mg_up2(const float *b, const float *d2, const float *e, float *d, int n, double c1,
                     double c2, int tile);
int maze_test_mg_up(const float *b, const float *d2, const float *e, float *d, int n, double c,
                    int tile);
/* The FP32 multigrid-preconditioned CG for the linear PB operator with the
 * given face maps and k2 (n^3 each), from x = 0 to the relative residual
 * rtol; returns the iterations, or -1 if the grid has no PB hierarchy. */
int maze_test_pb_mg_pcg(const double *b, double *x, double rtol, int n, const double *eps_x,
                        const double *eps_y, const double *eps_z, const double *k2);
/* The exact solve of a coarsest multigrid level of c^3 points (that of a
 * (2c)^3 grid: one work-group up to 16^3, axis passes above); returns c, or
 * -1 if (2c)^3 has no two-level hierarchy. */
int maze_test_mg_coarse(const float *b, float *e, int c);
/* The one-work-group coarse tail of the V-cycle set up for an n^3 grid, on the
 * first level it covers (its size is returned; -1 if the grid has no tail). */
int maze_test_mg_tail(const float *b, float *d, int n);

/* Allocation-free FP64 CG solve of A x = b (the implicit-Jacobi CG of
 * maze_test_conj_grad, library/src/plcg.c:28-90) on caller-owned DEVICE
 * buffers b and x (n^3 doubles each; x holds the start and receives the
 * solution) on `stream` (a hipStream_t, or NULL for the default stream);
 * blocks until the solve has finished.  No counterpart in the reference: for
 * callers that cannot allocate device memory (code around graph captures,
 * tight serving loops, harnesses that forbid hipMalloc).  Its work space is one
 * statically allocated device array per GPU (6.4 MB, held until the process
 * ends): grids of at most 64^3 points.  Calls that overlap on one GPU share
 * that array and race; serialise them.  Returns the 0-based iteration index
 * at convergence, -1 if not converged in n^2 iterations, -2 for n outside
 * 2..64 or a null pointer (nothing is touched), -3 if the work space cannot be
 * found on the current device. */
int maze_conj_grad_noalloc(const double *b, double *x, double tol, int n, void *stream);

/* ====================== Part 3: the replica engine ======================== */
/* Each request is one MaZe-LCG replica: NVE velocity Verlet, CIC charges,
 * Tosi-Fumi pair forces (cutoff L/2), vacuum dielectric, the reference's two
 * start-up steps with cold solves, then `steps` MD steps.  Requests are
 * admitted in arrival order while a slot is free and their device-memory
 * estimate fits the budget; admitted replicas advance concurrently, each
 * step queued as soon as the replica's previous solve state is published.
 * A replica's results do not depend on what else runs (own buffers,
 * deterministic kernels).  The engine's functions are not thread-safe: call
 * them from one thread (maze_engine_run starts its own workers). */
typedef struct maze_engine maze_engine;
/* budget <= 0: 90 % of the free device memory; max_slots: replicas at once. */
maze_engine *maze_engine_create(long long hbm_budget_bytes, int max_slots);
void maze_engine_destroy(maze_engine *e);
/* Host threads that drive the replicas (default 1): each admits and steps
 * its own share of the slots.  Measured on 64 replicas: one thread is
 * fastest (kernel launches from several threads contend in the runtime). */
void maze_engine_set_threads(maze_engine *e, int threads);
/* 1 (default): each replica captures its multigrid-PCG inner iteration (the
 * V-cycle kernels, direction and update) and its queued refinement passes
 * once as HIP graphs and replays them; 0: separate launches.  Results are
 * identical.  Runs with more than one host thread always launch separately
 * (concurrent HIP calls from other threads invalidate a stream capture). */
void maze_engine_set_graphs(maze_engine *e, int on);
/* 1 (default): the replicas' start-up cold solves use the mixed path of their
 * precision (solver_set_mixed_init); 0: FP64 CG as in the reference. */
void maze_engine_set_mixed_init(maze_engine *e, int on);
/* Device bytes one replica needs (grids, multigrid levels, particles, pair
 * partials, scratch), and the precision it runs: 2 (FP32 multigrid-PCG in
 * FP64 refinement) needs an even grid whose halvings end at or below 128
 * points, else 0 (FP64 CG) is used; 1 runs as asked. */
long long maze_engine_estimate_bytes(int n_grid, int np, int precision);
int maze_engine_precision_for(int n_grid, int precision);
/* Queues a request arriving arrival_s seconds after the start of the next
 * maze_engine_run.  Arrays as solver_initialize_particles (tf_params: n_typ^2
 * x 5 Tosi-Fumi parameters in atomic units); dt in a.u.; tol the absolute
 * residual tolerance of the field solve.  Returns the request id, -1 for
 * invalid arguments, -2 if the request alone exceeds the budget. */
int maze_engine_submit(maze_engine *e, double arrival_s, int n_grid, double L, double tol,
                       int precision, int np, int n_typ, const int *types, const double *pos,
                       const double *vel, const double *mass, const double *charges,
                       const double *tf_params, double dt, int steps);
/* Physics of a queued request, replacing submit's defaults (MAZE-LCG without
 * preconditioner, CIC, Tosi-Fumi, velocity Verlet without thermostat); the
 * single-replica path's functions with the same arguments are named.
 * grid_type, precond: solver_initialize_grid's numbering (LCG, MULTIGRID,
 * MAZE-LCG, MAZE-MULTIGRID; NONE, JACOBI, MG for the two CG methods); the
 * mixed precisions apply to MAZE-LCG without preconditioner, the rest runs in
 * FP64.  ca_scheme: solver_initialize_particles' CIC / SPL_QUADR / SPL_CUBIC.
 * eps_s, eps_int: solver_initialize_grid.  itg_type, T, gamma, thermostat:
 * solver_initialize_integrator.  seed (if seed_set): solver_set_rng_seed, the
 * OVRVO noise stream.  rescale: solver_rescale_velocities after the two
 * start-up steps (maze_poisson/solver.py:408-409).  Returns 0, -1 for an
 * unknown or already admitted request or an invalid value, -2 if the request
 * then exceeds the budget (the physics is left unchanged). */
int maze_engine_set_physics(maze_engine *e, int id, int grid_type, int precond, int ca_scheme,
                            double eps_s, double eps_int, int itg_type, double T, double gamma,
                            int thermostat, int seed_set, unsigned long long seed, int rescale);
/* Linear Poisson-Boltzmann on given maps: solver_initialize_grid_pois_boltz(0,
 * kbar2, 0) then solver_set_eps_k2 (n^3 face-centred values each).  Returns
 * as maze_engine_set_physics. */
int maze_engine_set_pb_maps(maze_engine *e, int id, double kbar2, const double *eps_x,
                            const double *eps_y, const double *eps_z, const double *k2);
/* Linear Poisson-Boltzmann with the maps rebuilt from the ion radii every step
 * and the boundary forces in the total: solver_initialize_grid_pois_boltz(w,
 * kbar2, nonpolar) then solver_initialize_particles_pois_boltz(gamma_np,
 * beta_np, radii: np values).  Returns as maze_engine_set_physics. */
int maze_engine_set_pb_radii(maze_engine *e, int id, double w, double kbar2, int nonpolar,
                             double gamma_np, double beta_np, const double *radii);
/* Withdraws a queued request before maze_engine_run.  Returns 0 or -1. */
int maze_engine_cancel(maze_engine *e, int id);
/* Keep the final field phi of a queued request for maze_engine_field (n^3
 * doubles of host memory until the engine is destroyed).  Returns 0 or -1. */
int maze_engine_keep_field(maze_engine *e, int id, int on);
/* Runs until every queued request has finished; returns 0. */
int maze_engine_run(maze_engine *e);
/* Observables of a finished request after its last step, as the
 * single-replica getters give them: get_kinetic_energy, get_temperature,
 * get_energy_elec, the Tosi-Fumi potential energy (solver_compute_forces_noel)
 * and get_momentum (3 values).  Writes up to max of these 7 values and returns
 * their number, or -1 if the request has not finished. */
int maze_engine_observables(maze_engine *e, int id, double *out, int max);
/* The final field phi (n^3) of a finished request kept with
 * maze_engine_keep_field.  Returns 0, or -1. */
int maze_engine_field(maze_engine *e, int id, double *phi);
/* Final positions, velocities and total forces of a finished request (np x 3
 * each; any pointer may be NULL).  Returns 0, or -1 if not finished. */
int maze_engine_result(maze_engine *e, int id, double *pos, double *vel, double *fcs_tot);
/* Per request (up to max values): arrival, admission and finish time (s from
 * the start of the run), steps done, mean / p50 / p99 step latency (ms),
 * estimated and measured device bytes, precision run, unconverged solves,
 * bytes used of the replica's arena (one device allocation per replica),
 * bytes that did not fit it, and the drop of free device memory when the
 * request's streams were created (0 when it reused pooled streams; streams
 * are made four pairs at a time, and the first ones carry the runtime's
 * per-process hardware-queue set-up).  Returns the number of values written. */
int maze_engine_request_stats(maze_engine *e, int id, double *out, int max);
/* Last run (up to max values): wall time (s), replica-steps/s, atom-steps/s,
 * p50 / p99 step latency over all steps (ms), peak estimated device bytes in
 * use, most replicas running at once, host seconds spent admitting (set-up
 * and start-up steps), issuing and finishing steps. */
int maze_engine_stats(maze_engine *e, double *out, int max);

#ifdef __cplusplus
}
#endif
#endif /* MAZE_POISSON_H */

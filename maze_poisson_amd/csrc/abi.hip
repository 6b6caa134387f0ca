// abi.hip -- the C-ABI of libmaze_poisson.so.
//
// The reference symbols keep their names, signatures and semantics
// (library/src/solver_md.c:19-354 and the enum tables of grid_general.c:8-22,
// particles.c:17-36, integrator_general.c:7-15), over one process-global
// GPU-resident replica, so maze_poisson/c_api binds this library unchanged.
// Inputs are copied to HBM at initialisation; getters copy out; nothing else
// crosses the PCIe bus during a step.  Extensions are declared in
// include/maze_poisson.h.
#include <cstring>
#include <vector>

#include "sim.hpp"

using maze::Sim;

static Sim* g_sim = nullptr;

static Sim& S() {
    if (!g_sim) g_sim = new Sim();
    return *g_sim;
}

static void d2h(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    S().sync();
    MAZE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost));
}

static void h2d(void* dst, const void* src, size_t bytes) {
    if (bytes == 0) return;
    S().sync();
    MAZE_HIP(hipMemcpy(dst, src, bytes, hipMemcpyHostToDevice));
}

static char g_grid_str[5][16] = {"LCG", "FFT", "MULTIGRID", "MAZE-LCG", "MAZE-MULTIGRID"};
static char g_precond_str[5][16] = {"NONE", "JACOBI", "MG", "SSOR", "BLOCKJACOBI"};
static char g_pot_str[3][16] = {"TF", "LJ", "SC"};
static char g_cas_str[3][16] = {"CIC", "SPL_QUADR", "SPL_CUBIC"};
static char g_itg_str[2][16] = {"OVRVO", "VERLET"};

extern "C" {

// ---- MPI hooks (serial semantics, library/src/mpi_base.c:169-177) --------
int init_mpi() { return 0; }
int get_rank() { return 0; }

// ---- enum discovery ----------------------------------------------------------
int get_grid_type_num() { return 5; }
char* get_grid_type_str(int i) { return g_grid_str[i]; }
int get_precond_type_num() { return 5; }
char* get_precond_type_str(int i) { return g_precond_str[i]; }
int get_potential_type_num() { return 3; }
char* get_potential_type_str(int i) { return g_pot_str[i]; }
int get_ca_scheme_type_num() { return 3; }
char* get_ca_scheme_type_str(int i) { return g_cas_str[i]; }
int get_integrator_type_num() { return 2; }
char* get_integrator_type_str(int i) { return g_itg_str[i]; }

// ---- initialisation ------------------------------------------------------------
void solver_initialize() {
    int dev = 0;
    hipDeviceProp_t prop;
    MAZE_HIP(hipGetDevice(&dev));
    MAZE_HIP(hipGetDeviceProperties(&prop, dev));
    std::printf("******************************************************\n");
    std::printf("* MAZE_POISSON (HIP, %s)\n", prop.gcnArchName);
    std::printf("******************************************************\n");
    std::fflush(stdout);
    S();
}

void solver_initialize_grid(int n, double L, double h, double tol, double eps, double eps_int,
                            int grid_type, int precond_type) {
    S().init_grid(n, L, h, tol, eps, eps_int, grid_type, precond_type);
}

void solver_initialize_grid_pois_boltz(double w, double kbar2, int nonpolar_enabled) {
    S().init_pb(w, kbar2, nonpolar_enabled);
}

void solver_initialize_particles(int n, int n_typ, double L, double h, int n_p, int pot_type,
                                 int cas_type, int* types, double* pos, double* vel, double* mass,
                                 double* charges, double* pot_params) {
    if (pot_type != 0) {
        std::fprintf(stderr, "maze: only the TF pair potential is provided (got %d)\n", pot_type);
        std::exit(1);
    }
    S().init_particles(n, n_typ, L, h, n_p, cas_type, types, pos, vel, mass, charges, pot_params);
}

void solver_initialize_particles_pois_boltz(double gamma_np, double beta_np, double* radii) {
    S().init_pb_particles(gamma_np, beta_np, radii);
}

void solver_initialize_integrator(int n_p, double dt, double T, double gamma, int itg_type,
                                  int itg_enabled) {
    (void)n_p;
    S().init_integrator(dt, T, gamma, itg_type, itg_enabled);
}

// ---- step ------------------------------------------------------------------------
int solver_update_charges() { return S().update_charges(); }
void solver_init_field() { S().init_field(); }
int solver_update_field() {
    const int r = S().update_field();
    S().check_pb_flag();
    return r;
}
void solver_update_eps_k2() { S().update_eps_k2(); }
void solver_compute_forces_elec() { S().forces_elec(); }
double solver_compute_forces_noel() { return S().forces_noel(true); }
double solver_compute_forces_pb() { return S().forces_pb(); }
void solver_compute_forces_tot() { S().forces_tot(); }
void integrator_part_1() { S().part1(); }
void integrator_part_2() { S().part2(); }
void solver_rescale_velocities() { S().rescale_velocities(); }
int solver_check_thermostat() { return S().check_thermostat(); }

void solver_set_field(double* phi) { h2d(S().phi, phi, S().G * sizeof(double)); }
void solver_set_field_prev(double* phi) { h2d(S().phip, phi, S().G * sizeof(double)); }

void solver_finalize() {
    if (g_sim) {
        delete g_sim;
        g_sim = nullptr;
    }
}

// ---- getters -------------------------------------------------------------------
void get_pos(double* o) { d2h(o, S().p.pos, 3L * S().p.np * sizeof(double)); }
void get_vel(double* o) { d2h(o, S().p.vel, 3L * S().p.np * sizeof(double)); }
void get_fcs_elec(double* o) { d2h(o, S().p.fe, 3L * S().p.np * sizeof(double)); }
void get_fcs_noel(double* o) { d2h(o, S().p.fn, 3L * S().p.np * sizeof(double)); }
void get_fcs_tot(double* o) { d2h(o, S().p.ft, 3L * S().p.np * sizeof(double)); }
// PB boundary forces: zeros until the radii are given (library/src/solver_md.c:263-285).
static void get_pb_force(double* o, const double* src) {
    if (S().pbd.active) d2h(o, src, 3L * S().p.np * sizeof(double));
    else std::memset(o, 0, 3L * S().p.np * sizeof(double));
}
void get_fcs_db(double* o) { get_pb_force(o, S().pbd.fdb); }
void get_fcs_ib(double* o) { get_pb_force(o, S().pbd.fib); }
void get_fcs_np(double* o) { get_pb_force(o, S().pbd.fnp); }
void get_types(int* o) { d2h(o, S().p.types, S().p.np * sizeof(int)); }
void get_charges(double* o) { d2h(o, S().p.charges, S().p.np * sizeof(double)); }
void get_masses(double* o) { d2h(o, S().p.mass, S().p.np * sizeof(double)); }
void get_radii(double* o) {
    if (S().pbd.active) d2h(o, S().pbd.radii, S().p.np * sizeof(double));
    else std::memset(o, 0, S().p.np * sizeof(double));
}
void get_field(double* o) { d2h(o, S().phi, S().G * sizeof(double)); }
void get_field_prev(double* o) { d2h(o, S().phip, S().G * sizeof(double)); }
void get_q(double* o) { d2h(o, S().q, S().G * sizeof(double)); }
double get_kinetic_energy() { return S().kinetic(); }
double get_energy_elec() { return S().energy_elec(); }
double get_energy_elec_qphi() { return S().energy_qphi(); }
void get_momentum(double* o) { S().momentum(o); }
double get_temperature() { return S().temperature(); }

// ---- extensions -------------------------------------------------------------------
void solver_md_loop_iter() { S().md_step(); }

// Block until every launched kernel of the replica has finished.
void solver_sync() { S().sync(); }

void solver_run_n_steps(int n_steps) {
    for (int i = 0; i < n_steps; i++) S().md_step();
}

void solver_set_rng_seed(uint64_t seed) {
    S().seed = seed;
    S().calls = 0;
}

void solver_set_eps_k2(double* ex, double* ey, double* ez, double* k2) {
    S().set_eps_k2(ex, ey, ez, k2);
}

void get_eps_k2(double* ex, double* ey, double* ez, double* k2) {
    Sim& s = S();
    if (!s.pb) {
        std::fprintf(stderr, "maze: get_eps_k2 needs solver_initialize_grid_pois_boltz first\n");
        std::exit(1);
    }
    d2h(ex, s.ex, s.G * sizeof(double));
    d2h(ey, s.ey, s.G * sizeof(double));
    d2h(ez, s.ez, s.G * sizeof(double));
    d2h(k2, s.k2, s.G * sizeof(double));
}

int solver_set_precision(int mode) {
    if (mode < 0 || mode > 3) return -1;
    S().precision = mode;
    S().sched.clear();
    return 0;
}

int solver_get_precision() { return S().precision; }

// Cold solves with the mixed path (1) or the reference's FP64 CG (0, default).
void solver_set_mixed_init(int on) { S().mixed_init = on != 0; }

// Refinement passes of the last mixed-precision solve and each pass's inner
// iterations (negative: stopped by the iteration budget), at most `max`.
int solver_get_last_passes(int* iters, int max) {
    const auto& v = S().sched;
    const int k = std::min((int)v.size(), max);
    for (int i = 0; i < k; i++) iters[i] = v[i];
    return (int)v.size();
}

void solver_set_inner_rtol(double rtol) {
    S().inner_rtol = rtol;
    S().sched.clear();
    S().drop_graphs();
}

// Start each mixed-precision MaZe solve from 2 y_n-1 - y_n-2 (1) or from
// y_n-1 as the reference does (0, default).
void solver_set_y_extrapolation(int on) {
    S().y_extrap = on ? 1 : 0;
    S().hist_ok = 0;
    S().sched.clear();
    S().drop_graphs();
}

// One damping for both sweeps of the multigrid smoother (plain damped Jacobi).
void solver_set_mg_omega(double omega) {
    S().mg_omega = S().mg_omega2 = omega;
    S().sched.clear();
    S().drop_graphs();
}

// Two different sweep weights (a degree-2 polynomial smoother).
void solver_set_mg_omegas(double omega1, double omega2) {
    S().mg_omega = omega1;
    S().mg_omega2 = omega2;
    S().sched.clear();
    S().drop_graphs();
}

// Number of multigrid levels for the current grid (0: unavailable).
int solver_mg_levels() { return S().setup_mg() ? (int)S().mg.size() : 0; }

void solver_set_charge_check(int strict) { S().strict_charge_check = strict; }

void get_y(double* o) { d2h(o, S().y, S().G * sizeof(double)); }

void get_neighbors(int64_t* o) {
    Sim& s = S();
    const int np = s.p.np;
    std::vector<int> base(3L * np);
    d2h(base.data(), s.p.base, 3L * np * sizeof(int));
    static const int off[8][3] = {{0, 0, 0}, {1, 0, 0}, {0, 1, 0}, {0, 0, 1},
                                  {1, 1, 0}, {1, 0, 1}, {0, 1, 1}, {1, 1, 1}};
    const int n = s.p.n, nn = s.p.nn;
    for (int i = 0; i < np; i++)
        for (int c = 0; c < nn; c++)
            for (int d = 0; d < 3; d++) {
                const int o_cd = nn == 8 ? off[c][d]
                                         : (d == 0 ? c >> 4 : (d == 1 ? (c >> 2) & 3 : c & 3)) - 1;
                o[((long)i * nn + c) * 3 + d] = (base[3 * i + d] + o_cd + n) % n;
            }
}

void get_tf_neighbor_counts(int* o) { d2h(o, S().p.counts, S().p.np * sizeof(int)); }

// Per-kernel timing with HIP events on the solver stream (only iterations
// that did work are counted).
void maze_prof_enable(int on) { S().prof = on; }

void maze_prof_reset() {
    for (auto& t : S().timers) {
        t.total_ms = 0.0;
        t.count = 0;
    }
}

int maze_prof_get(const char* name, double* total_ms, long* count) {
    S().prof_flush(1 << 30);
    for (auto& t : S().timers)
        if (std::strcmp(t.name, name) == 0) {
            *total_ms = t.total_ms;
            *count = t.count;
            return 0;
        }
    return -1;
}

int maze_grid_blocks() { return S().dims.blocks; }

// One launch of the plain copy kernel over n_doubles doubles (read n*8 bytes,
// write n*8 bytes): the calibration reference for hardware byte counters.
void maze_calib_copy(long n_doubles) {
    double *a = nullptr, *b = nullptr;
    MAZE_HIP(hipMalloc(&a, n_doubles * sizeof(double)));
    MAZE_HIP(hipMalloc(&b, n_doubles * sizeof(double)));
    maze::launch_fill(S().stream, n_doubles, 1.0, a);
    S().sync();
    maze::launch_copy(S().stream, n_doubles, a, b);
    S().sync();
    MAZE_HIP(hipFree(a));
    MAZE_HIP(hipFree(b));
}

// Bandwidth (GB/s, bytes read + written) of the library's copy kernel over
// n_doubles doubles, averaged over `reps` launches after one warm-up: the
// on-box reference a memory-bound kernel is compared with.
double maze_copy_bandwidth(long n_doubles, int reps) {
    if (n_doubles <= 0 || reps <= 0) return 0.0;
    double *a = nullptr, *b = nullptr;
    hipStream_t s;
    hipEvent_t e0, e1;
    MAZE_HIP(hipStreamCreateWithFlags(&s, hipStreamNonBlocking));
    MAZE_HIP(hipEventCreate(&e0));
    MAZE_HIP(hipEventCreate(&e1));
    MAZE_HIP(hipMalloc(&a, n_doubles * sizeof(double)));
    MAZE_HIP(hipMalloc(&b, n_doubles * sizeof(double)));
    maze::launch_fill(s, n_doubles, 1.0, a);
    maze::launch_copy(s, n_doubles, a, b);
    MAZE_HIP(hipEventRecord(e0, s));
    for (int r = 0; r < reps; r++) maze::launch_copy(s, n_doubles, a, b);
    MAZE_HIP(hipEventRecord(e1, s));
    MAZE_HIP(hipEventSynchronize(e1));
    float ms = 0.0f;
    MAZE_HIP(hipEventElapsedTime(&ms, e0, e1));
    MAZE_HIP(hipFree(a));
    MAZE_HIP(hipFree(b));
    MAZE_HIP(hipEventDestroy(e0));
    MAZE_HIP(hipEventDestroy(e1));
    MAZE_HIP(hipStreamDestroy(s));
    return 16.0 * (double)n_doubles * reps / (ms * 1e-3) / 1e9;
}

// ---- kernel-level entry points on host buffers (for parity tests) ------------------
// Each runs on a private replica and copies in and out; none touches the
// process-global state.
static Sim* scratch_sim(int n) {
    if (n < 2) {
        std::fprintf(stderr, "maze: test grids need n >= 2\n");
        std::exit(1);
    }
    Sim* s = new Sim();
    s->init_grid(n, 1.0, 1.0 / n, 0.0, 1.0, 1.0, n % 2 == 0 && n >= 32 ? 2 : 3, 0);
    return s;
}

void maze_test_laplace(const double* u, double* out, int n) {
    Sim* s = scratch_sim(n);
    MAZE_HIP(hipMemcpy(s->tmp, u, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_laplace(s->stream, s->dims, maze::PBCoef{}, s->tmp, s->r);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->r, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
}

void maze_test_laplace_pb(const double* u, double* out, int n, const double* ex, const double* ey,
                          const double* ez, const double* k2) {
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    MAZE_HIP(hipMemcpy(s->tmp, u, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_laplace(s->stream, s->dims, s->pbc(), s->tmp, s->r);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->r, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
}

int maze_test_conj_grad(const double* b, double* x, double tol, int n) {
    Sim* s = scratch_sim(n);
    MAZE_HIP(hipMemcpy(s->sigma, b, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, x, s->G * sizeof(double), hipMemcpyHostToDevice));
    int pred = 32;
    const int res = s->cg_solve(s->dims, maze::PBCoef{}, s->sigma, s->y, tol, n * n, s->r, s->p0,
                                s->p1, &pred);
    MAZE_HIP(hipMemcpy(x, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return res;
}

int maze_test_conj_grad_pb(const double* b, double* x, double tol, int n, const double* ex,
                           const double* ey, const double* ez, const double* k2) {
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    MAZE_HIP(hipMemcpy(s->sigma, b, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, x, s->G * sizeof(double), hipMemcpyHostToDevice));
    int pred = 32;
    const int res = s->cg_solve(s->dims, s->pbc(), s->sigma, s->y, tol, n * n, s->r, s->p0, s->p1,
                                &pred);
    MAZE_HIP(hipMemcpy(x, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return res;
}

void maze_test_rbgs(const double* in, double* out, int n, int iters) {
    Sim* s = scratch_sim(n);
    MAZE_HIP(hipMemcpy(s->sigma, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, out, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_rbgs(s->stream, s->dims, s->sigma, s->y, iters);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
}

// Multigrid-preconditioned FP32 CG for A x = b from x = 0 (relative stop);
// returns the iteration count, or -1 when no multigrid hierarchy exists for n.
int maze_test_mg_pcg(const double* b, double* x, double rtol, int n) {
    Sim* s = scratch_sim(n);
    s->precision = maze::kMixedMG;
    if (!s->setup_mg()) {
        delete s;
        return -1;
    }
    const long G = s->G;
    std::vector<float> bf(G), xf(G);
    for (long i = 0; i < G; i++) bf[i] = (float)b[i];
    float *db = nullptr, *dx = nullptr, *dr = nullptr, *pa = nullptr, *pb = nullptr;
    MAZE_HIP(hipMalloc(&db, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dx, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dr, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&pa, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&pb, G * sizeof(float)));
    MAZE_HIP(hipMemcpy(db, bf.data(), G * sizeof(float), hipMemcpyHostToDevice));
    s->r32 = dr;
    s->q32a = pa;
    s->q32b = pb;
    const int it = s->inner_solve(true, db, dx, rtol, n * n, 1);
    MAZE_HIP(hipMemcpy(xf.data(), dx, G * sizeof(float), hipMemcpyDeviceToHost));
    for (long i = 0; i < G; i++) x[i] = (double)xf[i];
    s->r32 = s->q32a = s->q32b = nullptr;
    for (float* a : {db, dx, dr, pa, pb}) MAZE_HIP(hipFree(a));
    delete s;
    return it;
}

int maze_test_restrict27(const double* in, double* out, int n) {
    if (n < 2 || n % 2 != 0) return -1;
    Sim* s = scratch_sim(n);
    MAZE_HIP(hipMemcpy(s->tmp, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_restrict27(s->stream, n, s->tmp, s->r);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->r, s->G / 8 * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return 0;
}

int maze_test_prolong_add(const double* in, double* out, int nc) {
    if (nc < 1) return -1;
    Sim* s = scratch_sim(2 * nc);
    MAZE_HIP(hipMemcpy(s->tmp, in, s->G / 8 * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->r, out, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_prolong_add(s->stream, nc, s->tmp, s->r);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->r, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return 0;
}

// Fused multigrid level kernels on FP32 host buffers (n even), sweep weights
// c1 then c2 (the *2 forms) or c for both.
int maze_test_mg_down2(const float* b, float* d2, float* bc, int n, double c1, double c2,
                       int tile) {
    if (n < 2 || n % 2 != 0 || (tile >= maze::kMgTiles && !maze::mg_tile_ok(tile, n))) return -1;
    if (tile < 0) tile = maze::mg_default_tile(n);
    Sim* s = scratch_sim(n);
    const long G = s->G, Gc = G / 8;
    float *db = nullptr, *dd = nullptr, *dc = nullptr;
    MAZE_HIP(hipMalloc(&db, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dd, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dc, Gc * sizeof(float)));
    MAZE_HIP(hipMemcpy(db, b, G * sizeof(float), hipMemcpyHostToDevice));
    maze::launch_mg_down(s->stream, n, tile, db, dd, dc, c1, c2, s->st);
    s->sync();
    MAZE_HIP(hipMemcpy(d2, dd, G * sizeof(float), hipMemcpyDeviceToHost));
    MAZE_HIP(hipMemcpy(bc, dc, Gc * sizeof(float), hipMemcpyDeviceToHost));
    for (float* a : {db, dd, dc}) MAZE_HIP(hipFree(a));
    delete s;
    return 0;
}

int maze_test_mg_down(const float* b, float* d2, float* bc, int n, double c, int tile) {
    return maze_test_mg_down2(b, d2, bc, n, c, c, tile);
}

int maze_test_mg_up2(const float* b, const float* d2, const float* e, float* d, int n, double c1,
                     double c2, int tile) {
    if (n < 2 || n % 2 != 0 || (tile >= maze::kMgTiles && !maze::mg_tile_ok(tile, n))) return -1;
    if (tile < 0) tile = maze::mg_default_tile_up(n);
    Sim* s = scratch_sim(n);
    const long G = s->G, Gc = G / 8;
    float *db = nullptr, *dd2 = nullptr, *de = nullptr, *dd = nullptr;
    MAZE_HIP(hipMalloc(&db, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dd2, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&de, Gc * sizeof(float)));
    MAZE_HIP(hipMalloc(&dd, G * sizeof(float)));
    MAZE_HIP(hipMemcpy(db, b, G * sizeof(float), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(dd2, d2, G * sizeof(float), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(de, e, Gc * sizeof(float), hipMemcpyHostToDevice));
    maze::launch_mg_up(s->stream, n, tile, db, dd2, de, dd, c1, c2, s->st, 0, s->red);
    s->sync();
    MAZE_HIP(hipMemcpy(d, dd, G * sizeof(float), hipMemcpyDeviceToHost));
    for (float* a : {db, dd2, de, dd}) MAZE_HIP(hipFree(a));
    delete s;
    return 0;
}

int maze_test_mg_up(const float* b, const float* d2, const float* e, float* d, int n, double c,
                    int tile) {
    return maze_test_mg_up2(b, d2, e, d, n, c, c, tile);
}

// The one-work-group tail of the V-cycle as set up for an n^3 fine grid: from
// the first level it takes over (returned, or -1 if there is none), on FP32
// host buffers of that level's size.
int maze_test_pb_mg_pcg(const double* b, double* x, double rtol, int n, const double* ex,
                        const double* ey, const double* ez, const double* k2) {
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    s->precision = maze::kMixedMG;
    if (!s->setup_pbmg()) {
        delete s;
        return -1;
    }
    const long G = s->G;
    std::vector<float> bf(G), xf(G);
    for (long i = 0; i < G; i++) bf[i] = (float)b[i];
    float *db = nullptr, *dx = nullptr, *dr = nullptr, *pa = nullptr, *pb = nullptr;
    MAZE_HIP(hipMalloc(&db, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dx, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&dr, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&pa, G * sizeof(float)));
    MAZE_HIP(hipMalloc(&pb, G * sizeof(float)));
    MAZE_HIP(hipMemcpy(db, bf.data(), G * sizeof(float), hipMemcpyHostToDevice));
    s->r32 = dr;
    s->q32a = pa;
    s->q32b = pb;
    const int it = s->inner_solve(true, db, dx, rtol, n * n, 1);
    MAZE_HIP(hipMemcpy(xf.data(), dx, G * sizeof(float), hipMemcpyDeviceToHost));
    for (long i = 0; i < G; i++) x[i] = (double)xf[i];
    s->r32 = s->q32a = s->q32b = nullptr;
    for (float* a : {db, dx, dr, pa, pb}) MAZE_HIP(hipFree(a));
    delete s;
    return it;
}

int maze_test_mg_coarse(const float* b, float* e, int c) {
    Sim* s = scratch_sim(2 * c);
    s->precision = maze::kMixedMG;
    if (!s->setup_mg() || s->mg.size() != 2 || s->mg_c != c) {
        delete s;
        return -1;
    }
    const long g = (long)c * c * c;
    MAZE_HIP(hipMemcpy(s->mg[1].b, b, g * sizeof(float), hipMemcpyHostToDevice));
    s->vcycle(1, 0);  // the coarsest level: its exact solve only
    s->sync();
    MAZE_HIP(hipMemcpy(e, s->mg[1].d, g * sizeof(float), hipMemcpyDeviceToHost));
    delete s;
    return c;
}

int maze_test_mg_tail(const float* b, float* d, int n) {
    Sim* s = scratch_sim(n);
    s->precision = maze::kMixedMG;
    s->mg_stop = 12;  // the hierarchy whose smallest levels run in the tail
    if (!s->setup_mg() || s->mg_tail >= (int)s->mg.size()) {
        delete s;
        return -1;
    }
    const int l = s->mg_tail, nl = (int)s->mg.size() - l;
    int ns[8];
    for (int k = 0; k < nl; k++) ns[k] = s->mg[l + k].n;
    const long g = (long)ns[0] * ns[0] * ns[0];
    MAZE_HIP(hipMemcpy(s->mg[l].b, b, g * sizeof(float), hipMemcpyHostToDevice));
    maze::launch_mg_tail(s->stream, nl, ns, s->mg[l].b, s->mg[l].d, s->mg_omega / 6.0,
                         s->mg_omega2 / 6.0, s->mg_Qf,
                         s->mg_ilam, s->st);
    s->sync();
    MAZE_HIP(hipMemcpy(d, s->mg[l].d, g * sizeof(float), hipMemcpyDeviceToHost));
    delete s;
    return ns[0];
}

void maze_test_rbgs_pb(const double* in, double* out, int n, int iters, const double* ex,
                       const double* ey, const double* ez, const double* k2) {
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    MAZE_HIP(hipMemcpy(s->sigma, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, out, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_rbgs_pb(s->stream, s->dims, s->pbc(), s->sigma, s->y, iters);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
}

int maze_test_restrict_pb_coef(const double* ex, const double* ey, const double* ez,
                               const double* k2, double* cex, double* cey, double* cez,
                               double* ck2, int n) {
    if (n < 2 || n % 2 != 0) return -1;
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    const long Gc = s->G / 8;
    double* d = nullptr;
    MAZE_HIP(hipMalloc(&d, 4 * Gc * sizeof(double)));
    maze::launch_restrict_pb_coef(s->stream, n, s->pbc(), d, d + Gc, d + 2 * Gc, d + 3 * Gc);
    s->sync();
    double* outs[4] = {cex, cey, cez, ck2};
    for (int a = 0; a < 4; a++)
        MAZE_HIP(hipMemcpy(outs[a], d + a * Gc, Gc * sizeof(double), hipMemcpyDeviceToHost));
    MAZE_HIP(hipFree(d));
    delete s;
    return 0;
}

int maze_test_mg_vcycle_pb(const double* in, double* out, int n, const double* ex,
                           const double* ey, const double* ez, const double* k2) {
    if (n < 32 || n % 2 != 0) return -1;
    Sim* s = scratch_sim(n);
    s->init_pb(0.0, 0.0);
    s->set_eps_k2(ex, ey, ez, k2);
    MAZE_HIP(hipMemcpy(s->sigma, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, out, s->G * sizeof(double), hipMemcpyHostToDevice));
    s->mg_vcycle(s->sigma, s->y);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return 0;
}

int maze_test_blockjacobi(const double* in, double* out, int n) {
    if (n < 10 || n % 10 != 0) return -1;
    Sim* s = scratch_sim(n);
    double* tab = nullptr;
    MAZE_HIP(hipMalloc(&tab, maze::kBlockJacobiTab * sizeof(double)));
    std::vector<double> t(maze::kBlockJacobiTab);
    maze::blockjacobi_tables(t.data());
    MAZE_HIP(hipMemcpy(tab, t.data(), maze::kBlockJacobiTab * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->sigma, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    maze::launch_blockjacobi(s->stream, n, s->sigma, s->y, tab);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    MAZE_HIP(hipFree(tab));
    delete s;
    return 0;
}

int maze_test_mg_vcycle(const double* in, double* out, int n) {
    if (n < 32 || n % 2 != 0) return -1;
    Sim* s = scratch_sim(n);
    MAZE_HIP(hipMemcpy(s->sigma, in, s->G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(s->y, out, s->G * sizeof(double), hipMemcpyHostToDevice));
    s->mg_vcycle(s->sigma, s->y);
    s->sync();
    MAZE_HIP(hipMemcpy(out, s->y, s->G * sizeof(double), hipMemcpyDeviceToHost));
    delete s;
    return 0;
}

}  // extern "C"

// sim.hpp -- one simulation replica resident on the GPU.
#pragma once

#include <vector>

#include "grid.hpp"
#include "particles.hpp"
#include "pb.hpp"

namespace maze {

enum GridType { kLCG = 0, kFFT = 1, kMG = 2, kMazeLCG = 3, kMazeMG = 4 };
enum PrecondType { kPNone = 0, kPJacobi = 1, kPMG = 2, kPSSOR = 3, kPBlockJacobi = 4 };
enum ItgType { kOVRVO = 0, kVerlet = 1 };
// Arithmetic of the field solve: the reference FP64 CG, or FP32-storage CG
// inside FP64 iterative refinement.
// kFP64MG: FP64 CG preconditioned by kMixedMG's V-cycle carried out in FP64
// (the same solver without the FP32 storage; Poisson operator only).
enum Precision { kFP64 = 0, kMixedCG = 1, kMixedMG = 2, kFP64MG = 3 };

// A block of device memory that replica allocations are carved from while it
// is installed (arena_install) on the allocating thread: one hipMalloc per
// replica and an exact account of its footprint.  dfree leaves arena memory
// alone; the owner frees the block after arena_forget.
constexpr size_t kArenaAlign = 256;
struct DevArena {
    char* base = nullptr;
    size_t size = 0, used = 0;
    size_t overflow = 0;  // bytes that did not fit and came from hipMalloc
};
void arena_install(DevArena* a);  // nullptr: plain hipMalloc again
void arena_forget(const DevArena* a);

// One level of the FP32 multigrid preconditioner.
struct MGLevel {
    int n;
    GridDims dims;
    float* b;  // right-hand side (level 0: the CG residual)
    float* d;  // correction / cycle output
    float* t;  // residual and Jacobi ping-pong buffer
    int tile;     // MgTile of the fused down kernel
    int tile_up;  // MgTile of the fused up kernel
};

// Per-kernel timing collected with HIP events on the solver stream.
enum TimerId {
    kTCgDir = 0, kTCgUpdate, kTCgDirF32, kTCgUpdateF32, kTMgDown, kTMgUp, kTRefine, kTPred, kTTf,
    kTCic, kTPbMaps, kTPbForces
};
struct KernelTimer {
    const char* name;
    double total_ms = 0.0;
    long count = 0;
    double bytes_per_launch = 0.0;
};

// The launch shape of the grid passes of an n^3 grid (planes per work-group
// sized for about 512 work-groups per sweep); every replica uses it.
GridDims grid_dims_for(int n);

struct Sim {
    hipStream_t stream = nullptr;

    // ---- grid ----
    int n = 0;
    double L = 0, h = 0, tol = 0, eps_s = 1, eps_int = 1;
    int grid_type = kMazeLCG, precond = kPNone;
    int precision = kFP64;
    // Relative residual reduction of each inner solve; <= 0 selects the
    // default of the mode (kCgInnerRtol for Jacobi-CG, kMgInnerRtol for the
    // multigrid-preconditioned CG).
    double inner_rtol = 0.0;
    // Iteration budget of each inner solve of the mixed path (<= 0: the
    // default of the mode, n^2 for Jacobi-CG and kMgInnerMax for multigrid-
    // PCG); MAZE_INNER_MAX sets it.
    int inner_max = 0;
    // Multigrid-PCG inner solves reduce their residual by 1e-6: one
    // refinement pass of 4 V-cycles then meets the FP64 stop at 4096 ions /
    // 128^3, where 1e-3 took two passes of 2 + 3 cycles (2124 against 1729
    // steps/s; 3e-6 .. 5e-7 measured alike, 1e-5 needs 3 + 5 cycles); the
    // PB paths and the ensemble gain alike.  An FP32 solve does not stall in
    // its recursive residual, but the budget bounds a pass anyway.
    static constexpr double kMgInnerRtol = 1e-6, kCgInnerRtol = 1e-4;
    static constexpr int kMgInnerMax = 40;
    double inner_rtol_for(bool use_mg) const {
        return inner_rtol > 0 ? inner_rtol : (use_mg ? kMgInnerRtol : kCgInnerRtol);
    }
    int inner_limit_for(bool use_mg) const {
        return inner_max > 0 ? inner_max : (use_mg ? kMgInnerMax : n * n);
    }
    int pb = 0;
    int nonpolar = 0;
    double w = 0, kbar2 = 0;
    long G = 0;
    GridDims dims{};
    double *q = nullptr, *y = nullptr, *phi = nullptr, *phip = nullptr, *tmp = nullptr;
    double *sigma = nullptr, *r = nullptr, *p0 = nullptr, *p1 = nullptr, *y_alt = nullptr;
    double* y_hist = nullptr;  // y of the step before last (extrapolated start)
    // Coarse-grid PB coefficients of the reference two-grid cycle.
    double *cex = nullptr, *cey = nullptr, *cez = nullptr, *ck2 = nullptr;
    int y_extrap = 0;          // start the mixed solve from 2 y_n-1 - y_n-2
    int hist_ok = 0;           // y_hist holds the previous step's y
    double *ex = nullptr, *ey = nullptr, *ez = nullptr, *k2 = nullptr;
    float *d32 = nullptr, *r32 = nullptr, *q32a = nullptr, *q32b = nullptr;
    // reference multigrid coarse level
    int nc = 0;
    GridDims cdims{};
    double *mg_r = nullptr, *c_rhs = nullptr, *c_eps = nullptr, *c_r = nullptr, *c_p0 = nullptr,
           *c_p1 = nullptr;

    // FP32 multigrid preconditioner
    std::vector<MGLevel> mg;
    double* mg_Qf = nullptr;
    double* mg_ilam = nullptr;
    double *mg_w0 = nullptr, *mg_w1 = nullptr;  // work arrays of a coarsest level above 12^3
    int mg_c = 0;
    // Arrival counter of the down kernel that also solves a 16^3 coarsest level
    // (launch_mg_down_coarse, opt-in with MAZE_MG_FUSE_COARSE=1: bit-identical to
    // the two kernels, measured 0.9 % slower on the 4096-ion / 128^3 step).
    unsigned int* mg_coarse_count = nullptr;
    int mg_fuse_coarse = 0;
    // Weights of the two Jacobi smoothing sweeps (pre: omega then omega2;
    // post: omega2 then omega).  Default: the degree-2 Chebyshev polynomial
    // for the Jacobi-scaled spectrum [1/3, 2] (the weights are the reciprocal
    // Chebyshev nodes), which damps the upper two thirds of the spectrum
    // by >= 1/T2(7/5) = 0.34 per pair of sweeps against 0.51 for two sweeps
    // at 6/7; measured: 5 V-cycles per MD step at 128^3 instead of 6.
    double mg_omega = cheb_weight(1.0 / 3.0, 0), mg_omega2 = cheb_weight(1.0 / 3.0, 1);
    static double cheb_weight(double lo, int k) {
        const double mid = (2.0 + lo) / 2.0, half = (2.0 - lo) / 2.0 * 0.70710678118654752;
        return 1.0 / (k == 0 ? mid + half : mid - half);
    }
    int mg_ok = -1;             // -1 not set up, 0 unavailable for this n, 1 ready
    // FP32 multigrid for the linear PB operator (precision kMixedMG with PB):
    // per level FP32 face coefficients (level 0: a copy of the maps), the
    // inverse diagonal and the cycle's vectors.
    struct PbLevel {
        int n = 0;
        GridDims dims{};
        PBCoef32 cf{};
        float *ex = nullptr, *ey = nullptr, *ez = nullptr, *k2 = nullptr;
        float *g = nullptr, *b = nullptr, *d = nullptr, *t = nullptr;
    };
    std::vector<PbLevel> pbmg;
    int pbmg_ok = -1;           // -1 not set up, 0 unavailable, 1 ready
    int pbmg_dirty = 1;         // the maps changed since the coefficients were built
    int pb_coarse_iters = 24;   // Chebyshev iterations on the coarsest level
    // The two smallest PB levels in one work-group (launch_pb_tail) when the
    // level above the coarsest has at most 20 points per side and is not the
    // finest (MAZE_PB_TAIL=1; the same values as the separate passes).  Off by
    // default: at 8000 ions / 160^3 the work-group took 62 us per cycle
    // against 44.5 us for the seven launches it replaces (one CU streams the
    // level's coefficient arrays from L2), 644 against 688 steps/s.
    int pb_tail = 0;
    double pb_coarse_lo = 0.02; // their lower spectrum bound (upper: 2)
    // Replay each multigrid-PCG inner iteration as a captured HIP graph (three
    // variants: the first iteration, odd and even ones) instead of its seven
    // or more separate launches: the kernels and their arguments are fixed
    // for a replica (buffers, weights; the stop flags are read on the device).
    // Dropped with the multigrid levels (free_mg) and by set_mg_omegas.
    int use_graphs = 0;
    hipGraphExec_t inner_graph[3] = {nullptr, nullptr, nullptr};
    // Whole queued refinement passes (m inner iterations and the refinement)
    // as graphs, per length m <= kPassGraphs and order of the two y buffers.
    static constexpr int kPassGraphs = 6;
    hipGraphExec_t pass_graph[2 * kPassGraphs] = {};
    // The refinement arguments each pass graph was captured with (they are
    // kernel arguments, fixed in the graph): a pass whose arguments differ
    // (new tolerance, inner stop or budget) is captured again.
    RefineArgs pass_graph_ra[2 * kPassGraphs] = {};
    int mg_tail = 0;            // first level handled by the one-work-group tail kernel

    CGState* st = nullptr;      // device
    // State of the coarse CG inside the reference V-cycle, apart from st so
    // that a V-cycle can precondition an outer CG (MAZE-LCG with precond MG).
    CGState* st_c = nullptr;    // device
    double* pv = nullptr;       // preconditioned residual v of the explicit PCG
    double* bj_tab = nullptr;   // block-Jacobi eigenbasis tables (blockjacobi_tables)
    CGState* st_host = nullptr; // pinned
    StatePub* pub_host = nullptr;  // host-coherent, written by the GPU (read_state)
    StatePub* pub_dev = nullptr;   // its device address
    unsigned int pub_seq = 0;
    unsigned long pub_count = 0;   // states read so far
    unsigned int want_seq = 0;     // sequence number of the last publication queued

    // A CG solve in progress (cg_issue .. cg_finish).
    struct CgRun {
        GridDims d;
        PBCoef pbc;
        double* x;
        double* rbuf;
        double *pin, *pout;
        int limit, launched, chunk;
        int* pred;
    } cg{};
    // A mixed-precision solve in progress (mixed_issue .. mixed_finish).
    struct MixedRun {
        bool use_mg = false, extrap = false;
        double *ya = nullptr, *yb = nullptr;
        RefineArgs ra{};
        int inner_limit = 0, queued = 0;
    } mx;
    // An MD step in progress (step_issue .. step_finish).  spec_post: the
    // field update and the force / half-kick were queued behind the solve,
    // gated on the device by the outer loop's stop flag; spec_hit: they ran
    // (the published state says the loop had converged).
    int step_tf = 0, step_wait = 0, step_res = 0, spec_post = 0, spec_hit = 0;
    Reduce red{};
    // Partials of the multigrid-PCG inner iteration (deferred finalisation).
    DeferRed dred{};
    // The Tosi-Fumi forces depend on the positions only, so one MD step
    // (md_step) runs them on a second stream (default priority) concurrently
    // with the charge spread and the field solve, with their own reduction
    // buffers; the total force waits for them.  MAZE_TF_OVERLAP=0 turns it off.
    // They start after the charge spread (tf_late), so they share the GPU with
    // the field solve instead of with the spread's latency-bound gather:
    // measured 1420 against 1403 steps/s at 4096 ions / 128^3.  MAZE_TF_LATE=0
    // starts them with the spread.
    hipStream_t side = nullptr;
    hipEvent_t ev_pos = nullptr, ev_tf = nullptr;
    Reduce red_side{};
    int tf_overlap = 1;
    int tf_late = 1;

    double* dscalar = nullptr;  // device scratch scalar
    double* hscalar = nullptr;  // pinned

    int last_iters = 32;        // prediction for the chunked convergence polling
    int last_inner = 32;
    // Inner iterations of each refinement pass of the last mixed solve
    // (negative: the pass stopped on its iteration budget, not its tolerance).
    std::vector<int> sched;
    static constexpr int kMaxOuter = 50;  // refinement passes per solve
    int last_coarse = 16;
    int last_res = 0;

    // ---- particles ----
    PDev p{};
    PScalars* ps = nullptr;      // device
    PScalars* ps_host = nullptr; // pinned
    // The last charge spread left q zero except at its own nodes (the next one
    // clears those instead of the whole grid).
    int spread_zero_old = 0;
    double q_tot = 0.0;          // library/src/solver_md.c:16
    int strict_charge_check = 1;
    // Poisson-Boltzmann boundary from ion radii (maps and boundary forces).
    PBDev pbd;
    int pb_flag_pending = 0;     // a map build whose overflow flag is still to be read

    // ---- integrator ----
    int itype = kVerlet, enabled = 0;
    double dt = 0, T = 0, c1 = 1, c2 = 1;
    uint64_t seed = 42, calls = 0;

    // ---- profiling ----
    int prof = 0;
    std::vector<hipEvent_t> ev_pool;
    size_t ev_next = 0;
    struct Pending { int timer; hipEvent_t a, b; int iter, pass; };
    int cur_pass = -1;  // refinement pass being queued (-1: none)
    std::vector<Pending> pending;
    std::vector<KernelTimer> timers = {{"cg_dir"},     {"cg_update"},  {"cg_dir_f32"},
                                       {"cg_update_f32"}, {"mg_down_l0"}, {"mg_up_l0"},
                                       {"refine"},     {"pred_sigma"}, {"tf_pairs"},
                                       {"cic"},        {"pb_maps"},    {"pb_forces"}};
    // When profiling is on, the first kernel `launch` issues records its own
    // start and end in a pair of HIP events (MAZE_LAUNCH); `iter` tags solver
    // iterations so that launches after the stop (no-ops) can be left out.
    template <class F>
    void timed(int timer, int iter, F&& launch) {
        if (!prof) {
            launch();
            return;
        }
        hipEvent_t a = next_event(), b = next_event();
        g_launch_timer.start = a;
        g_launch_timer.stop = b;
        launch();
        if (g_launch_timer.start) {  // no kernel was launched: an empty interval
            g_launch_timer.start = nullptr;
            MAZE_HIP(hipEventRecord(a, stream));
            MAZE_HIP(hipEventRecord(b, stream));
        }
        pending.push_back({timer, a, b, iter, cur_pass});
    }

    Sim();
    Sim(hipStream_t main, hipStream_t side);
    ~Sim();
    int own_streams = 1;

    void init_grid(int n, double L, double h, double tol, double eps, double eps_int, int type,
                   int precond);
    void init_pb(double w, double kbar2, int nonpolar = 0);
    void set_eps_k2(const double* ex, const double* ey, const double* ez, const double* k2);
    void init_particles(int n, int n_typ, double L, double h, int np, int cas, const int* types,
                        const double* pos, const double* vel, const double* mass,
                        const double* charges, const double* pot_params);
    void init_integrator(double dt, double T, double gamma, int type, int enabled);
    void init_pb_particles(double gamma_np, double beta_np, const double* radii);
    void update_eps_k2();
    double forces_pb();
    void check_pb_flag();
    void free_pb();

    // base_done: the base cells were formed by part1(true).
    int update_charges(bool defer = false, bool base_done = false);
    int check_charge();
    void init_field();
    void init_field_issue();
    bool init_field_ready() const;
    void init_field_finish();
    void cold_issue();
    void cold_finish();
    // Cold solves (init_field) with the mixed path when precision != kFP64
    // (replica engine; solver_set_mixed_init), instead of the reference's FP64 CG.
    int mixed_init = 0;
    int cold_mixed = 0;  // the cold solve in progress is the mixed one
    int cold_pred = 32;  // FP64 cold solve: iterations of the first chunk
    static constexpr int kColdPasses = 4, kColdIters = 6;
    int update_field();
    void forces_elec();
    double forces_noel(bool read_back = true);
    void forces_tot();
    // with_base (the MD step's path): Verlet's first half also forms the base
    // cells of the charge assignment; returns whether it did.
    bool part1(bool with_base = false);
    void part2();
    void md_step();
    // md_step in two halves for callers that advance several replicas from
    // one thread: step_issue queues the step up to the solver-state
    // publication and returns at once, step_ready polls for that state, and
    // step_finish completes the step (md_step = step_issue + step_finish).
    void step_issue();
    bool step_ready() const;
    int step_finish();
    void rescale_velocities();
    int check_thermostat();
    double kinetic();
    double temperature();  // 2 K / (3 N k_B)
    double pe_tf();        // Tosi-Fumi potential energy of the last pair-force pass
    double energy_elec();
    double energy_qphi();
    void momentum(double* out);

    // solvers
    int cg_solve(const GridDims& d, const PBCoef& pbc, const double* b, double* x, double tol,
                 int limit, double* rbuf, double* pa, double* pbuf, int* pred);
    int cg_count_solve(const GridDims& d, const PBCoef& cpb, const double* b, double* x,
                       double rtol, int maxit, double* rbuf, double* pa, double* pbuf, int* pred,
                       CGState* cst);
    // MAZE-LCG with the reference's multigrid preconditioner (FP64 only).
    int maze_pcg_mg();
    int pcg_mg_solve(const double* b, double* x, double tol, int limit);
    void precond_apply(const double* r, double* v);
    // MAZE-LCG in FP64 with an explicit preconditioner (MG or BLOCKJACOBI):
    // the reference's conj_grad_precond (library/src/plcg.c:104-166).
    bool explicit_pcg() const {
        return grid_type == kMazeLCG && precision == kFP64 &&
               (precond == kPMG || precond == kPBlockJacobi) && !pb;
    }
    void restrict_pb_coef();
    void inner_iter(bool use_mg, float* x, int it);
    void inner_body(bool use_mg, float* x, int it);
    void drop_graphs();
    void inner_start(bool use_mg, const float* b, float* x, double rtol, int limit, int gated);
    int inner_solve(bool use_mg, const float* b, float* x, double rtol, int limit, int first_chunk);
    bool setup_mg();
    void tune_mg_tiles();
    // Pick each level's tile shapes by timing (replica engine): the down
    // kernels and the up kernels below the finest level, which reduce nothing,
    // so the results do not depend on the choice.  Cached per level size.
    int mg_autotune = 0;
    bool setup_pbmg();
    void free_pbmg();
    void pbmg_coefs();
    void vcycle_pb(int level);
    // Level sizes of the FP32 multigrid for an n^3 grid (empty: unavailable).
    static std::vector<int> mg_sizes(int n, int stop = default_mg_stop());
    static int default_mg_stop();
    int mg_stop = default_mg_stop();  // see mg_sizes
    void free_mg();
    void vcycle(int level, int iter);
    // FP64 multigrid-PCG (precision kFP64MG): the levels of mg with FP64
    // arrays; the V-cycle of vcycle() with FP64 operands preconditions FP64
    // CG from the previous solution to the reference's stop test.
    struct MGLevel64 {
        int n;
        double *b, *d, *t;
    };
    std::vector<MGLevel64> mg64;
    int mg64_ok = -1;
    int mg64_launched = 0;
    int mg64_pred = 4;  // iterations queued before the first poll (last step's count)
    bool setup_mg64();
    void vcycle64(int level);
    void fp64mg_issue();
    void fp64mg_chunk(int count);
    int fp64mg_finish();
    bool use_fp64mg();
    int step_mg64 = 0;
    // The inner iteration's stop test is deferred to the next consumer
    // (multigrid-PCG on the Poisson operator); a host read of the state in the
    // middle of an inner solve applies it first.
    bool inner_deferred(bool use_mg) const { return use_mg && !pb; }
    void inner_read_state(bool use_mg);
    int maze_fp64();
    void fp64_issue();
    int fp64_finish();
    int maze_mixed();
    void mixed_issue();
    int mixed_finish();
    void mixed_refine(int add);
    void mixed_pass(int k, int m, bool poll);
    void cg_issue(const CgRun& c, const double* b, double tol);
    void cg_chunk();
    int cg_finish();
    int lcg_update();
    int mg_update(bool maze);
    void mg_vcycle(const double* in, double* out, int sm = 0);
    PBCoef pbc() const;

    void sync() { MAZE_HIP(hipStreamSynchronize(stream)); }
    void read_state(const CGState* which = nullptr);  // publish + wait_published
    // Queue the publication of the state (sequence want_seq): st, or `which`.
    void publish(const CGState* which = nullptr);
    bool published() const;    // the state of want_seq has arrived
    void wait_published();     // spin until it has, then copy it to st_host
    hipEvent_t next_event();
    void prof_flush(int iters_done);
    void prof_flush_passes(const int* iters, int npass);
    void free_grid();
    void free_particles();
};

}  // namespace maze

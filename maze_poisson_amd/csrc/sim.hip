// sim.hip -- host-side driver of one GPU-resident replica.
//
// All state lives in HBM; the host only launches kernels and, once per
// solve (in chunks of iterations), reads the small CGState block to learn
// whether the device-side stop test has fired.  Kernels launched after the
// stop are no-ops (they test st->done first), so the state is exactly the
// state at convergence.
#include <algorithm>
#include <chrono>
#include <cmath>
#include <cstring>
#include <map>
#include <mutex>

#include "sim.hpp"

namespace maze {

thread_local LaunchTimer g_launch_timer;

static constexpr double kBoltz = 3.1668e-6;  // Hartree / K, library/include/constants.h:10
static constexpr int kMaxPartials = 1 << 15;

static void fatal(const char* msg) {
    std::fprintf(stderr, "%s\n", msg);
    std::fflush(stderr);
    std::exit(1);
}

// Device arenas: while one is installed on the calling thread, replica
// allocations are carved from it (256-byte aligned) instead of hipMalloc;
// allocations that do not fit fall back to hipMalloc and are counted.
namespace {
thread_local DevArena* t_arena = nullptr;
std::mutex g_arena_mu;
std::vector<const DevArena*> g_arenas;  // live arenas, for dfree

bool in_arena(const void* p) {
    const char* c = static_cast<const char*>(p);
    std::lock_guard<std::mutex> g(g_arena_mu);
    for (const DevArena* a : g_arenas)
        if (c >= a->base && c < a->base + a->size) return true;
    return false;
}
}  // namespace

void arena_install(DevArena* a) {
    t_arena = a;
    std::lock_guard<std::mutex> g(g_arena_mu);
    if (a && std::find(g_arenas.begin(), g_arenas.end(), a) == g_arenas.end()) g_arenas.push_back(a);
}

void arena_forget(const DevArena* a) {
    if (t_arena == a) t_arena = nullptr;
    std::lock_guard<std::mutex> g(g_arena_mu);
    g_arenas.erase(std::remove(g_arenas.begin(), g_arenas.end(), a), g_arenas.end());
}

template <class T>
static T* dalloc(long count) {
    T* p = nullptr;
    if (count <= 0) count = 1;
    const size_t bytes = sizeof(T) * (size_t)count;
    if (DevArena* a = t_arena) {
        const size_t off = (a->used + kArenaAlign - 1) & ~(kArenaAlign - 1);
        if (off + bytes <= a->size) {
            p = reinterpret_cast<T*>(a->base + off);
            a->used = off + bytes;
        } else {
            a->overflow += bytes;
        }
    }
    if (!p) MAZE_HIP(hipMalloc(&p, bytes));
    MAZE_HIP(hipMemset(p, 0, bytes));
    return p;
}

template <class T>
static void dfree(T*& p) {
    if (p && !in_arena(p)) MAZE_HIP(hipFree(p));
    p = nullptr;
}

// Planes per work-group so that one sweep launches about 512 work-groups.
// Measured at 128^3: 16 planes per work-group (512 work-groups) is fastest;
// 8 planes is 1.3x slower and 4 planes 1.8x slower (each work-group reloads
// its two halo planes and pays the reduction hand-off), 64 planes is 2.6x
// slower (too few waves).
static int choose_slab(int n) {
    const long chunks = ((long)n * n + kBlock - 1) / kBlock;
    static const long target = [] {
        const char* e = std::getenv("MAZE_SLAB_WG");  // tuning override
        const long v = e ? std::atol(e) : 0;
        return v > 0 ? v : 512L;
    }();
    long s = chunks * n / target;
    if (s < 1) s = 1;
    if (s > 128) s = 128;
    return (int)s;
}

GridDims grid_dims_for(int n) { return make_dims(n, choose_slab(n)); }

Sim::Sim() : Sim(nullptr, nullptr) {}

// main / side: streams the caller keeps (both or neither), for callers that
// create many replicas (stream creation costs milliseconds).
Sim::Sim(hipStream_t main, hipStream_t side_) {
    own_streams = main == nullptr;
    if (own_streams) {
        MAZE_HIP(hipStreamCreateWithFlags(&stream, hipStreamNonBlocking));
        MAZE_HIP(hipStreamCreateWithFlags(&side, hipStreamNonBlocking));
    } else {
        stream = main;
        side = side_;
    }
    MAZE_HIP(hipEventCreateWithFlags(&ev_pos, hipEventDisableTiming));
    MAZE_HIP(hipEventCreateWithFlags(&ev_tf, hipEventDisableTiming));
    red_side.psum = dalloc<double>(kMaxPartials);
    red_side.pmax = dalloc<double>(kMaxPartials);
    red_side.counter = dalloc<unsigned int>(kArrivalWords);
    if (const char* e = std::getenv("MAZE_TF_OVERLAP")) tf_overlap = std::atoi(e) != 0;
    if (const char* e = std::getenv("MAZE_TF_LATE")) tf_late = std::atoi(e) != 0;
    if (const char* e = std::getenv("MAZE_INNER_MAX")) inner_max = std::atoi(e);
    if (const char* e = std::getenv("MAZE_INNER_RTOL")) inner_rtol = std::atof(e);  // tuning
    if (const char* e = std::getenv("MAZE_PB_TAIL")) pb_tail = std::atoi(e) != 0;
    if (const char* e = std::getenv("MAZE_MG_FUSE_COARSE")) mg_fuse_coarse = std::atoi(e) != 0;
    st = dalloc<CGState>(1);
    st_c = dalloc<CGState>(1);
    MAZE_HIP(hipHostMalloc(&st_host, sizeof(CGState)));
    std::memset(st_host, 0, sizeof(CGState));
    red.psum = dalloc<double>(kMaxPartials);
    red.pmax = dalloc<double>(kMaxPartials);
    red.counter = dalloc<unsigned int>(kArrivalWords);
    dred.pz = dalloc<double>(kMaxPartials);
    dred.pp = dalloc<double>(kMaxPartials);
    dred.pus = dalloc<double>(kMaxPartials);
    dred.pum = dalloc<double>(kMaxPartials);
    dscalar = dalloc<double>(8);
    MAZE_HIP(hipHostMalloc(&hscalar, 8 * sizeof(double)));
    ps = dalloc<PScalars>(1);
    MAZE_HIP(hipHostMalloc(&ps_host, sizeof(PScalars)));
    MAZE_HIP(hipHostMalloc(&pub_host, sizeof(StatePub), hipHostMallocCoherent | hipHostMallocMapped));
    std::memset(pub_host, 0, sizeof(StatePub));
    MAZE_HIP(hipHostGetDevicePointer(reinterpret_cast<void**>(&pub_dev), pub_host, 0));
}

Sim::~Sim() {
    // Nothing may still run on either stream, or sit in a graph, when the
    // buffers it uses are freed.
    sync();
    MAZE_HIP(hipStreamSynchronize(side));
    drop_graphs();
    free_grid();
    free_particles();
    dfree(st);
    dfree(st_c);
    dfree(red.psum);
    dfree(red.pmax);
    dfree(red.counter);
    dfree(dred.pz);
    dfree(dred.pp);
    dfree(dred.pus);
    dfree(dred.pum);
    dfree(red_side.psum);
    dfree(red_side.pmax);
    dfree(red_side.counter);
    dfree(dscalar);
    dfree(ps);
    if (st_host) MAZE_HIP(hipHostFree(st_host));
    if (hscalar) MAZE_HIP(hipHostFree(hscalar));
    if (ps_host) MAZE_HIP(hipHostFree(ps_host));
    if (pub_host) MAZE_HIP(hipHostFree(pub_host));
    for (auto e : ev_pool) MAZE_HIP(hipEventDestroy(e));
    MAZE_HIP(hipEventDestroy(ev_pos));
    MAZE_HIP(hipEventDestroy(ev_tf));
    if (own_streams) {
        MAZE_HIP(hipStreamDestroy(side));
        MAZE_HIP(hipStreamDestroy(stream));
    }
}

void Sim::free_grid() {
    for (double** a : {&q, &y, &phi, &phip, &tmp, &sigma, &r, &p0, &p1, &y_alt, &ex, &ey, &ez, &k2,
                       &mg_r, &c_rhs, &c_eps, &c_r, &c_p0, &c_p1, &pv, &bj_tab})
        dfree(*a);
    for (float** a : {&d32, &r32, &q32a, &q32b}) dfree(*a);
    for (double** a : {&cex, &cey, &cez, &ck2}) dfree(*a);
    free_mg();
    free_pbmg();
    sched.clear();
    dfree(y_hist);
    hist_ok = 0;
    G = 0;
}

void Sim::free_particles() {
    dfree(p.types);
    for (double** a : {&p.pos, &p.vel, &p.mass, &p.charges, &p.fe, &p.fn, &p.ft, &p.table,
                       &p.tpart, &p.eion, &p.contrib})
        dfree(*a);
    dfree(p.base);
    dfree(p.head);
    dfree(p.node);
    dfree(p.link);
    dfree(p.counts);
    p.np = 0;
    free_pb();
}

void Sim::free_pb() {
    for (double** a : {&pbd.radii, &pbd.rp2, &pbd.rm2, &pbd.fdb, &pbd.fib, &pbd.fnp}) dfree(*a);
    for (int** a : {&pbd.rmap, &pbd.rfrc, &pbd.ids, &pbd.ids_sorted}) dfree(*a);
    dfree(pbd.keys);
    dfree(pbd.keys_sorted);
    dfree(pbd.sort_tmp);
    pbd = PBDev{};
    pb_flag_pending = 0;
}

PBCoef Sim::pbc() const {
    if (!pb) return PBCoef{nullptr, nullptr, nullptr, nullptr};
    return PBCoef{ex, ey, ez, k2};
}

// library/src/grid_general.c:24-83 and the per-type init functions.
void Sim::init_grid(int n_, double L_, double h_, double tol_, double eps, double eps_int_,
                    int type, int precond_) {
    if (type == kFFT)
        fatal("maze: the FFT grid (continuous-symbol FFTW solve) is not provided by this library");
    if (type < 0 || type > 4) fatal("maze: invalid grid type");
    if ((type == kMG || type == kMazeMG) && (n_ % 2 != 0 || n_ < 32))
        fatal("Multigrid: requires an even grid of at least 32 points per side");
    if (type == kMazeLCG && precond_ == kPSSOR)
        fatal("maze: MAZE-LCG provides the NONE, JACOBI, MG and BLOCKJACOBI preconditioners "
              "(SSOR is not provided)");
    if (type == kMazeLCG && precond_ > kPBlockJacobi) fatal("maze: invalid preconditioner type");
    // library/src/precond_blockjacobi.c:91-95 exits unless the grid is whole blocks.
    if (type == kMazeLCG && precond_ == kPBlockJacobi && n_ % 10 != 0)
        fatal("Error: The size of the input vector is not a multiple of the block size (10)");
    // library/src/multigrid.c:114-121 exits when the first halving falls below 16.
    if (type == kMazeLCG && precond_ == kPMG && (n_ % 2 != 0 || n_ < 32))
        fatal("Multigrid preconditioner: requires an even grid of at least 32 points per side");
    free_grid();
    n = n_; L = L_; h = h_; tol = tol_; eps_s = eps; eps_int = eps_int_;
    grid_type = type; precond = precond_;
    pb = 0;
    G = (long)n * n * n;
    dims = grid_dims_for(n);
    if (dims.blocks > kMaxPartials) fatal("maze: grid too large for the reduction buffers");
    q = dalloc<double>(G); y = dalloc<double>(G); phi = dalloc<double>(G);
    spread_zero_old = 0;
    phip = dalloc<double>(G); tmp = dalloc<double>(G); sigma = dalloc<double>(G);
    r = dalloc<double>(G); p0 = dalloc<double>(G); p1 = dalloc<double>(G);
    if (type == kMG || type == kMazeMG || (type == kMazeLCG && precond_ == kPMG)) {
        nc = n / 2;
        cdims = make_dims(nc, choose_slab(nc));
        const long Gc = (long)nc * nc * nc;
        mg_r = dalloc<double>(G);
        if (type == kMazeLCG) pv = dalloc<double>(G);
        c_rhs = dalloc<double>(Gc); c_eps = dalloc<double>(Gc); c_r = dalloc<double>(Gc);
        c_p0 = dalloc<double>(Gc); c_p1 = dalloc<double>(Gc);
    }
    if (type == kMazeLCG && precond_ == kPBlockJacobi) {
        pv = dalloc<double>(G);
        bj_tab = dalloc<double>(kBlockJacobiTab);
        std::vector<double> t(kBlockJacobiTab);
        blockjacobi_tables(t.data());
        MAZE_HIP(hipMemcpy(bj_tab, t.data(), kBlockJacobiTab * sizeof(double),
                           hipMemcpyHostToDevice));
    }
    last_iters = 32;
    sync();
}

// library/src/grid_general.c:85-100 (maps default to vacuum until set).
void Sim::init_pb(double w_, double kbar2_, int nonpolar_) {
    pb = 1; w = w_; kbar2 = kbar2_; nonpolar = nonpolar_;
    pbmg_dirty = 1;
    ex = dalloc<double>(G); ey = dalloc<double>(G); ez = dalloc<double>(G); k2 = dalloc<double>(G);
    launch_fill(stream, G, 1.0, ex);
    launch_fill(stream, G, 1.0, ey);
    launch_fill(stream, G, 1.0, ez);
    launch_fill(stream, G, 0.0, k2);
    restrict_pb_coef();
    sync();
}

void Sim::set_eps_k2(const double* ex_, const double* ey_, const double* ez_, const double* k2_) {
    pbmg_dirty = 1;
    if (!pb) fatal("maze: solver_set_eps_k2 needs solver_initialize_grid_pois_boltz first");
    MAZE_HIP(hipMemcpy(ex, ex_, G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(ey, ey_, G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(ez, ez_, G * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(k2, k2_, G * sizeof(double), hipMemcpyHostToDevice));
    restrict_pb_coef();
    sync();
}

// library/src/solver_md.c:51-65, particles.c:68-109, :181-228.
void Sim::init_particles(int n_, int n_typ, double L_, double h_, int np, int cas,
                         const int* types, const double* pos, const double* vel,
                         const double* mass, const double* charges, const double* pp) {
    if (cas < 0 || cas > 2) fatal("Invalid charge assignment scheme type");
    if (cas != 0 && n_ < 4) fatal("maze: spline charge assignment needs n >= 4");
    if (n_typ < 1 || n_typ > 16) fatal("maze: between 1 and 16 particle types are supported");
    if (np < 0) fatal("maze: negative particle count");
    free_particles();
    p.np = np; p.n_typ = n_typ; p.n = n_; p.L = L_; p.h = h_;
    p.cas = cas;
    p.nn = cas == 0 ? 8 : 64;
    p.r_cut = L_ / 2.0;
    p.types = dalloc<int>(np);
    p.pos = dalloc<double>(3L * np); p.vel = dalloc<double>(3L * np);
    p.mass = dalloc<double>(np); p.charges = dalloc<double>(np);
    p.fe = dalloc<double>(3L * np); p.fn = dalloc<double>(3L * np); p.ft = dalloc<double>(3L * np);
    p.base = dalloc<int>(3L * np);
    p.head = dalloc<unsigned long long>((long)n_ * n_ * n_);
    p.node = dalloc<int>((long)p.nn * np);
    p.link = dalloc<int>((long)p.nn * np);
    p.contrib = dalloc<double>((long)p.nn * np);
    spread_zero_old = 0;
    p.tpart = dalloc<double>(tf_part_doubles(np));
    p.eion = dalloc<double>(np);
    p.counts = dalloc<int>(np);
    // Tosi-Fumi shifted-force constants per type pair, with the reference's
    // expressions (library/src/particles.c:191-215), evaluated on the host.
    std::vector<double> table(7 * n_typ * n_typ);
    const double rc = p.r_cut;
    const double rc6 = std::pow(rc, 6), rc7 = rc6 * rc, rc8 = rc7 * rc, rc9 = rc8 * rc;
    for (int t = 0; t < n_typ * n_typ; t++) {
        const double A = pp[t * 5], B = pp[t * 5 + 1], C = pp[t * 5 + 2], D = pp[t * 5 + 3],
                     s = pp[t * 5 + 4];
        const double v_shift = A * std::exp(B * (s - rc)) - C / rc6 - D / rc8;
        const double alpha = A * B * std::exp(B * (s - rc)) - 6 * C / rc7 - 8 * D / rc9;
        const double beta = -v_shift - alpha * rc;
        double* o = &table[t * 7];
        o[0] = A; o[1] = B; o[2] = C; o[3] = D; o[4] = s; o[5] = alpha; o[6] = beta;
    }
    p.table = dalloc<double>(7L * n_typ * n_typ);
    MAZE_HIP(hipMemcpy(p.table, table.data(), table.size() * sizeof(double), hipMemcpyHostToDevice));
    if (np > 0) {
        MAZE_HIP(hipMemcpy(p.types, types, np * sizeof(int), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(p.pos, pos, 3L * np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(p.vel, vel, 3L * np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(p.mass, mass, np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(p.charges, charges, np * sizeof(double), hipMemcpyHostToDevice));
    }
    sync();
}

// library/src/particles.c:111-125: the ion radii (solvation + probe, Bohr)
// of the radius-based PB boundary, with the per-ion constants of the map and
// force kernels: (r + w)^2 and (r - w)^2 (as the reference evaluates them)
// and the half-widths of the two boxes.
void Sim::init_pb_particles(double gamma_np, double beta_np, const double* radii) {
    if (!pb) fatal("maze: solver_initialize_grid_pois_boltz must come before the PB radii");
    free_pb();
    const int np = p.np;
    pbd.np = np; pbd.n = n; pbd.h = h; pbd.w = w;
    pbd.eps_s = eps_s; pbd.eps_int = eps_int; pbd.kbar2 = kbar2;
    pbd.gamma_np = gamma_np; pbd.beta_np = beta_np; pbd.nonpolar = nonpolar;
    std::vector<double> rp2(np), rm2(np);
    std::vector<int> rmap(np), rfrc(np);
    int rmax = 0;
    for (int i = 0; i < np; i++) {
        const double r = radii[i];
        rp2[i] = std::pow(r + w, 2);
        rm2[i] = std::pow(r - w, 2);
        rmap[i] = (int)std::floor((r + w) / h) + 1;
        rfrc[i] = (int)std::floor((r + w) / h) + 2;
        if (!(r >= 0.0) || 2 * rfrc[i] + 1 > n)
            fatal("maze: PB radius + transition width must be >= 0 and span less than the grid");
        rmax = std::max(rmax, rmap[i]);
    }
    pbd.rmax = rmax;
    pbd.nb = (n + kPbTile - 1) / kPbTile;
    const unsigned long bins = (unsigned long)pbd.nb * pbd.nb * pbd.nb;
    pbd.key_bits = 1;
    while ((1ul << pbd.key_bits) < bins) pbd.key_bits++;
    pbd.radii = dalloc<double>(np); pbd.rp2 = dalloc<double>(np); pbd.rm2 = dalloc<double>(np);
    pbd.rmap = dalloc<int>(np); pbd.rfrc = dalloc<int>(np);
    pbd.keys = dalloc<unsigned int>(np); pbd.keys_sorted = dalloc<unsigned int>(np);
    pbd.ids = dalloc<int>(np); pbd.ids_sorted = dalloc<int>(np);
    pbd.fdb = dalloc<double>(3L * np); pbd.fib = dalloc<double>(3L * np);
    pbd.fnp = dalloc<double>(3L * np);
    pbd.sort_bytes = pb_sort_bytes(std::max(np, 1), pbd.key_bits);
    pbd.sort_tmp = dalloc<char>((long)std::max<size_t>(pbd.sort_bytes, 16));
    pbd.flag = &ps->pb_flag;
    if (np > 0) {
        MAZE_HIP(hipMemcpy(pbd.radii, radii, np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(pbd.rp2, rp2.data(), np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(pbd.rm2, rm2.data(), np * sizeof(double), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(pbd.rmap, rmap.data(), np * sizeof(int), hipMemcpyHostToDevice));
        MAZE_HIP(hipMemcpy(pbd.rfrc, rfrc.data(), np * sizeof(int), hipMemcpyHostToDevice));
    }
    pbd.active = 1;
    sync();
}

// library/src/grid_general.c:138-301 (the maps of the current positions);
// without radii the maps set by set_eps_k2 are kept.
void Sim::update_eps_k2() {
    if (pbd.active) pbmg_dirty = 1;
    if (!pb || !pbd.active) return;
    launch_pb_bin(stream, pbd, p.pos);
    timed(kTPbMaps, 0, [&] { launch_pb_maps(stream, pbd, p.pos, ex, ey, ez, k2); });
    restrict_pb_coef();
    MAZE_HIP(hipMemcpyAsync(&ps_host->pb_flag, &ps->pb_flag, sizeof(int), hipMemcpyDeviceToHost,
                            stream));
    pb_flag_pending = 1;
}

void Sim::check_pb_flag() {
    if (!pb_flag_pending) return;
    sync();
    pb_flag_pending = 0;
    if (ps_host->pb_flag)
        fatal("maze: more than 2048 ions within reach of one 8^3 tile of the PB maps");
}

// library/src/particles.c:420-775; returns the non-polar energy, which the
// reference leaves at 0.
double Sim::forces_pb() {
    if (!pb) fatal("Poisson-Boltzmann forces are not enabled in the grid.");
    if (!pbd.active)
        fatal("maze: PB boundary forces need the ion radii (solver_initialize_particles_pois_boltz)");
    check_pb_flag();
    timed(kTPbForces, 0, [&] { launch_pb_forces(stream, pbd, p.pos, phi, ex, ey, ez, k2); });
    return 0.0;
}

// library/src/solver_md.c:71-89, integrator_ovrvo.c:140-148.
void Sim::init_integrator(double dt_, double T_, double gamma, int type, int en) {
    itype = type; dt = dt_; T = T_; enabled = 0; c1 = 1.0; c2 = 1.0; calls = 0;
    if (en == 1) {
        enabled = 1;
        if (type == kOVRVO) {
            c1 = std::exp(-gamma * dt);
            c2 = std::sqrt(2 / (gamma * dt) * std::tanh(0.5 * gamma * dt));
        }
    }
}

// The solver state (and the last charge sum) on the host.  A one-wave kernel
// writes them to host-coherent memory with a sequence number last; the host
// polls the number instead of synchronising the stream, which saves the
// wake-up of a blocking wait on the critical path of every solve.  Should the
// number not arrive within a second, the stream is synchronised and the
// number checked once more.
void Sim::read_state(const CGState* which) {
    publish(which);
    wait_published();
}

void Sim::publish(const CGState* which) {
    want_seq = ++pub_seq;
    launch_publish_state(stream, which ? which : st, &ps->q_sum, pub_dev, want_seq);
}

bool Sim::published() const {
    return __atomic_load_n(&pub_host->seq, __ATOMIC_ACQUIRE) == want_seq;
}

void Sim::wait_published() {
    const unsigned int want = want_seq;
    const auto t0 = std::chrono::steady_clock::now();
    for (long spin = 0;; spin++) {
        if (__atomic_load_n(&pub_host->seq, __ATOMIC_ACQUIRE) == want) break;
        if ((spin & 4095) == 4095 &&
            std::chrono::steady_clock::now() - t0 > std::chrono::seconds(1)) {
            sync();
            if (__atomic_load_n(&pub_host->seq, __ATOMIC_ACQUIRE) != want)
                fatal("maze: the solver state was not published");
            break;
        }
        __builtin_ia32_pause();
    }
    std::memcpy(st_host, &pub_host->st, sizeof(CGState));
    ps_host->q_sum = pub_host->q_sum;
    pub_count++;
}

hipEvent_t Sim::next_event() {
    if (ev_next == ev_pool.size()) {
        hipEvent_t e;
        MAZE_HIP(hipEventCreate(&e));
        ev_pool.push_back(e);
    }
    return ev_pool[ev_next++];
}

// Accumulate the durations of the first `iters_done` iterations' kernels
// (later launches were no-ops after the stop and are not counted).
void Sim::prof_flush(int iters_done) {
    if (!pending.empty()) {  // timed launches may sit on either stream
        sync();
        MAZE_HIP(hipStreamSynchronize(side));
    }
    for (auto& pe : pending) {
        if (pe.iter >= iters_done) continue;
        float ms = 0.f;
        MAZE_HIP(hipEventElapsedTime(&ms, pe.a, pe.b));
        KernelTimer& t = timers[pe.timer];
        t.total_ms += ms;
        t.count += 1;
    }
    pending.clear();
    ev_next = 0;
}

// As prof_flush for a refinement loop: launches tagged with pass k count if
// that pass was applied and, for inner iterations, if the iteration ran
// (iters[k] iterations if converged, -iters[k] if stopped by its budget).
void Sim::prof_flush_passes(const int* iters, int npass) {
    if (!pending.empty()) {  // timed launches may sit on either stream
        sync();
        MAZE_HIP(hipStreamSynchronize(side));
    }
    for (auto& pe : pending) {
        if (pe.pass >= 0) {
            if (pe.pass >= npass) continue;
            const int done = iters[pe.pass] >= 0 ? iters[pe.pass] : -iters[pe.pass];
            if (pe.iter >= std::max(done, 1)) continue;
        }
        float ms = 0.f;
        MAZE_HIP(hipEventElapsedTime(&ms, pe.a, pe.b));
        KernelTimer& t = timers[pe.timer];
        t.total_ms += ms;
        t.count += 1;
    }
    pending.clear();
    ev_next = 0;
}

// Implicit-Jacobi CG, library/src/plcg.c:28-90: x holds the warm start.
// Iterations are launched in chunks sized from the previous solve and the
// device-side stop flag is polled between chunks.
int Sim::cg_solve(const GridDims& d, const PBCoef& pbc_, const double* b, double* x, double tol_,
                  int limit, double* rbuf, double* pa, double* pbuf, int* pred) {
    cg_issue(CgRun{d, pbc_, x, rbuf, pa, pbuf, limit, 0, std::max(4, *pred + 1), pred}, b, tol_);
    return cg_finish();
}

// The first chunk of a CG solve, up to the publication of its state.
void Sim::cg_issue(const CgRun& c, const double* b, double tol_) {
    cg = c;
    launch_cg_reset(stream, st, tol_, 0.0, cg.limit, 0);
    launch_cg_init<double>(stream, cg.d, cg.pbc, cg.x, b, cg.rbuf, st, red);
    cg_chunk();
    publish();
}

void Sim::cg_chunk() {
    for (int c = 0; c < cg.chunk && cg.launched < cg.limit; c++, cg.launched++) {
        const int it = cg.launched;
        timed(kTCgDir, it, [&] {
            launch_cg_dir<double>(stream, cg.d, cg.pbc, cg.pin, cg.rbuf, cg.pout, st, red, it == 0);
        });
        timed(kTCgUpdate, it,
              [&] { launch_cg_update<double>(stream, cg.d, cg.pbc, cg.pout, cg.x, cg.rbuf, st, red); });
        std::swap(cg.pin, cg.pout);
    }
}

// Waits for the published state; further chunks until the stop.
int Sim::cg_finish() {
    wait_published();
    while (!st_host->done && cg.launched < cg.limit) {
        cg.chunk = std::max(4, cg.chunk / 4);
        cg_chunk();
        read_state();
    }
    const int done_iters = st_host->res >= 0 ? st_host->res + 1 : cg.launched;
    prof_flush(done_iters);
    *cg.pred = done_iters;
    return st_host->res;
}

// Coarse CG with a relative stop, returning the iteration count; x must hold
// the initial guess.  Poisson: implicit-Jacobi recurrence (library/src/
// multigrid.c:13-96); PB (coarse coefficients cpb): plain CG (library/src/
// multigrid_pb.c:16-98).
int Sim::cg_count_solve(const GridDims& d, const PBCoef& cpb, const double* b, double* x,
                        double rtol, int maxit, double* rbuf, double* pa, double* pbuf, int* pred,
                        CGState* st) {
    const int plain = cpb.ex ? 1 : 0;
    launch_cg_reset(stream, st, 0.0, rtol, maxit, 1, 0, 0, plain);
    launch_cg_init<double>(stream, d, cpb, x, b, rbuf, st, red, nullptr, plain);
    int launched = 0;
    int chunk = std::max(4, *pred + 1);
    double* pin = pa;
    double* pout = pbuf;
    for (;;) {
        for (int c = 0; c < chunk && launched < maxit; c++, launched++) {
            launch_cg_dir<double>(stream, d, cpb, pin, rbuf, pout, st, red, launched == 0);
            launch_cg_update<double>(stream, d, cpb, pout, x, rbuf, st, red);
            std::swap(pin, pout);
        }
        read_state(st);
        if (st_host->done || launched >= maxit) break;
        chunk = std::max(4, chunk / 4);
    }
    *pred = std::max(1, st_host->res);
    return st_host->res;
}

// One iteration of the FP32-storage inner solve (Jacobi-CG, or CG with the
// multigrid preconditioner of Sim::vcycle); `it` is its 0-based index.
void Sim::inner_iter(bool use_mg, float* x, int it) {
    if (use_graphs && use_mg && !pb && !prof && x == d32) {
        const int v = it == 0 ? 0 : ((it & 1) ? 1 : 2);
        if (!inner_graph[v]) {
            hipGraph_t g = nullptr;
            MAZE_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            inner_body(use_mg, x, it);
            MAZE_HIP(hipStreamEndCapture(stream, &g));
            MAZE_HIP(hipGraphInstantiate(&inner_graph[v], g, nullptr, nullptr, 0));
            MAZE_HIP(hipGraphDestroy(g));
        }
        MAZE_HIP(hipGraphLaunch(inner_graph[v], stream));
        return;
    }
    inner_body(use_mg, x, it);
}

void Sim::drop_graphs() {
    for (hipGraphExec_t& g : inner_graph) {
        if (g) MAZE_HIP(hipGraphExecDestroy(g));
        g = nullptr;
    }
    for (hipGraphExec_t& g : pass_graph) {
        if (g) MAZE_HIP(hipGraphExecDestroy(g));
        g = nullptr;
    }
}

void Sim::inner_body(bool use_mg, float* x, int it) {
    float* pin = (it & 1) ? q32b : q32a;
    float* pout = (it & 1) ? q32a : q32b;
    if (use_mg && pb) {  // PB multigrid-PCG
        vcycle_pb(0);
        timed(kTCgDirF32, it, [&] {
            launch_pcg_dir(stream, dims, pin, pbmg[0].d, pout, st, red, it == 0, pbmg[0].cf);
        });
        timed(kTCgUpdateF32, it, [&] {
            launch_cg_update32(stream, dims, pbmg[0].cf, pout, x, r32, st, red, it == 0);
        });
    } else if (use_mg) {
        // Deferred finalisation: the finest up kernel, the direction pass and
        // the update write per-work-group partials that the next kernel folds.
        const int zcnt = mg_up_blocks(mg[0].n, mg[0].tile_up);
        if (zcnt > kMaxPartials) fatal("maze: finest multigrid tiles exceed the reduction buffers");
        vcycle(0, it);
        timed(kTCgDirF32, it, [&] {
            launch_pcg_dir_df(stream, dims, pin, mg[0].d, pout, st, dred, zcnt, it == 0, it & 1);
        });
        timed(kTCgUpdateF32, it, [&] {
            launch_cg_update_df(stream, dims, pout, x, r32, st, dred, it == 0, it & 1);
        });
    } else {
        timed(kTCgDirF32, it,
              [&] { launch_cg_dir<float>(stream, dims, pbc(), pin, r32, pout, st, red, it == 0); });
        timed(kTCgUpdateF32, it,
              [&] { launch_cg_update<float>(stream, dims, pbc(), pout, x, r32, st, red, it == 0); });
    }
}

// Start of an inner solve A x = b from x = 0 with the relative stop
// max|r| <= rtol max|b| (library/src/multigrid.c:74-78); the multigrid
// variant uses the explicit-preconditioner recurrence of
// library/src/plcg.c:104-166.  gated: skipped once the outer loop is done.
void Sim::inner_start(bool use_mg, const float* b, float* x, double rtol, int limit, int gated) {
    launch_cg_reset(stream, st, 0.0, rtol, limit, 1, use_mg ? 1 : 0, gated);
    launch_cg_init<float>(stream, dims, use_mg && !pb ? PBCoef{} : pbc(), nullptr, b, r32, st, red,
                          x);
    if (use_mg && !pb) mg[0].b = r32;
    if (use_mg && pb) {
        pbmg_coefs();
        pbmg[0].b = r32;
    }
}

// Complete inner solve, polling the device after every iteration past the
// first `first_chunk`; returns the iterations performed.
int Sim::inner_solve(bool use_mg, const float* b, float* x, double rtol, int limit,
                     int first_chunk) {
    inner_start(use_mg, b, x, rtol, limit, 0);
    int launched = 0;
    for (; launched < std::min(first_chunk, limit); launched++) inner_iter(use_mg, x, launched);
    inner_read_state(use_mg);
    while (!st_host->done && launched < limit) {
        inner_iter(use_mg, x, launched++);
        inner_read_state(use_mg);
    }
    const int it = st_host->res;
    prof_flush(it);
    return it;
}

void Sim::inner_read_state(bool use_mg) {
    if (inner_deferred(use_mg)) launch_df_finalize(stream, st, dred, dims.blocks);
    read_state();
}

void Sim::free_mg() {
    drop_graphs();
    for (size_t l = 0; l < mg.size(); l++) {
        if (l > 0) dfree(mg[l].b);
        dfree(mg[l].d);
        dfree(mg[l].t);
    }
    mg.clear();
    for (size_t l = 0; l < mg64.size(); l++) {
        if (l > 0) dfree(mg64[l].b);
        dfree(mg64[l].d);
        dfree(mg64[l].t);
    }
    mg64.clear();
    mg64_ok = -1;
    dfree(mg_Qf);
    dfree(mg_ilam);
    dfree(mg_w0);
    dfree(mg_w1);
    dfree(mg_coarse_count);
    mg_ok = -1;
}

// Levels n, n/2, ...: halved once, then while the size is even and above
// `stop`; the last level is solved exactly (in one work-group up to 16 points
// per side, by axis passes over global work arrays up to 128).  Unavailable
// when halving stops at an odd size above 128.  stop = 16 (default): the
// one-work-group exact solve replaces the two smallest levels; stop = 12 is
// the earlier hierarchy, whose levels of at most 16^3 run in the one-
// work-group tail (MAZE_MG_COARSE=12).
int Sim::default_mg_stop() {
    static const int stop = [] {
        const char* e = std::getenv("MAZE_MG_COARSE");
        return e && std::atoi(e) == 12 ? 12 : kMgCoarseLds;
    }();
    return stop;
}

std::vector<int> Sim::mg_sizes(int n, int stop) {
    std::vector<int> sizes{n};
    while (sizes.back() % 2 == 0 && sizes.back() >= 8 && (sizes.size() == 1 || sizes.back() > stop))
        sizes.push_back(sizes.back() / 2);
    if (sizes.back() > kMgCoarseMax || sizes.size() < 2) sizes.clear();
    return sizes;
}

bool Sim::setup_mg() {
    if (mg_ok >= 0) return mg_ok == 1;
    const std::vector<int> sizes = mg_sizes(n, mg_stop);
    if (sizes.empty()) {
        mg_ok = 0;
        return false;
    }
    for (size_t l = 0; l < sizes.size(); l++) {
        MGLevel L;
        L.n = sizes[l];
        L.dims = make_dims(L.n, choose_slab(L.n));
        const long g = (long)L.n * L.n * L.n;
        L.b = l == 0 ? nullptr : dalloc<float>(g);
        L.tile = mg_default_tile(L.n);
        L.tile_up = mg_default_tile_up(L.n);
        if (const char* e = std::getenv("MAZE_MG_TILES")) {  // "t0,t1,...": tuning override
            const char* q = e;
            for (size_t k = 0; k < l && q; k++) {
                q = std::strchr(q, ',');
                if (q) q++;
            }
            if (q && *q >= '0' && *q < '0' + kMgTiles) L.tile = L.tile_up = *q - '0';
        }
        L.d = dalloc<float>(g);
        L.t = dalloc<float>(g);
        mg.push_back(L);
    }
    // Real orthonormal eigenbasis of the periodic 1-D second difference on c
    // points: constant, cos/sin pairs, and the alternating mode for even c;
    // eigenvalue 2 cos(2 pi k / c) - 2 per row.
    const int c = sizes.back();
    mg_c = c;
    std::vector<double> Q(c * c), mu(c);
    int row = 0;
    for (int m = 0; m < c; m++) Q[row * c + m] = 1.0 / std::sqrt((double)c);
    mu[row++] = 0.0;
    for (int k = 1; 2 * k < c; k++) {
        const double lam = 2 * std::cos(2 * kPi * k / c) - 2;
        for (int m = 0; m < c; m++) Q[row * c + m] = std::sqrt(2.0 / c) * std::cos(2 * kPi * k * m / c);
        mu[row++] = lam;
        for (int m = 0; m < c; m++) Q[row * c + m] = std::sqrt(2.0 / c) * std::sin(2 * kPi * k * m / c);
        mu[row++] = lam;
    }
    if (c % 2 == 0) {
        for (int m = 0; m < c; m++) Q[row * c + m] = (m % 2 ? -1.0 : 1.0) / std::sqrt((double)c);
        mu[row++] = -4.0;
    }
    std::vector<double> il(c * c * c);
    for (int a = 0; a < c; a++)
        for (int b = 0; b < c; b++)
            for (int d = 0; d < c; d++) {
                const double l3 = mu[a] + mu[b] + mu[d];
                il[(a * c + b) * c + d] = (a == 0 && b == 0 && d == 0) ? 0.0 : 1.0 / l3;
            }
    mg_Qf = dalloc<double>(c * c);
    mg_ilam = dalloc<double>(c * c * c);
    mg_coarse_count = dalloc<unsigned int>(1);  // zeroed
    if (c > kMgCoarseLds) {
        mg_w0 = dalloc<double>(c * c * c);
        mg_w1 = dalloc<double>(c * c * c);
    }
    MAZE_HIP(hipMemcpy(mg_Qf, Q.data(), Q.size() * sizeof(double), hipMemcpyHostToDevice));
    MAZE_HIP(hipMemcpy(mg_ilam, il.data(), il.size() * sizeof(double), hipMemcpyHostToDevice));
    // The one-work-group tail takes over from the first level below the
    // finest whose remaining levels fit its LDS (<= 16^3); without such a
    // level the exact coarsest solve runs as its own kernel.
    mg_tail = (int)mg.size();
    for (int l = 1; l < (int)mg.size() && mg_stop <= 12; l++) {
        int ns[8];
        const int nl = (int)mg.size() - l;
        if (nl > 8) continue;
        for (int k = 0; k < nl; k++) ns[k] = mg[l + k].n;
        if (mg_tail_fits(nl, ns)) {
            mg_tail = l;
            break;
        }
    }
    mg_ok = 1;
    if (mg_autotune) tune_mg_tiles();
    return true;
}

namespace {
std::mutex g_tune_mu;
std::map<int, std::pair<int, int>> g_tuned;  // level size -> (down tile, up tile)
}  // namespace

// Times every tile shape that can run a level (five launches each, on the
// level's own buffers with a scratch solver state) and keeps the fastest.
void Sim::tune_mg_tiles() {
    CGState* tst = dalloc<CGState>(1);  // zeroed: not done
    hipEvent_t a, b;
    MAZE_HIP(hipEventCreate(&a));
    MAZE_HIP(hipEventCreate(&b));
    const double c1 = mg_omega / 6.0, c2 = mg_omega2 / 6.0;
    for (int l = 0; l + 1 < (int)mg.size() && l < mg_tail; l++) {
        MGLevel& L = mg[l];
        {
            std::lock_guard<std::mutex> g(g_tune_mu);
            auto it = g_tuned.find(L.n);
            if (it != g_tuned.end()) {
                L.tile = it->second.first;
                if (l > 0) L.tile_up = it->second.second;
                continue;
            }
        }
        auto time_it = [&](auto&& launch) {
            launch();  // warm-up
            MAZE_HIP(hipEventRecord(a, stream));
            for (int r = 0; r < 5; r++) launch();
            MAZE_HIP(hipEventRecord(b, stream));
            MAZE_HIP(hipEventSynchronize(b));
            float ms = 0.0f;
            MAZE_HIP(hipEventElapsedTime(&ms, a, b));
            return ms;
        };
        float best_d = 1e30f, best_u = 1e30f;
        int td = L.tile, tu = L.tile_up;
        for (int t = 0; t < kMgTiles; t++) {
            if (!mg_tile_ok(t, L.n)) continue;
            const float md = time_it(
                [&] { launch_mg_down(stream, L.n, t, L.t, L.d, mg[l + 1].b, c1, c2, tst); });
            if (md < best_d) best_d = md, td = t;
            if (l == 0) continue;  // the finest up kernel reduces <r, z>: default shape
            const float mu = time_it([&] {
                launch_mg_up(stream, L.n, t, L.t, L.d, mg[l + 1].d, L.b, c1, c2, tst, 0, red);
            });
            if (mu < best_u) best_u = mu, tu = t;
        }
        L.tile = td;
        if (l > 0) L.tile_up = tu;
        std::lock_guard<std::mutex> g(g_tune_mu);
        g_tuned[L.n] = {td, tu};
    }
    MAZE_HIP(hipEventDestroy(a));
    MAZE_HIP(hipEventDestroy(b));
    sync();
    dfree(tst);
}

// The PB multigrid needs a hierarchy whose coarsest level fits one work-group.
bool Sim::setup_pbmg() {
    if (pbmg_ok >= 0) return pbmg_ok == 1;
    const std::vector<int> sizes = mg_sizes(n, mg_stop);
    if (!pb || sizes.empty() || sizes.back() > kMgCoarseLds) {
        pbmg_ok = 0;
        return false;
    }
    for (size_t l = 0; l < sizes.size(); l++) {
        PbLevel L;
        L.n = sizes[l];
        L.dims = make_dims(L.n, choose_slab(L.n));
        const long g = (long)L.n * L.n * L.n;
        L.ex = dalloc<float>(g); L.ey = dalloc<float>(g);
        L.ez = dalloc<float>(g); L.k2 = dalloc<float>(g);
        L.cf = PBCoef32{L.ex, L.ey, L.ez, L.k2};
        if (l > 0) L.b = dalloc<float>(g);
        L.g = dalloc<float>(g); L.d = dalloc<float>(g); L.t = dalloc<float>(g);
        pbmg.push_back(L);
    }
    pbmg_dirty = 1;
    pbmg_ok = 1;
    return true;
}

void Sim::free_pbmg() {
    for (PbLevel& L : pbmg) {
        for (float** a : {&L.ex, &L.ey, &L.ez, &L.k2, &L.g, &L.d, &L.t}) dfree(*a);
        if (&L != &pbmg[0]) dfree(L.b);
    }
    pbmg.clear();
    pbmg_ok = -1;
    pbmg_dirty = 1;
}

// FP32 coefficients of every level from the current maps (after each map change).
void Sim::pbmg_coefs() {
    if (!pbmg_dirty) return;
    for (size_t l = 0; l < pbmg.size(); l++) {
        PbLevel& L = pbmg[l];
        if (l == 0) launch_pb_to32(stream, G, pbc(), L.ex, L.ey, L.ez, L.k2);
        else launch_pb_coarsen(stream, pbmg[l - 1].n, pbmg[l - 1].cf, L.ex, L.ey, L.ez, L.k2);
        launch_pb_idg(stream, L.n, L.cf, L.g);
    }
    pbmg_dirty = 0;
}

// V(2,2) for the PB operator: two damped-Jacobi sweeps from zero, residual,
// restriction, the coarser cycle, prolongation-correction, two sweeps (the
// last one on the finest level with <r, z> and the PCG beta).
void Sim::vcycle_pb(int l) {
    PbLevel& L = pbmg[l];
    if (l == (int)pbmg.size() - 1) {
        launch_pb_coarse(stream, L.n, L.cf, L.g, L.b, L.d, pb_coarse_iters, pb_coarse_lo, 2.0, st);
        return;
    }
    PbLevel& C = pbmg[l + 1];
    if (pb_tail && l > 0 && l + 2 == (int)pbmg.size() && L.n <= kPbTailMaxN) {
        launch_pb_tail(stream, L.n, L.cf, L.g, L.b, L.d, C.cf, C.g, mg_omega, mg_omega2,
                       pb_coarse_iters, pb_coarse_lo, 2.0, st);
        return;
    }
    launch_pb_pre(stream, L.dims, L.cf, L.g, L.b, L.d, mg_omega, mg_omega2, st);
    launch_pb_resid(stream, L.dims, L.cf, L.b, L.d, L.t, st);
    launch_mg_restrict(stream, L.n, L.t, C.b, st);
    vcycle_pb(l + 1);
    launch_mg_prolong_add(stream, C.n, C.d, L.d, st);
    launch_pb_smooth(stream, L.dims, L.cf, L.g, L.b, L.d, L.t, mg_omega2, st, 0, red);
    launch_pb_smooth(stream, L.dims, L.cf, L.g, L.b, L.t, L.d, mg_omega, st, l == 0 ? 1 : 0, red);
}

// V(2,2): pre-smooth, residual, restrict, recurse, prolong-correct, post-smooth,
// as one fused down and one fused up kernel per level above the tail.
void Sim::vcycle(int l, int iter) {
    MGLevel& L = mg[l];
    const double c1 = mg_omega / 6.0, c2 = mg_omega2 / 6.0;
    if (l >= mg_tail) {
        int ns[8];
        const int nl = (int)mg.size() - l;
        for (int k = 0; k < nl; k++) ns[k] = mg[l + k].n;
        launch_mg_tail(stream, nl, ns, L.b, L.d, c1, c2, mg_Qf, mg_ilam, st);
        return;
    }
    if (l == (int)mg.size() - 1) {
        if (mg_c <= kMgCoarseLds) launch_mg_coarse(stream, mg_c, L.b, L.d, mg_Qf, mg_ilam, st);
        else launch_mg_coarse_big(stream, mg_c, L.b, L.d, mg_w0, mg_w1, mg_Qf, mg_ilam, st);
        return;
    }
    const bool top = l == 0;
    auto t = [&](int timer, auto&& f) {
        if (top) timed(timer, iter, f);
        else f();
    };
    // Above a 16^3 coarsest level the down kernel's last work-group solves it.
    const bool fused = mg_fuse_coarse && l + 2 == (int)mg.size() && l + 1 < mg_tail &&
                       mg_down_coarse_ok(L.n, L.tile, mg_c);
    if (fused)
        t(kTMgDown, [&] {
            launch_mg_down_coarse(stream, L.n, L.tile, L.b, L.t, mg[l + 1].b, c1, c2, st, mg_c,
                                  CoarseTail{mg_coarse_count, mg_Qf, mg_ilam, mg[l + 1].d});
        });
    else {
        // The finest up kernel writes the <r, z> partials the direction pass folds.
        t(kTMgDown, [&] { launch_mg_down(stream, L.n, L.tile, L.b, L.t, mg[l + 1].b, c1, c2, st); });
        vcycle(l + 1, iter);
    }
    t(kTMgUp, [&] {
        launch_mg_up(stream, L.n, L.tile_up, L.b, L.t, mg[l + 1].d, L.d, c1, c2, st, top,
                     Reduce{dred.pz, nullptr, nullptr});
    });
}

// MaZe step in FP64, library/src/verlet.c:31-71.
int Sim::maze_fp64() {
    fp64_issue();
    return fp64_finish();
}

void Sim::fp64_issue() {
    const double coef = pb ? (4 * kPi) / h : (4 * kPi) / (h * eps_s);
    timed(kTPred, 0,
          [&] { launch_pred_sigma(stream, dims, pbc(), phi, phip, q, coef, tmp, sigma); });
    double* old_phip = phip;
    phip = phi;
    phi = tmp;
    tmp = old_phip;
    cg_issue(CgRun{dims, pbc(), y, r, p0, p1, n * n, 0, std::max(4, last_iters + 1), &last_iters},
             sigma, tol);
}

int Sim::fp64_finish() {
    const int res = cg_finish();
    launch_axpy(stream, G, -1.0, y, phi);
    return res;
}

// ---- FP64 multigrid-PCG (precision kFP64MG) ----------------------------------
// The FP32 hierarchy's level sizes with FP64 arrays.  Unavailable for PB, for
// grids without the FP32 multigrid, and for the hierarchies whose smallest
// levels run in the one-work-group tail or whose coarsest level is above 16^3
// (those solves stay FP32); the MaZe step then runs the reference's FP64 CG.
bool Sim::setup_mg64() {
    if (mg64_ok >= 0) return mg64_ok == 1;
    mg64_ok = 0;
    if (pb || !setup_mg() || mg_tail < (int)mg.size() || mg_c > kMgCoarseLds) return false;
    for (size_t l = 0; l < mg.size(); l++) {
        MGLevel64 L;
        L.n = mg[l].n;
        const long g = (long)L.n * L.n * L.n;
        L.b = l == 0 ? nullptr : dalloc<double>(g);
        L.d = dalloc<double>(g);
        L.t = dalloc<double>(g);
        mg64.push_back(L);
    }
    mg64_ok = 1;
    return true;
}

bool Sim::use_fp64mg() {
    return precision == kFP64MG && grid_type == kMazeLCG && setup_mg64();
}

// vcycle() with FP64 operands: V(2,2) with the same sweep weights, the
// 27-point restriction, trilinear prolongation and the exact coarsest solve;
// the finest up kernel writes the <r, z> partials the direction pass folds.
void Sim::vcycle64(int l) {
    MGLevel64& L = mg64[l];
    const double c1 = mg_omega / 6.0, c2 = mg_omega2 / 6.0;
    if (l == (int)mg64.size() - 1) {
        launch_mg_coarse(stream, mg_c, L.b, L.d, mg_Qf, mg_ilam, st);
        return;
    }
    launch_mg_down(stream, L.n, L.b, L.t, mg64[l + 1].b, c1, c2, st);
    vcycle64(l + 1);
    launch_mg_up(stream, L.n, L.b, L.t, mg64[l + 1].d, L.d, c1, c2, st, l == 0 ? 1 : 0,
                 Reduce{dred.pz, nullptr, nullptr});
}

// MaZe step with the FP64 multigrid-PCG: the prediction of library/src/
// verlet.c:31-71, then the explicit-preconditioner CG of library/src/
// plcg.c:104-166 (r = A y - sigma, p = -z, x += alpha p, r += alpha A p) with
// P^-1 = vcycle64, from the previous solution y to max|r| <= tol, with the
// deferred finalisation of the FP32 inner iteration.  The previous step's
// iteration count is queued at once; further iterations are polled.
void Sim::fp64mg_issue() {
    const double coef = (4 * kPi) / (h * eps_s);
    timed(kTPred, 0,
          [&] { launch_pred_sigma(stream, dims, pbc(), phi, phip, q, coef, tmp, sigma); });
    double* old_phip = phip;
    phip = phi;
    phi = tmp;
    tmp = old_phip;
    mg64[0].b = r;
    launch_cg_reset(stream, st, tol, 0.0, n * n, 0, 1);
    launch_cg_init<double>(stream, dims, pbc(), y, sigma, r, st, red);
    mg64_launched = 0;
    fp64mg_chunk(mg64_pred);
    launch_df_finalize(stream, st, dred, dims.blocks);
    publish();
}

void Sim::fp64mg_chunk(int count) {
    const int zcnt = mg_up_blocks64(n);
    if (zcnt > kMaxPartials) fatal("maze: finest multigrid tiles exceed the reduction buffers");
    for (int c = 0; c < count && mg64_launched < n * n; c++, mg64_launched++) {
        const int it = mg64_launched;
        double* pin = (it & 1) ? p1 : p0;
        double* pout = (it & 1) ? p0 : p1;
        vcycle64(0);
        launch_pcg_dir_df(stream, dims, pin, mg64[0].d, pout, st, dred, zcnt, it == 0, it & 1);
        launch_cg_update_df(stream, dims, pout, y, r, st, dred, 0, it & 1);
    }
}

int Sim::fp64mg_finish() {
    wait_published();
    while (!st_host->done && mg64_launched < n * n) {
        fp64mg_chunk(1);
        launch_df_finalize(stream, st, dred, dims.blocks);
        read_state();
    }
    const int res = st_host->res;
    mg64_pred = std::max(1, res >= 0 ? res + 1 : mg64_pred);
    launch_axpy(stream, G, -1.0, y, phi);
    return res;
}

// MaZe step with FP32-storage inner solves inside FP64 iterative refinement:
// each refinement pass recomputes the true residual sigma - A y in FP64, and
// the loop stops on the reference's criterion max|A y - sigma| <= tol; each
// inner solve reduces its residual by inner_rtol.  The passes of the previous
// step (with their inner iteration counts) are queued without waiting: the
// device advances the loop and turns passes after convergence into no-ops,
// so a step normally costs one host round trip.  Further passes, if needed,
// are driven from the host.  Returns the total number of inner iterations, or
// -1 if the outer loop did not converge in kMaxOuter passes.
//
// y_extrap: the loop starts from 2 y_n-1 - y_n-2 instead of y_n-1 (the
// converged y satisfies the same stop test; only the start differs).  The
// three y buffers then rotate: y (y_n-1), y_hist (y_n-2) and y_alt (spare).
int Sim::maze_mixed() {
    spec_post = 0;  // nothing was queued behind this solve
    mixed_issue();
    return mixed_finish();
}

// Everything up to the publication of the state after the previous step's
// schedule of passes.
void Sim::mixed_issue() {
    if (!d32) {
        d32 = dalloc<float>(G); r32 = dalloc<float>(G);
        q32a = dalloc<float>(G); q32b = dalloc<float>(G); y_alt = dalloc<double>(G);
    }
    if (y_extrap && !y_hist) y_hist = dalloc<double>(G);
    MixedRun& m = mx;
    m.use_mg = precision == kMixedMG && (pb ? setup_pbmg() : setup_mg());
    if (m.use_mg && !pb) mg[0].b = r32;
    if (m.use_mg && pb) {
        pbmg_coefs();
        pbmg[0].b = r32;
    }
    const double coef = pb ? (4 * kPi) / h : (4 * kPi) / (h * eps_s);
    timed(kTPred, 0,
          [&] { launch_pred_sigma(stream, dims, pbc(), phi, phip, q, coef, tmp, sigma); });
    double* old_phip = phip;
    phip = phi;
    phi = tmp;
    tmp = old_phip;
    m.inner_limit = inner_limit_for(m.use_mg);
    m.ra = RefineArgs{tol, inner_rtol_for(m.use_mg), m.inner_limit, m.use_mg ? 1 : 0};
    // The two buffers the refinement passes alternate between (st->ysel).
    m.extrap = y_extrap && hist_ok;
    m.ya = m.extrap ? y_alt : y;
    m.yb = m.extrap ? y_hist : y_alt;
    if (y_extrap && !hist_ok) launch_copy(stream, G, y, y_hist);
    if (m.extrap)
        timed(kTRefine, 0, [&] {
            launch_refine_extrap(stream, dims, pbc(), y, y_hist, y_alt, sigma, r32, m.ra, st, red);
        });
    else
        mixed_refine(0);
    for (m.queued = 0; m.queued < (int)sched.size(); m.queued++) {
        const int k = sched[m.queued];
        mixed_pass(m.queued, k >= 0 ? std::max(k, 1) : 1 - k, false);
    }
    publish();
}

// Each refinement pass also starts the next pass's inner solve.
void Sim::mixed_refine(int add) {
    timed(kTRefine, 0, [&] {
        launch_refine(stream, dims, pbc(), mx.ya, mx.yb, d32, sigma, r32, mx.ra, st, red, add,
                      inner_deferred(mx.use_mg) ? &dred : nullptr);
    });
}

// One refinement pass of `m` inner iterations; poll: iterate until the
// inner stop (reading the state after each iteration).
void Sim::mixed_pass(int k, int m, bool poll) {
    // A queued pass of up to kPassGraphs iterations replays one captured graph
    // (its iterations and the refinement), one per length and y buffer order.
    const bool md_buffers = (mx.ya == y || mx.ya == y_alt) && (mx.yb == y || mx.yb == y_alt);
    if (!poll && use_graphs && mx.use_mg && !pb && !prof && !mx.extrap && md_buffers && m >= 1 &&
        m <= kPassGraphs && m <= mx.inner_limit) {
        const int v = 2 * (m - 1) + (mx.ya < mx.yb ? 0 : 1);  // y and y_alt swap roles
        const RefineArgs& ga = pass_graph_ra[v];
        if (pass_graph[v] && (ga.otol != mx.ra.otol || ga.rtol != mx.ra.rtol ||
                              ga.limit != mx.ra.limit || ga.pcg != mx.ra.pcg)) {
            MAZE_HIP(hipGraphExecDestroy(pass_graph[v]));
            pass_graph[v] = nullptr;
        }
        if (!pass_graph[v]) {
            pass_graph_ra[v] = mx.ra;
            hipGraph_t g = nullptr;
            MAZE_HIP(hipStreamBeginCapture(stream, hipStreamCaptureModeRelaxed));
            for (int it = 0; it < m; it++) inner_body(mx.use_mg, d32, it);
            mixed_refine(1);
            MAZE_HIP(hipStreamEndCapture(stream, &g));
            MAZE_HIP(hipGraphInstantiate(&pass_graph[v], g, nullptr, nullptr, 0));
            MAZE_HIP(hipGraphDestroy(g));
        }
        MAZE_HIP(hipGraphLaunch(pass_graph[v], stream));
        return;
    }
    cur_pass = k;
    int launched = 0;
    for (; launched < std::min(m, mx.inner_limit); launched++) inner_iter(mx.use_mg, d32, launched);
    if (poll) {
        inner_read_state(mx.use_mg);
        while (!st_host->done && launched < mx.inner_limit) {
            inner_iter(mx.use_mg, d32, launched++);
            inner_read_state(mx.use_mg);
        }
    }
    mixed_refine(1);
    cur_pass = -1;
}

int Sim::mixed_finish() {
    MixedRun& m = mx;
    wait_published();
    spec_hit = spec_post && st_host->outer_done;
    while (!st_host->outer_done && st_host->passes < kMaxOuter) {
        mixed_pass(m.queued++, 1, true);
        read_state();
    }
    const int passes = std::min(st_host->passes, kMaxPasses);
    sched.assign(st_host->pass_iters, st_host->pass_iters + passes);
    prof_flush_passes(st_host->pass_iters, passes);
    double* fin = st_host->ysel ? m.yb : m.ya;
    if (!m.extrap) {
        if (fin != y) std::swap(y, y_alt);
        if (y_extrap) hist_ok = 1;  // y_hist holds y_n-1 (copied above)
    } else {
        // y_n-1 becomes the history, the result the current y.
        double* other = fin == m.ya ? m.yb : m.ya;
        y_hist = y;
        y = fin;
        y_alt = other;
    }
    if (!spec_hit) launch_axpy(stream, G, -1.0, y, phi);
    const int res = st_host->outer_done ? st_host->inner_total : -1;
    last_res = res;
    return res;
}

// library/src/grid_lcg.c:126-149.
int Sim::lcg_update() {
    double c = -4 * kPi / h;
    if (!pb) c /= eps_s;
    launch_copy(stream, G, phi, phip);
    launch_scale_copy(stream, G, c, q, y);
    return cg_solve(dims, pbc(), y, phi, tol, n * n, r, p0, p1, &last_iters);
}

// Reference two-grid V-cycle, library/src/multigrid.c:99-168 (3 RBGS sweeps
// for the multigrid solvers, 5 as the MAZE-LCG preconditioner,
// library/src/precond_mg.c:3-5), or its Poisson-Boltzmann form,
// library/src/multigrid_pb.c:275-355 (4 RBGS-PB sweeps, coarse coefficients
// restricted once per coefficient update).  sm <= 0: the solver's default.
void Sim::mg_vcycle(const double* in, double* out, int sm) {
    if (sm <= 0) sm = pb ? 4 : 3;
    if (pb)
        launch_rbgs_pb(stream, dims, pbc(), in, out, sm);
    else
        launch_rbgs(stream, dims, in, out, sm);
    launch_mg_residual(stream, dims, pbc(), in, out, mg_r);
    launch_restrict27(stream, n, mg_r, c_rhs);
    launch_fill(stream, (long)nc * nc * nc, 0.0, c_eps);
    const PBCoef cpb = pb ? PBCoef{cex, cey, cez, ck2} : PBCoef{};
    cg_count_solve(cdims, cpb, c_rhs, c_eps, 1e-5, 50, c_r, c_p0, c_p1, &last_coarse, st_c);
    launch_prolong_add(stream, nc, c_eps, out);
    if (pb)
        launch_rbgs_pb(stream, dims, pbc(), in, out, sm);
    else
        launch_rbgs(stream, dims, in, out, sm);
}

// MaZe step with the reference's multigrid preconditioner,
// library/src/verlet.c:31-71 with conj_grad_precond (grid_maze_lcg.c:130-145).
int Sim::maze_pcg_mg() {
    const double coef = (4 * kPi) / (h * eps_s);
    launch_pred_sigma(stream, dims, PBCoef{}, phi, phip, q, coef, tmp, sigma);
    double* old_phip = phip;
    phip = phi;
    phi = tmp;
    tmp = old_phip;
    const int res = pcg_mg_solve(sigma, y, tol, n * n);
    launch_axpy(stream, G, -1.0, y, phi);
    return res;
}

// v = P^-1 r of the explicit-preconditioner CG: one reference V-cycle of 5
// RBGS sweeps per side from the previous v (library/src/precond_mg.c:3-5), or
// the block-Jacobi inverse on every 10^3 block
// (library/src/precond_blockjacobi.c:170-206), which overwrites v.
void Sim::precond_apply(const double* rr, double* v) {
    if (precond == kPBlockJacobi)
        launch_blockjacobi(stream, n, rr, v, bj_tab);
    else
        mg_vcycle(rr, v, 5);
}

// Explicit-preconditioner CG, library/src/plcg.c:104-166, with P^-1 = one
// reference V-cycle of 5 RBGS sweeps per side started from the previous v
// (library/src/precond_mg.c:3-5; multigrid.c:170-173 does not clear `out`).
// The first V-cycle of a solve starts from v = 0 (the reference's v is fresh
// malloc'd memory there, library/src/plcg.c:120).  Every iteration waits for
// the state: the V-cycle's coarse CG is itself polled from the host.
// Returns the 0-based iteration index at convergence, or -1 after `limit`.
int Sim::pcg_mg_solve(const double* b, double* x, double tol_, int limit) {
    launch_cg_reset(stream, st, tol_, 0.0, limit, 0, /*pcg=*/1, 0, /*plain=*/1);
    launch_cg_init<double>(stream, dims, PBCoef{}, x, b, r, st, red, nullptr, 1);  // r = A x - b
    launch_fill(stream, G, 0.0, pv);
    read_state();
    if (st_host->done) return st_host->res;  // zero residual: x unchanged
    precond_apply(r, pv);
    launch_dot_rv(stream, dims, r, pv, st, red, 1);
    double* pin = p0;
    double* pout = p1;
    for (int it = 0; it < limit; it++) {
        // p = -v (first) or beta p - v, then <p, A p> and alpha; the plain
        // recurrence's coefficient -1 multiplies v in place of r.
        launch_cg_dir<double>(stream, dims, PBCoef{}, pin, pv, pout, st, red, it == 0);
        launch_cg_update<double>(stream, dims, PBCoef{}, pout, x, r, st, red);
        std::swap(pin, pout);
        read_state();
        if (st_host->done) break;
        precond_apply(r, pv);
        launch_dot_rv(stream, dims, r, pv, st, red, 0);
    }
    return st_host->res;
}

// Coarse PB coefficients of the reference two-grid cycle.
void Sim::restrict_pb_coef() {
    if (!pb || !(grid_type == kMG || grid_type == kMazeMG)) return;
    const long Gc = (long)nc * nc * nc;
    if (!cex) {
        cex = dalloc<double>(Gc); cey = dalloc<double>(Gc);
        cez = dalloc<double>(Gc); ck2 = dalloc<double>(Gc);
    }
    launch_restrict_pb_coef(stream, n, pbc(), cex, cey, cez, ck2);
}

// library/src/grid_multigrid.c:110-212 (maze == false) and
// library/src/verlet.c:86-149 (maze == true); with PB, grid_multigrid.c:149-175
// and verlet.c:211-271.
int Sim::mg_update(bool maze) {
    double* x;
    if (maze) {
        const double coef = pb ? (4 * kPi) / h : (4 * kPi) / (h * eps_s);
        launch_pred_sigma(stream, dims, pbc(), phi, phip, q, coef, tmp, sigma);
        double* old_phip = phip;
        phip = phi;
        phi = tmp;
        tmp = old_phip;
        x = y;
    } else {
        launch_pred_sigma(stream, dims, pbc(), phi, phip, q, 0.0, tmp, sigma);
        double* old_phip = phip;
        phip = phi;
        phi = tmp;
        tmp = old_phip;
        double c = -4 * kPi / h;
        if (!pb) c /= eps_s;
        launch_scale_copy(stream, G, c, q, sigma);
        x = phi;
    }
    int res = -1, k = 0;
    while (k < 1000) {
        mg_vcycle(sigma, x);
        launch_true_residual_max(stream, dims, pbc(), x, sigma, st, red);
        read_state();
        k++;
        if (st_host->resid <= tol) {
            res = k;
            break;
        }
    }
    if (maze) launch_axpy(stream, G, -1.0, y, phi);
    if (res == -1 && maze) std::fprintf(stderr, "Warning: Multigrid did not converge after 1000 iterations.\n");
    return res;
}

// library/src/solver_md.c:91-107.  defer: the charge-conservation check
// is made by check_charge() after the next read_state() instead of here.
int Sim::update_charges(bool defer, bool base_done) {
    if (!base_done) launch_cic_base(stream, p);
    // q holds the previous spread, which the spread clears node by node; after
    // a new grid or particle set, clear it whole once.
    if (p.np > 0 && p.n != n) fatal("maze: the particles' grid size differs from the field grid");
    if (!spread_zero_old) launch_fill(stream, G, 0.0, q);
    launch_assign(stream, p, ps, q, spread_zero_old);
    spread_zero_old = p.np > 0 ? 1 : 0;
    launch_contrib_sum(stream, p, ps, red);
    if (defer) return 0;  // read_state() brings q_sum with the solver state
    MAZE_HIP(hipMemcpyAsync(&ps_host->q_sum, &ps->q_sum, sizeof(double), hipMemcpyDeviceToHost,
                            stream));
    sync();
    return check_charge();
}

int Sim::check_charge() {
    const double loc = ps_host->q_sum;
    if (std::fabs(q_tot - loc) > 1e-6) {
        std::printf("Charge conservation error: q_tot = %.6f, q_tot_loc = %.6f\n", q_tot, loc);
        std::fflush(stdout);
        if (strict_charge_check) std::exit(1);
        return 1;
    }
    q_tot = loc;
    return 0;
}

// library/src/grid_maze_lcg.c:100-124 and the matching init_field of the
// other grid types: cold solve A phi = -4 pi q / (h eps) from zero.
void Sim::init_field() {
    init_field_issue();
    init_field_finish();
}

// The cold solve in two halves, like an MD step: init_field_issue queues it up
// to the publication of its state (init_field_ready polls for it) and
// init_field_finish completes it.
void Sim::init_field_issue() {
    hist_ok = 0;
    double c = -4 * kPi / h;
    if (!pb) c /= eps_s;
    launch_fill(stream, G, 0.0, y);
    launch_copy(stream, G, phi, phip);
    launch_scale_copy(stream, G, c, q, sigma);
    launch_fill(stream, G, 0.0, phi);
    cold_mixed = mixed_init && (precision == kMixedCG || precision == kMixedMG) &&
                 grid_type == kMazeLCG;
    if (cold_mixed) {
        cold_issue();
        return;
    }
    cold_pred = std::max(32, last_iters);
    cg_issue(CgRun{dims, pbc(), phi, r, p0, p1, n * n, 0, std::max(4, cold_pred + 1), &cold_pred},
             sigma, tol);
}

bool Sim::init_field_ready() const { return published(); }

void Sim::init_field_finish() {
    if (cold_mixed) cold_finish();
    else cg_finish();
}

// The mixed cold solve: the refinement loop from phi = 0 to the FP64 stop test
// max|A phi - sigma| <= tol, with FP32 inner solves; phi* (tmp) is free here
// and serves as the second iterate.  kColdPasses passes of up to kColdIters
// inner iterations are queued at once (iterations after an inner solve has
// met its tolerance, and passes after the loop has, are no-ops); further
// passes are polled.  The MD steps' pass schedule is left alone.
void Sim::cold_issue() {
    if (!d32) {
        d32 = dalloc<float>(G); r32 = dalloc<float>(G);
        q32a = dalloc<float>(G); q32b = dalloc<float>(G); y_alt = dalloc<double>(G);
    }
    MixedRun& m = mx;
    m.use_mg = precision == kMixedMG && (pb ? setup_pbmg() : setup_mg());
    if (m.use_mg && !pb) mg[0].b = r32;
    if (m.use_mg && pb) {
        pbmg_coefs();
        pbmg[0].b = r32;
    }
    m.inner_limit = inner_limit_for(m.use_mg);
    m.ra = RefineArgs{tol, inner_rtol_for(m.use_mg), m.inner_limit, m.use_mg ? 1 : 0};
    m.extrap = false;
    m.ya = phi;
    m.yb = tmp;
    mixed_refine(0);
    for (m.queued = 0; m.queued < kColdPasses; m.queued++)
        mixed_pass(m.queued, std::min(kColdIters, m.inner_limit), false);
    publish();
}

void Sim::cold_finish() {
    MixedRun& m = mx;
    wait_published();
    while (!st_host->outer_done && st_host->passes < kMaxOuter) {
        mixed_pass(m.queued++, 1, true);
        read_state();
    }
    const bool converged = st_host->outer_done;
    if (st_host->ysel) launch_copy(stream, G, m.yb, phi);
    if (!converged) {  // finish with FP64 CG from the refined iterate
        int pred = std::max(32, last_iters);
        cg_solve(dims, pbc(), sigma, phi, tol, n * n, r, p0, p1, &pred);
    }
}

int Sim::update_field() {
    switch (grid_type) {
        case kLCG:
            return lcg_update();
        case kMazeLCG:
            if (explicit_pcg()) return maze_pcg_mg();
            if (use_fp64mg()) {
                fp64mg_issue();
                return fp64mg_finish();
            }
            if (precision == kFP64 || precision == kFP64MG) return maze_fp64();
            return maze_mixed();  // kMixedCG / kMixedMG
        case kMG:
            return mg_update(false);
        case kMazeMG: {
            const int res = mg_update(true);
            if (precond != kPNone) fatal("Maze Multigrid with preconditioner not implemented yet.");
            return res;
        }
    }
    return -1;
}

void Sim::forces_elec() { launch_force_fd(stream, p, phi); }

double Sim::forces_noel(bool read_back) {
    timed(kTTf, 0, [&] { launch_tf(stream, p, ps, red); });
    if (!read_back) return 0.0;
    MAZE_HIP(hipMemcpyAsync(&ps_host->pe_tf, &ps->pe_tf, sizeof(double), hipMemcpyDeviceToHost,
                            stream));
    sync();
    return ps_host->pe_tf;
}

void Sim::forces_tot() {
    if (pbd.active) launch_force_total(stream, p, pbd.fdb, pbd.fib, pbd.fnp);
    else launch_force_total(stream, p);
}

// library/src/integrator_verlet.c:12-71, integrator_ovrvo.c:45-138.
bool Sim::part1(bool with_base) {
    if (itype == kVerlet) {
        if (enabled) launch_kinetic(stream, p, ps, red, T, 1);
        launch_verlet_part1(stream, p, dt, ps, enabled, with_base);
        return with_base;
    }
    if (enabled) launch_o_block(stream, p, seed, calls++, std::sqrt(c1), (1 - c1) * kBoltz * T);
    launch_kick(stream, p, 0.5 * dt * c2);
    launch_drift(stream, p, c2 * dt);
    return false;
}

void Sim::part2() {
    if (itype == kVerlet) {
        launch_kick(stream, p, 0.5 * dt);
    } else {
        launch_kick(stream, p, 0.5 * dt * c2);
        if (enabled) launch_o_block(stream, p, seed, calls++, std::sqrt(c1), (1 - c1) * kBoltz * T);
    }
}

// One MD step in the order of maze_poisson/solver.py:483-492; the Tosi-Fumi
// forces overlap the field solve on the side stream.
void Sim::md_step() {
    step_issue();
    step_finish();
}

// The first half of an MD step, queued without waiting: the first integrator
// half, the pair forces (side stream), the charge spread and the field solve
// up to the publication of the solver state.
void Sim::step_issue() {
    const bool base_done = part1(true);
    step_tf = tf_overlap && p.np > 0;
    auto issue_tf = [&] {
        MAZE_HIP(hipEventRecord(ev_pos, stream));
        MAZE_HIP(hipStreamWaitEvent(side, ev_pos, 0));
        timed(kTTf, 0, [&] { launch_tf(side, p, ps, red_side); });
        MAZE_HIP(hipEventRecord(ev_tf, side));
    };
    if (step_tf && !tf_late) issue_tf();
    update_charges(true, base_done);
    if (step_tf && tf_late) issue_tf();
    update_eps_k2();
    step_wait = 1;
    spec_post = spec_hit = 0;
    const bool pcg_mg = explicit_pcg();
    if (grid_type == kMazeLCG && !pcg_mg) {
        step_mg64 = use_fp64mg() ? 1 : 0;
        if (step_mg64) fp64mg_issue();
        else if (precision == kFP64 || precision == kFP64MG) fp64_issue();
        else mixed_issue();
        // The queued passes usually converge: the field update and the
        // force / half-kick go into the queue now, each doing nothing unless
        // the device's outer stop flag is set when it runs, so the GPU does
        // not idle while the host reads the published state.  step_finish
        // launches them itself when they did nothing.
        if (!step_mg64 && precision != kFP64 && precision != kFP64MG && step_tf && !pb &&
            !pbd.active && itype == kVerlet) {
            MAZE_HIP(hipStreamWaitEvent(stream, ev_tf, 0));
            launch_phi_update_gated(stream, G, st, mx.ya, mx.yb, phi);
            launch_force_fd_kick(stream, p, phi, 0.5 * dt, &st->outer_done);
            spec_post = 1;
        }
    } else {
        // The other solvers wait for their state themselves; one more read
        // only if none was made (it brings the charge sum).
        const unsigned long reads = pub_count;
        step_res = update_field();
        if (pub_count == reads) publish();
        else step_wait = 0;
    }
}

// Whether the state published by step_issue has arrived (never blocks).
bool Sim::step_ready() const { return !step_wait || published(); }

// The second half: completes the solve (further passes if the queued ones
// fell short), then the forces and the second integrator half.  Returns the
// solver's result.
int Sim::step_finish() {
    int res = step_res;
    const bool pcg_mg = explicit_pcg();
    if (grid_type == kMazeLCG && !pcg_mg) {
        if (step_mg64) res = fp64mg_finish();
        else res = precision == kFP64 || precision == kFP64MG ? fp64_finish() : mixed_finish();
    }
    else if (step_wait) wait_published();
    check_charge();
    check_pb_flag();
    if (!pbd.active) {
        // The field force, the total force and the velocity half-kick in one
        // launch (the same values as forces_elec, forces_tot and part2's kick).
        if (!step_tf) forces_noel(false);
        if (!spec_hit) {
            if (step_tf) MAZE_HIP(hipStreamWaitEvent(stream, ev_tf, 0));
            launch_force_fd_kick(stream, p, phi, itype == kVerlet ? 0.5 * dt : 0.5 * dt * c2);
        }
        if (itype != kVerlet && enabled)
            launch_o_block(stream, p, seed, calls++, std::sqrt(c1), (1 - c1) * kBoltz * T);
        return res;
    }
    forces_elec();
    if (!step_tf) forces_noel(false);
    forces_pb();
    if (step_tf) MAZE_HIP(hipStreamWaitEvent(stream, ev_tf, 0));
    forces_tot();
    part2();
    return res;
}

void Sim::rescale_velocities() { launch_rescale_velocities(stream, p, ps); }

double Sim::kinetic() {
    launch_kinetic(stream, p, ps, red, 0.0, 0);
    MAZE_HIP(hipMemcpyAsync(&ps_host->kinetic, &ps->kinetic, sizeof(double), hipMemcpyDeviceToHost,
                            stream));
    sync();
    return ps_host->kinetic;
}

double Sim::temperature() { return 2 * kinetic() / (3 * p.np * kBoltz); }

double Sim::pe_tf() {
    MAZE_HIP(hipMemcpyAsync(&ps_host->pe_tf, &ps->pe_tf, sizeof(double), hipMemcpyDeviceToHost,
                            stream));
    sync();
    return ps_host->pe_tf;
}

int Sim::check_thermostat() {
    if (enabled) {
        const double temp = 2 * kinetic() / (3 * p.np * kBoltz);
        if (std::fabs(temp - T) < 100) {
            enabled = 0;
            if (itype == kOVRVO) { c1 = 1.0; c2 = 1.0; }
            return 1;
        }
    }
    return 0;
}

// 1/2 sum_g q_g phi_g: the electrostatic energy of the grid charges in the
// field (Hartree; phi solves A phi = -4 pi q / (h eps), i.e. the potential in
// atomic units).
double Sim::energy_qphi() {
    launch_dot_scaled(stream, G, 0.5, q, phi, dscalar, red);
    MAZE_HIP(hipMemcpyAsync(hscalar, dscalar, sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    return hscalar[0];
}

double Sim::energy_elec() {
    launch_half_sum(stream, G, phi, dscalar, red);
    MAZE_HIP(hipMemcpyAsync(hscalar, dscalar, sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    return hscalar[0];
}

void Sim::momentum(double* out) {
    launch_momentum(stream, p, ps);
    MAZE_HIP(hipMemcpyAsync(ps_host->mom, ps->mom, 3 * sizeof(double), hipMemcpyDeviceToHost, stream));
    sync();
    out[0] = ps_host->mom[0];
    out[1] = ps_host->mom[1];
    out[2] = ps_host->mom[2];
}

}  // namespace maze

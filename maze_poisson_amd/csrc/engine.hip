// engine.hip -- many independent replicas on one GPU, advanced from one host
// thread: the batched ensemble and the continuous-batching request loop.
//
// Each replica is a full Sim (its own device state, streams and reduction
// buffers), so replicas never share memory and a replica's trajectory is the
// same alone and in any mix.  The loop keeps every admitted replica's next
// step queued: a replica whose solver state has been published (Sim::
// step_ready) is finished (forces, second integrator half) and its next step
// is issued at once, so the GPU always holds work of several replicas on
// several streams and small grids fill the chip together.  Requests are
// admitted in arrival order when a slot is free and their device-memory
// estimate fits the budget; a finished replica's results are copied out and
// its memory returned before the next admission.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <cmath>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <memory>
#include <mutex>
#include <thread>
#include <vector>

#include "sim.hpp"

namespace maze {

namespace {

double now_s() {
    using C = std::chrono::steady_clock;
    return std::chrono::duration<double>(C::now().time_since_epoch()).count();
}

double quantile(std::vector<double> v, double q) {
    if (v.empty()) return 0.0;
    std::sort(v.begin(), v.end());
    const double x = q * (double)(v.size() - 1);
    const size_t i = (size_t)x;
    if (i + 1 >= v.size()) return v.back();
    return v[i] + (x - (double)i) * (v[i + 1] - v[i]);
}

}  // namespace

// Physics of a request beyond maze_engine_submit's defaults (MAZE-LCG without
// a preconditioner, CIC, Tosi-Fumi, velocity Verlet without thermostat).
struct Phys {
    int grid_type = kMazeLCG, precond = kPNone, cas = 0;
    double eps_s = 1.0, eps_int = 1.0;
    int itype = kVerlet, thermostat = 0, rescale = 0;
    double T = 0.0, gamma = 0.0;
    int seed_set = 0;
    unsigned long long seed = 0;
    int pb = 0;  // 0 Poisson, 1 PB on given maps, 2 PB maps from the ion radii every step
    double w = 0.0, kbar2 = 0.0, gamma_np = 0.0, beta_np = 0.0;
    int nonpolar = 0;
    std::vector<double> ex, ey, ez, k2, radii;
};

struct Request {
    // inputs
    double arrival = 0;
    int n = 0, np = 0, n_typ = 0, precision = kMixedMG, steps = 0;
    double L = 0, tol = 0, dt = 0;
    std::vector<int> types;
    std::vector<double> pos, vel, mass, charges, tf;
    Phys ph;
    int keep_field = 0;
    // progress
    int state = 0;  // 0 waiting, 1 running, 2 done, 3 withdrawn
    int used_precision = kFP64;
    int done = 0, failures = 0;
    long long est_bytes = 0, meas_bytes = -1, arena_used = 0, arena_overflow = 0, free0 = 0;
    long long stream_bytes = 0;  // drop of free memory when its streams were created (0: reused)
    int boot = 0;  // start-up stage: 1 / 2 while the first / second cold solve runs
    std::unique_ptr<DevArena> arena;
    double t_admit = 0, t_done = 0, t_last = 0;
    std::vector<double> lat;  // seconds per step, as observed by the loop
    std::unique_ptr<Sim> sim;
    std::pair<hipStream_t, hipStream_t> streams{nullptr, nullptr};
    std::vector<double> out_pos, out_vel, out_f, out_phi;
    double obs[7] = {0, 0, 0, 0, 0, 0, 0};  // kinetic, temperature, E_elec, V_TF, momentum
};

// The precision a replica of an n^3 grid runs: the FP32 multigrid needs an
// even grid that halves to at most 128 points (Sim::mg_sizes), and with PB a
// coarsest level of at most 16 points; otherwise the FP64 CG, which is faster
// than the FP32 Jacobi-CG inside refinement.  The mixed paths replace the
// MAZE-LCG solve without preconditioner; the reference's other methods and
// preconditioners run in FP64 as asked.
static int precision_for(int n, int want, const Phys& ph) {
    if (ph.grid_type != kMazeLCG || ph.precond != kPNone) return kFP64;
    if (want == kMixedMG) {
        const std::vector<int> sz = Sim::mg_sizes(n);
        if (sz.empty() || (ph.pb && sz.back() > kMgCoarseLds)) return kFP64;
    }
    return want;
}

// Device bytes a replica carves from its arena: the Sim's reduction buffers
// and solver state, the grids, the multigrid levels, the particle arrays, the
// Tosi-Fumi partials and the charge-spread lists (each allocation 256-byte
// aligned).
static long long arena_bytes(int n, int np, int precision, const Phys& ph) {
    const long long G = (long long)n * n * n;
    // reduction partials of both streams and of the deferred inner iteration, scalars
    long long b = 8LL * 8 * (1 << 15) + (64 << 10);
    int allocs = 64;
    b += 9 * 8 * G;                                   // q, y, phi, phi_prev, phi*, sigma, r, p0, p1
    if (ph.grid_type == kMG || ph.grid_type == kMazeMG ||
        (ph.grid_type == kMazeLCG && ph.precond == kPMG)) {
        const long long Gc = G / 8;                   // the reference's two-grid cycle
        b += 8 * G + 5 * 8 * Gc;                      // fine residual, coarse rhs / eps / r / p
        if (ph.grid_type == kMazeLCG) b += 8 * G;     // preconditioned residual
        if (ph.pb && ph.grid_type != kMazeLCG) b += 4 * 8 * Gc;  // coarse PB coefficients
    }
    if (ph.grid_type == kMazeLCG && ph.precond == kPBlockJacobi)
        b += 8 * G + 8 * kBlockJacobiTab;             // preconditioned residual, eigenbasis tables
    if (ph.pb) b += 4 * 8 * G;                        // dielectric maps and screening
    if (precision != kFP64) b += 8 * G + 4 * 4 * G;   // second y, FP32 d, r, p0, p1
    if (precision == kMixedMG) {
        const std::vector<int> sz = Sim::mg_sizes(n);
        for (size_t l = 0; l < sz.size(); l++) {
            const long long g = (long long)sz[l] * sz[l] * sz[l];
            if (ph.pb)  // FP32 coefficients, inverse diagonal, d, t (+ b below the finest)
                b += (l == 0 ? 7 : 8) * 4 * g;
            else
                b += (l == 0 ? 2 : 3) * 4 * g;  // level 0 borrows the FP32 residual as b
        }
        allocs += 8 * (int)sz.size();
        if (!sz.empty() && !ph.pb) {
            const long long c = sz.back();
            b += 8 * (c * c + c * c * c);  // coarsest eigenbasis and inverse eigenvalues
            if (c > kMgCoarseLds) b += 2 * 8 * c * c * c;  // its work arrays
        }
    }
    // types, pos/vel/forces, mass/charges, base cell, pair counts, the charge
    // spread's contributions with their node and link entries (8 nodes per ion
    // for CIC, 64 for the splines), the pair energies
    const long long nn = ph.cas == 0 ? 8 : 64;
    b += (long long)np * (4 + 5 * 24 + 2 * 8 + 12 + 4 + nn * (8 + 4 + 4) + 8);
    b += 8 * G;                                           // spread list heads
    b += 8LL * tf_part_doubles(np) + 7LL * 8 * 16 * 16;  // pair table, up to 16 types
    if (ph.pb == 2) {
        // radii, squared reach, map / force reach, bin keys and ids (sorted and
        // not), the three boundary forces, the key sort's work space
        const int nb = (n + kPbTile - 1) / kPbTile;
        const unsigned long bins = (unsigned long)nb * nb * nb;
        int key_bits = 1;
        while ((1ul << key_bits) < bins) key_bits++;
        b += (long long)np * (3 * 8 + 2 * 4 + 4 * 4 + 3 * 24);
        b += (long long)std::max<size_t>(pb_sort_bytes(std::max(np, 1), key_bits), 16);
        allocs += 16;
    }
    return b + (long long)allocs * (long long)kArenaAlign;  // alignment of each allocation
}

// Device memory a replica takes: its arena (hipMalloc rounds to 2 MiB) plus
// what the runtime allocates beyond it for the replica's events and launches
// (measured: at most 2 MiB). Per-process costs stay outside: the first
// kernels load their code objects (17 MiB) and the first streams set up the
// hardware queues (184 MiB).
constexpr long long kReplicaExtra = 10LL << 20;
static long long estimate_bytes(int n, int np, int precision, const Phys& ph) {
    const long long mib2 = 2LL << 20;
    return (arena_bytes(n, np, precision, ph) + mib2 - 1) / mib2 * mib2 + kReplicaExtra;
}

// Hardware queues the runtime gives one process (its GPU_MAX_HW_QUEUES, 4 by
// default).
static int hw_queues() {
    const char* v = std::getenv("GPU_MAX_HW_QUEUES");
    const int q = v ? std::atoi(v) : 4;
    return std::min(std::max(q, 1), 32);
}

struct Engine {
    long long budget = 0;
    int max_slots = 1;
    int threads = 1;
    int graphs = 1;  // replicas replay their inner iterations as HIP graphs
    int mixed_init = 1;  // start-up cold solves with the mixed path (Sim::mixed_init)
    std::vector<Request> reqs;
    // last run; the admission queue and the memory account are shared by the
    // worker threads under `mu`
    std::mutex mu;
    std::vector<int> order;
    size_t next = 0;
    std::atomic<int> left{0};
    int running = 0, max_running = 0;
    long long used = 0, peak = 0;
    double t0 = 0, t_end = 0;
    double host_admit = 0, host_issue = 0, host_finish = 0;  // host seconds per phase
    int run_threads = 1;  // worker threads of the current run

    // Streams of retired replicas, reused by the next ones (creating a stream
    // costs milliseconds); destroyed with the engine.
    std::vector<std::pair<hipStream_t, hipStream_t>> stream_pool;

    // New streams come Q main streams, then their Q side streams, Q being the
    // runtime's hardware queues per process (GPU_MAX_HW_QUEUES, four by
    // default): the runtime hands streams to its queues in turn, so the
    // replicas' solve streams spread over all queues instead of alternating
    // with the short pair-force streams. The pool so holds at most Q - 1 idle
    // pairs beyond those of the replicas that ever ran at once.
    // `created` gets the drop of free device memory over a new batch: the
    // runtime sets up its hardware queues (a per-process cost, at most four
    // queues) as the first streams are made, so it is not a per-replica cost.
    std::pair<hipStream_t, hipStream_t> take_streams(long long& created) {
        created = 0;
        std::lock_guard<std::mutex> g(mu);
        if (stream_pool.empty()) {
            const int q = hw_queues();
            size_t f0 = 0, f1 = 0, total = 0;
            MAZE_HIP(hipMemGetInfo(&f0, &total));
            std::vector<hipStream_t> mains(q), sides(q);
            for (auto& m : mains) MAZE_HIP(hipStreamCreateWithFlags(&m, hipStreamNonBlocking));
            for (auto& sd : sides) MAZE_HIP(hipStreamCreateWithFlags(&sd, hipStreamNonBlocking));
            for (int k = q - 1; k >= 0; k--) stream_pool.push_back({mains[k], sides[k]});
            MAZE_HIP(hipMemGetInfo(&f1, &total));
            created = (long long)f0 - (long long)f1;
        }
        auto p = stream_pool.back();
        stream_pool.pop_back();
        return p;
    }

    void give_streams(std::pair<hipStream_t, hipStream_t> p) {
        std::lock_guard<std::mutex> g(mu);
        stream_pool.push_back(p);
    }

    ~Engine() {
        for (auto& p : stream_pool) {
            MAZE_HIP(hipStreamDestroy(p.first));
            MAZE_HIP(hipStreamDestroy(p.second));
        }
    }

    // Set-up of an admitted request: its arena, the replica state, the two
    // start-up steps with cold solves and the first step's issue.
    void admit(Request& r) {
        r.used_precision = precision_for(r.n, r.precision, r.ph);
        // Streams come from the pool, created four pairs at a time and counted
        // apart from the replica's own memory.
        r.streams = take_streams(r.stream_bytes);
        size_t free0 = 0, total = 0;
        MAZE_HIP(hipMemGetInfo(&free0, &total));
        r.arena = std::make_unique<DevArena>();
        r.arena->size = (size_t)arena_bytes(r.n, r.np, r.used_precision, r.ph);
        MAZE_HIP(hipMalloc(&r.arena->base, r.arena->size));
        arena_install(r.arena.get());
        r.sim = std::make_unique<Sim>(r.streams.first, r.streams.second);
        Sim& s = *r.sim;
        const double h = r.L / r.n;
        const Phys& ph = r.ph;
        // The single-replica set-up order (maze_poisson/solver.py:97-106):
        // grid, PB grid and maps, particles, PB radii, integrator.
        s.init_grid(r.n, r.L, h, r.tol, ph.eps_s, ph.eps_int, ph.grid_type, ph.precond);
        if (ph.pb) {
            s.init_pb(ph.w, ph.kbar2, ph.nonpolar);
            if (ph.pb == 1) s.set_eps_k2(ph.ex.data(), ph.ey.data(), ph.ez.data(), ph.k2.data());
        }
        s.precision = r.used_precision;
        s.mg_autotune = 1;
        s.init_particles(r.n, r.n_typ, r.L, h, r.np, ph.cas, r.types.data(), r.pos.data(),
                         r.vel.data(), r.mass.data(), r.charges.data(), r.tf.data());
        if (ph.pb == 2) s.init_pb_particles(ph.gamma_np, ph.beta_np, ph.radii.data());
        s.init_integrator(r.dt, ph.T, ph.gamma, ph.itype, ph.thermostat);
        if (ph.seed_set) {
            s.seed = ph.seed;
            s.calls = 0;
        }
        // Stream capture is invalidated by HIP calls that other threads make
        // meanwhile (hipStreamEndCapture reports it), so only a one-thread run
        // replays graphs.
        s.use_graphs = graphs && run_threads == 1;
        s.mixed_init = mixed_init;
        // The two start-up steps with cold solves (maze_poisson/solver.py:376-398)
        // run through the loop like MD steps: the first cold solve is queued
        // here, and boot() continues each time its state has arrived.
        s.update_charges(true);  // checked when the solve's state arrives
        s.update_eps_k2();       // maps from the radii (a no-op otherwise)
        s.init_field_issue();
        arena_install(nullptr);
        r.free0 = (long long)free0;
        r.state = 1;
        r.boot = 1;
    }

    // Next start-up stage of a replica whose cold solve has published its state;
    // returns true once the replica is stepping (its first MD step queued).
    bool boot(Request& r) {
        Sim& s = *r.sim;
        arena_install(r.arena.get());
        s.init_field_finish();
        s.check_charge();
        s.forces_elec();
        s.forces_noel(false);
        if (r.ph.pb == 2) s.forces_pb();
        s.forces_tot();
        if (r.boot == 1) {
            s.part1();
            s.update_charges(true);
            s.update_eps_k2();
            s.init_field_issue();
            r.boot = 2;
            arena_install(nullptr);
            return false;
        }
        s.part2();
        if (r.ph.rescale) s.rescale_velocities();  // maze_poisson/solver.py:408-409
        r.boot = 0;
        r.t_admit = r.t_last = now_s() - t0;
        if (r.steps > 0) s.step_issue();  // allocates the mixed solver's arrays
        arena_install(nullptr);
        size_t free1 = 0, total = 0;
        MAZE_HIP(hipMemGetInfo(&free1, &total));
        r.meas_bytes = r.free0 - (long long)free1;  // exact with one replica set up at a time
        r.arena_used = (long long)r.arena->used;
        r.arena_overflow = (long long)r.arena->overflow;
        return true;
    }

    // Results out, replica and arena freed.
    void retire(Request& r) {
        Sim& s = *r.sim;
        r.out_pos.resize(3L * r.np);
        r.out_vel.resize(3L * r.np);
        r.out_f.resize(3L * r.np);
        if (r.np > 0) {
            MAZE_HIP(hipMemcpyAsync(r.out_pos.data(), s.p.pos, 24L * r.np, hipMemcpyDeviceToHost,
                                    s.stream));
            MAZE_HIP(hipMemcpyAsync(r.out_vel.data(), s.p.vel, 24L * r.np, hipMemcpyDeviceToHost,
                                    s.stream));
            MAZE_HIP(hipMemcpyAsync(r.out_f.data(), s.p.ft, 24L * r.np, hipMemcpyDeviceToHost,
                                    s.stream));
        }
        s.sync();
        // Observables at the last step: kinetic energy, temperature, electrostatic
        // energy, Tosi-Fumi potential energy, momentum (the single-replica getters).
        r.obs[0] = s.kinetic();
        r.obs[1] = r.np > 0 ? s.temperature() : 0.0;
        r.obs[2] = s.energy_elec();
        r.obs[3] = s.pe_tf();
        s.momentum(r.obs + 4);
        if (r.keep_field) {
            r.out_phi.resize((size_t)s.G);
            MAZE_HIP(hipMemcpy(r.out_phi.data(), s.phi, s.G * sizeof(double),
                               hipMemcpyDeviceToHost));
        }
        r.sim.reset();  // waits for both streams
        give_streams(r.streams);
        arena_forget(r.arena.get());
        MAZE_HIP(hipFree(r.arena->base));
        r.arena.reset();
        r.state = 2;
        r.t_done = now_s() - t0;
        std::lock_guard<std::mutex> g(mu);
        used -= r.est_bytes;
        running--;
        left--;
    }

    // Next request of the queue for worker `mine` (holding `held` replicas of
    // at most `share`), or -1: arrived, a slot free, its estimate in budget.
    int take(int held, int share) {
        std::lock_guard<std::mutex> g(mu);
        if (next >= order.size() || held >= share || running >= max_slots) return -1;
        Request& r = reqs[order[next]];
        if (r.arrival > now_s() - t0 || used + r.est_bytes > budget) return -1;
        used += r.est_bytes;
        peak = std::max(peak, used);
        running++;
        max_running = std::max(max_running, running);
        return order[next++];
    }

    // One worker: admits requests while it has room, finishes the steps whose
    // state has arrived and issues their next steps.
    void worker(int share) {
        std::vector<int> mine;
        double h_admit = 0, h_issue = 0, h_finish = 0;
        while (left.load() > 0) {
            for (int id; (id = take((int)mine.size(), share)) >= 0;) {
                const double ta = now_s();
                admit(reqs[id]);
                h_admit += now_s() - ta;
                mine.push_back(id);
            }
            bool progressed = false;
            for (size_t a = 0; a < mine.size();) {
                Request& r = reqs[mine[a]];
                if (r.boot) {
                    if (!r.sim->init_field_ready()) {
                        a++;
                        continue;
                    }
                    progressed = true;
                    const double ta = now_s();
                    const bool stepping = boot(r);
                    h_admit += now_s() - ta;
                    if (stepping && r.steps == 0) {
                        retire(r);
                        mine.erase(mine.begin() + (long)a);
                        continue;
                    }
                    a++;
                    continue;
                }
                if (!r.sim->step_ready()) {
                    a++;
                    continue;
                }
                progressed = true;
                const double tf = now_s();
                if (r.sim->step_finish() < 0) r.failures++;
                const double tn = now_s() - t0;
                h_finish += tn + t0 - tf;
                r.lat.push_back(tn - r.t_last);
                r.t_last = tn;
                if (++r.done == r.steps) {
                    retire(r);
                    mine.erase(mine.begin() + (long)a);
                    continue;
                }
                const double ti = now_s();
                r.sim->step_issue();
                h_issue += now_s() - ti;
                a++;
            }
            if (!progressed) {
                if (mine.empty()) std::this_thread::yield();
                else __builtin_ia32_pause();
            }
        }
        std::lock_guard<std::mutex> g(mu);
        host_admit += h_admit;
        host_issue += h_issue;
        host_finish += h_finish;
    }

    // Advances every waiting request to the end of its steps, with `threads`
    // host threads each driving its own share of the replicas.
    int run() {
        t0 = now_s();
        order.clear();
        for (int i = 0; i < (int)reqs.size(); i++)
            if (reqs[i].state == 0) order.push_back(i);
        std::stable_sort(order.begin(), order.end(),
                         [&](int a, int b) { return reqs[a].arrival < reqs[b].arrival; });
        next = 0;
        left = (int)order.size();
        running = max_running = 0;
        used = peak = 0;
        host_admit = host_issue = host_finish = 0;
        const int nt = std::max(1, std::min(threads, max_slots));
        run_threads = nt;
        const int share = (max_slots + nt - 1) / nt;
        if (nt == 1) {
            worker(share);
        } else {
            std::vector<std::thread> ts;
            for (int w = 0; w < nt; w++) ts.emplace_back([this, share] { worker(share); });
            for (auto& t : ts) t.join();
        }
        t_end = now_s() - t0;
        return 0;
    }
};

}  // namespace maze

using maze::Engine;
using maze::Request;

extern "C" {

struct maze_engine;

maze_engine* maze_engine_create(long long hbm_budget_bytes, int max_slots) {
    auto* e = new Engine();
    if (hbm_budget_bytes <= 0) {
        size_t free_b = 0, total = 0;
        MAZE_HIP(hipMemGetInfo(&free_b, &total));
        hbm_budget_bytes = (long long)(0.9 * (double)free_b);
    }
    e->budget = hbm_budget_bytes;
    e->max_slots = std::max(1, max_slots);
    e->threads = 1;
    return reinterpret_cast<maze_engine*>(e);
}

void maze_engine_set_threads(maze_engine* h, int threads) {
    reinterpret_cast<Engine*>(h)->threads = std::max(1, std::min(threads, 64));
}

void maze_engine_set_graphs(maze_engine* h, int on) { reinterpret_cast<Engine*>(h)->graphs = on != 0; }

void maze_engine_set_mixed_init(maze_engine* h, int on) {
    reinterpret_cast<Engine*>(h)->mixed_init = on != 0;
}

void maze_engine_destroy(maze_engine* h) { delete reinterpret_cast<Engine*>(h); }

long long maze_engine_estimate_bytes(int n_grid, int np, int precision) {
    const maze::Phys ph;
    return maze::estimate_bytes(n_grid, np, maze::precision_for(n_grid, precision, ph), ph);
}

int maze_engine_precision_for(int n_grid, int precision) {
    return maze::precision_for(n_grid, precision, maze::Phys{});
}

int maze_engine_submit(maze_engine* h, double arrival_s, int n_grid, double L, double tol,
                       int precision, int np, int n_typ, const int* types, const double* pos,
                       const double* vel, const double* mass, const double* charges,
                       const double* tf_params, double dt, int steps) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (n_grid < 4 || L <= 0 || np < 0 || n_typ < 1 || n_typ > 16 || steps < 0 || precision < 0 ||
        precision > 2) {
        std::fprintf(stderr, "maze_engine_submit: invalid request\n");
        return -1;
    }
    Request r;
    r.arrival = arrival_s;
    r.n = n_grid;
    r.L = L;
    r.tol = tol;
    r.precision = precision;
    r.np = np;
    r.n_typ = n_typ;
    r.dt = dt;
    r.steps = steps;
    r.types.assign(types, types + np);
    r.pos.assign(pos, pos + 3L * np);
    r.vel.assign(vel, vel + 3L * np);
    r.mass.assign(mass, mass + np);
    r.charges.assign(charges, charges + np);
    r.tf.assign(tf_params, tf_params + 5L * n_typ * n_typ);
    r.est_bytes = maze::estimate_bytes(n_grid, np, maze::precision_for(n_grid, precision, r.ph), r.ph);
    if (r.est_bytes > e.budget) {
        std::fprintf(stderr,
                     "maze_engine_submit: request needs about %lld bytes, above the budget of "
                     "%lld\n",
                     r.est_bytes, e.budget);
        return -2;
    }
    e.reqs.push_back(std::move(r));
    return (int)e.reqs.size() - 1;
}

// A queued request whose physics may still change, or nullptr.
static Request* queued(maze_engine* h, int id) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (id < 0 || id >= (int)e.reqs.size() || e.reqs[id].state != 0) return nullptr;
    return &e.reqs[id];
}

// The estimate of a request after a change of its physics: -2 (and the change
// undone) if it no longer fits the budget.
static int reestimate(maze_engine* h, Request& r, const maze::Phys& old) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    const long long est = maze::estimate_bytes(r.n, r.np, maze::precision_for(r.n, r.precision, r.ph), r.ph);
    if (est > e.budget) {
        std::fprintf(stderr, "maze_engine: request needs about %lld bytes, above the budget of %lld\n",
                     est, e.budget);
        r.ph = old;
        return -2;
    }
    r.est_bytes = est;
    return 0;
}

int maze_engine_set_physics(maze_engine* h, int id, int grid_type, int precond, int ca_scheme,
                            double eps_s, double eps_int, int itg_type, double T, double gamma,
                            int thermostat, int seed_set, unsigned long long seed, int rescale) {
    Request* r = queued(h, id);
    const bool mg_type = grid_type == maze::kMG || grid_type == maze::kMazeMG ||
                         (grid_type == maze::kMazeLCG && precond == maze::kPMG);
    const bool ok = r && (grid_type == maze::kLCG || grid_type == maze::kMG ||
                          grid_type == maze::kMazeLCG || grid_type == maze::kMazeMG) &&
                    precond >= maze::kPNone && precond <= maze::kPBlockJacobi &&
                    precond != maze::kPSSOR &&
                    (precond == maze::kPNone || grid_type == maze::kLCG || grid_type == maze::kMazeLCG) &&
                    (precond != maze::kPBlockJacobi || grid_type != maze::kMazeLCG || r->n % 10 == 0) &&
                    (!mg_type || (r->n % 2 == 0 && r->n >= 32)) && ca_scheme >= 0 && ca_scheme <= 2 &&
                    (itg_type == maze::kOVRVO || itg_type == maze::kVerlet) && eps_s > 0 && eps_int > 0;
    if (!ok) {
        std::fprintf(stderr, "maze_engine_set_physics: invalid request or value\n");
        return -1;
    }
    const maze::Phys old = r->ph;
    maze::Phys& ph = r->ph;
    ph.grid_type = grid_type;
    ph.precond = precond;
    ph.cas = ca_scheme;
    ph.eps_s = eps_s;
    ph.eps_int = eps_int;
    ph.itype = itg_type;
    ph.T = T;
    ph.gamma = gamma;
    ph.thermostat = thermostat != 0;
    ph.seed_set = seed_set != 0;
    ph.seed = seed;
    ph.rescale = rescale != 0;
    return reestimate(h, *r, old);
}

int maze_engine_set_pb_maps(maze_engine* h, int id, double kbar2, const double* eps_x,
                            const double* eps_y, const double* eps_z, const double* k2) {
    Request* r = queued(h, id);
    if (!r || !eps_x || !eps_y || !eps_z || !k2) {
        std::fprintf(stderr, "maze_engine_set_pb_maps: invalid request or map\n");
        return -1;
    }
    const maze::Phys old = r->ph;
    maze::Phys& ph = r->ph;
    const size_t G = (size_t)r->n * r->n * r->n;
    ph.pb = 1;
    ph.w = 0.0;
    ph.kbar2 = kbar2;
    ph.nonpolar = 0;
    ph.ex.assign(eps_x, eps_x + G);
    ph.ey.assign(eps_y, eps_y + G);
    ph.ez.assign(eps_z, eps_z + G);
    ph.k2.assign(k2, k2 + G);
    ph.radii.clear();
    return reestimate(h, *r, old);
}

int maze_engine_set_pb_radii(maze_engine* h, int id, double w, double kbar2, int nonpolar,
                             double gamma_np, double beta_np, const double* radii) {
    Request* r = queued(h, id);
    bool ok = r && radii && w >= 0;
    // The radius plus the transition width must span less than the grid (Sim::init_pb_particles).
    for (int i = 0; ok && i < r->np; i++) {
        const double hh = r->L / r->n;
        ok = radii[i] >= 0 && 2 * ((int)std::floor((radii[i] + w) / hh) + 2) + 1 <= r->n;
    }
    if (!ok) {
        std::fprintf(stderr, "maze_engine_set_pb_radii: invalid request, width or radius\n");
        return -1;
    }
    const maze::Phys old = r->ph;
    maze::Phys& ph = r->ph;
    ph.pb = 2;
    ph.w = w;
    ph.kbar2 = kbar2;
    ph.nonpolar = nonpolar != 0;
    ph.gamma_np = gamma_np;
    ph.beta_np = beta_np;
    ph.radii.assign(radii, radii + r->np);
    ph.ex.clear();
    ph.ey.clear();
    ph.ez.clear();
    ph.k2.clear();
    return reestimate(h, *r, old);
}

int maze_engine_cancel(maze_engine* h, int id) {
    Request* r = queued(h, id);
    if (!r) return -1;
    r->state = 3;  // withdrawn: never admitted
    return 0;
}

int maze_engine_keep_field(maze_engine* h, int id, int on) {
    Request* r = queued(h, id);
    if (!r) return -1;
    r->keep_field = on != 0;
    return 0;
}

int maze_engine_run(maze_engine* h) { return reinterpret_cast<Engine*>(h)->run(); }

int maze_engine_result(maze_engine* h, int id, double* pos, double* vel, double* fcs_tot) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (id < 0 || id >= (int)e.reqs.size() || e.reqs[id].state != 2) return -1;
    const Request& r = e.reqs[id];
    const size_t b = 24L * r.np;
    if (pos) std::memcpy(pos, r.out_pos.data(), b);
    if (vel) std::memcpy(vel, r.out_vel.data(), b);
    if (fcs_tot) std::memcpy(fcs_tot, r.out_f.data(), b);
    return 0;
}

int maze_engine_observables(maze_engine* h, int id, double* out, int max) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (id < 0 || id >= (int)e.reqs.size() || e.reqs[id].state != 2 || !out) return -1;
    const int k = std::max(0, std::min(max, 7));
    std::memcpy(out, e.reqs[id].obs, k * sizeof(double));
    return k;
}

int maze_engine_field(maze_engine* h, int id, double* phi) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (id < 0 || id >= (int)e.reqs.size() || e.reqs[id].state != 2 || !phi ||
        e.reqs[id].out_phi.empty())
        return -1;
    const Request& r = e.reqs[id];
    std::memcpy(phi, r.out_phi.data(), r.out_phi.size() * sizeof(double));
    return 0;
}

// out: arrival, admitted, finished (s from the start of the run), steps done,
// mean, p50 and p99 step latency (ms), estimated and measured device bytes,
// precision run, solver failures, arena bytes used, bytes beyond the arena.
int maze_engine_request_stats(maze_engine* h, int id, double* out, int max) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    if (id < 0 || id >= (int)e.reqs.size()) return -1;
    const Request& r = e.reqs[id];
    double mean = 0;
    for (double x : r.lat) mean += x;
    if (!r.lat.empty()) mean /= (double)r.lat.size();
    const double v[14] = {r.arrival, r.t_admit, r.t_done, (double)r.done, 1e3 * mean,
                          1e3 * maze::quantile(r.lat, 0.5), 1e3 * maze::quantile(r.lat, 0.99),
                          (double)r.est_bytes, (double)r.meas_bytes, (double)r.used_precision,
                          (double)r.failures, (double)r.arena_used, (double)r.arena_overflow,
                          (double)r.stream_bytes};
    const int k = std::min(max, 14);
    for (int i = 0; i < k; i++) out[i] = v[i];
    return k;
}

// out: wall time of the last run (s), replica-steps/s, atom-steps/s, p50 and
// p99 step latency over all steps (ms), peak estimated device bytes, most
// replicas running at once, host seconds spent admitting, issuing and
// finishing steps.
int maze_engine_stats(maze_engine* h, double* out, int max) {
    Engine& e = *reinterpret_cast<Engine*>(h);
    std::vector<double> all;
    double steps = 0, atom_steps = 0;
    for (const Request& r : e.reqs) {
        if (r.state != 2) continue;
        all.insert(all.end(), r.lat.begin(), r.lat.end());
        steps += r.done;
        atom_steps += (double)r.done * r.np;
    }
    const double wall = e.t_end;
    const double v[10] = {wall,
                          wall > 0 ? steps / wall : 0.0,
                          wall > 0 ? atom_steps / wall : 0.0,
                          1e3 * maze::quantile(all, 0.5),
                          1e3 * maze::quantile(all, 0.99),
                          (double)e.peak,
                          (double)e.max_running,
                          e.host_admit,
                          e.host_issue,
                          e.host_finish};
    const int k = std::min(max, 10);
    for (int i = 0; i < k; i++) out[i] = v[i];
    return k;
}

}  // extern "C"

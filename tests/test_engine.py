"""The replica engine (many independent replicas on one GPU, C-ABI maze_engine_*):
every replica's trajectory is bit-identical to the single-replica C-ABI path (the
driver's sequence through solver_md_loop_iter), alone and in any mix, under any
memory budget; the admission estimate covers what a replica allocates."""
import numpy as np
import pytest

from maze_poisson_amd import driver, workloads as W
from maze_poisson_amd.engine import ReplicaEngine

STEPS = 6

# (ions, grid, precision asked): 48 halves to 12 (one-work-group coarse solve),
# 66 and 50 halve to odd 33 and 25 (axis-pass coarse solve); 40 runs FP64 as asked.
CASES = [(216, 48, 2), (512, 66, 2), (216, 50, 2), (216, 40, 0)]


def systems():
    return [W.nacl_system(n_ions, n, seed=40 + i) for i, (n_ions, n, _) in enumerate(CASES)]


def single(lib, s, precision, steps):
    """The single-replica C-ABI path with the engine's defaults (mixed cold solves)."""
    cfg = driver.RunConfig(tol=W.poisson_tol(s))
    driver.initialize(lib, s, cfg)
    lib.lib.solver_set_precision(precision)
    lib.lib.solver_set_mixed_init(1)
    driver.bootstrap(lib, cfg)
    for _ in range(steps):
        lib.solver_md_loop_iter()
    snap = driver.snapshot(lib, s, fields=False)
    lib.solver_finalize()
    return snap


def same(a, b):
    return all(np.array_equal(a[k], b[k]) for k in ("pos", "vel", "fcs_tot"))


pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def reference_runs(hip):
    return [single(hip, s, prec, STEPS) for s, (_, _, prec) in zip(systems(), CASES)]


def test_engine_matches_single_replica_path(hip, reference_runs):
    with ReplicaEngine(hip, max_slots=8, threads=1) as e:
        ids = [e.submit(s, STEPS, precision=prec) for s, (_, _, prec) in zip(systems(), CASES)]
        st = e.run()
        assert st["max_running"] == len(CASES)
        for rid, ref in zip(ids, reference_runs):
            rs = e.request_stats(rid)
            assert rs["steps"] == STEPS and rs["failures"] == 0
            assert same(e.result(rid), ref), rid
        precs = [e.request_stats(r)["precision"] for r in ids]
        assert precs == [2, 2, 2, 0]


def test_engine_alone_and_under_budget(hip, reference_runs):
    ss = systems()
    with ReplicaEngine(hip, max_slots=1) as e:  # one at a time
        ids = [e.submit(s, STEPS, precision=prec) for s, (_, _, prec) in zip(ss, CASES)]
        assert e.run()["max_running"] == 1
        for rid, ref in zip(ids, reference_runs):
            assert same(e.result(rid), ref)
    est = [ReplicaEngine.estimate_bytes(hip, s.n_grid, s.n_p, prec)
           for s, (_, _, prec) in zip(ss, CASES)]
    budget = int(max(est) * 1.5)  # room for one replica at a time, whatever the order
    with ReplicaEngine(hip, budget_bytes=budget, max_slots=64) as e:
        ids = [e.submit(s, STEPS, precision=prec, arrival=0.001 * i)
               for i, (s, (_, _, prec)) in enumerate(zip(ss, CASES))]
        st = e.run()
        assert st["peak_bytes"] <= budget
        for rid, ref in zip(ids, reference_runs):
            assert same(e.result(rid), ref)
        with pytest.raises(MemoryError):
            e.submit(W.nacl_system(4096, 128, seed=2), 1)


def test_engine_mix_of_many(hip, reference_runs):
    """The four reference replicas inside a mix of twelve of uneven sizes, driven by
    one host thread replaying graphs (all twelve admitted before the first step)
    and by four threads, which launch separately even with graphs asked for."""
    ss = systems()
    extra = W.ensemble_systems(8, sizes=(216, 512), seed0=4100)
    results = []
    for threads, graphs in ((1, True), (4, True)):  # four threads launch separately
        with ReplicaEngine(hip, max_slots=16, threads=threads, graphs=graphs) as e:
            mine = [e.submit(s, STEPS, precision=prec) for s, (_, _, prec) in zip(ss, CASES)]
            others = [e.submit(s, STEPS + 3 * i) for i, s in enumerate(extra)]
            st = e.run()
            if threads == 1:
                assert st["max_running"] == 12
            for rid, ref in zip(mine, reference_runs):
                assert same(e.result(rid), ref)
            for rid in others:
                rs = e.request_stats(rid)
                assert rs["failures"] == 0 and rs["p99_ms"] >= rs["p50_ms"] > 0
            results.append([e.result(r) for r in others])
    for a, b in zip(*results):
        assert same(a, b)


def test_engine_reference_cold_solves(hip):
    """With mixed cold solves off, a replica equals the reference call sequence of the
    single-replica path bit for bit (FP64 CG start-up)."""
    s = systems()[0]
    cfg = driver.RunConfig(tol=W.poisson_tol(s))
    driver.initialize(hip, s, cfg)
    hip.lib.solver_set_precision(2)
    driver.bootstrap(hip, cfg)
    for _ in range(STEPS):
        hip.solver_md_loop_iter()
    ref = driver.snapshot(hip, s, fields=False)
    hip.solver_finalize()
    with ReplicaEngine(hip, mixed_init=False) as e:
        rid = e.submit(s, STEPS)
        e.run()
        assert same(e.result(rid), ref)


def test_engine_memory_estimate(hip):
    """Every allocation of a replica fits its arena (one device allocation sized by
    the admission estimate), and the estimate covers the drop of free device memory
    over the set-up and first step of the first replica of a run (later ones can
    see memory of retired replicas coming back).  Streams are made four pairs at a
    time, outside that drop: one replica at a time reuses the first pair.  A
    warm-up run first loads the kernels' code objects, a per-process cost."""
    ss = systems()
    with ReplicaEngine(hip, max_slots=1) as w:
        w.submit(ss[0], 1)
        w.run()
    with ReplicaEngine(hip, max_slots=1) as e:
        ids = [e.submit(s, 1, precision=prec) for s, (_, _, prec) in zip(ss, CASES)]
        e.run()
        for k, rid in enumerate(ids):
            rs = e.request_stats(rid)
            assert rs["arena_overflow"] == 0, rs
            assert 0 < rs["arena_used"] <= rs["est_bytes"] <= rs["arena_used"] + (14 << 20), rs
            if k == 0:
                assert 0 < rs["meas_bytes"] <= rs["est_bytes"], rs
                assert rs["stream_bytes"] >= 0, rs
            else:
                assert rs["stream_bytes"] == 0, rs


# Physics beyond the defaults, each through the engine and through the single-replica
# path with the same configuration: (ions, grid, system seed, density, precision,
# RunConfig fields, OVRVO seed, rescale after start-up).
GAMMA_1PS = 2.4189e-5  # 1 / ps in a.u.
PHYS_CASES = {
    # OVRVO at 1500 K with the thermostat, linear PB on a synthetic dielectric, FP32 PB
    # multigrid-PCG in FP64 refinement
    "ovrvo_pb_maps": (216, 32, 3, 1.379, 2, dict(integrator="OVRVO", thermostat=True,
                                                 gamma=GAMMA_1PS, pb=True, eps_seed=4), 11, False),
    # cubic splines, PB maps rebuilt from the radii with the non-polar term
    "spline_pb_radii": (216, 36, 32, 1.379, 2, dict(cas="SPL_CUBIC", eps_s=80.0, pb_radii=True,
                                                    pb_w=1.0, nonpolar=True, gamma_np=0.05),
                        None, False),
    # the reference LCG method with the Jacobi preconditioner, quadratic splines, the
    # velocity rescale after start-up
    "lcg_jacobi_rescale": (216, 32, 5, 1.379, 0, dict(method="LCG", precond="JACOBI",
                                                      cas="SPL_QUADR"), None, True),
    # MAZE-LCG with the reference's multigrid preconditioner, Verlet thermostat
    "maze_lcg_mg_thermostat": (216, 32, 6, 1.379, 0, dict(precond="MG", thermostat=True,
                                                          T=1300.0), None, False),
    # the reference's MaZe multigrid
    "maze_multigrid": (216, 32, 7, 1.379, 0, dict(method="MAZE-MULTIGRID"), None, False),
}


def phys_system(name):
    n_ions, n, seed, dens, prec, fields, rng, rescale = PHYS_CASES[name]
    s = W.nacl_system(n_ions, n, seed=seed, density=dens)
    return s, driver.RunConfig(tol=W.poisson_tol(s), **fields), prec, rng, rescale


def phys_single(lib, name, steps):
    s, cfg, prec, rng, rescale = phys_system(name)
    driver.initialize(lib, s, cfg)
    if rng is not None:
        lib.solver_set_rng_seed(rng)
    lib.lib.solver_set_precision(prec)
    lib.lib.solver_set_mixed_init(1)
    driver.bootstrap(lib, cfg)
    if rescale:
        lib.solver_rescale_velocities()
    for _ in range(steps):
        lib.solver_md_loop_iter()
    snap = driver.snapshot(lib, s)
    snap["potential_tf"] = lib.solver_compute_forces_noel()  # same positions, same value
    lib.solver_finalize()
    return snap


@pytest.fixture(scope="module")
def phys_runs(hip):
    return {name: phys_single(hip, name, STEPS) for name in PHYS_CASES}


def test_engine_physics_matches_single_replica_path(hip, phys_runs):
    """Every physics option of the engine (integrator and thermostat, charge assignment,
    method and preconditioner, PB on given maps or from the radii, the start-up rescale),
    all replicas in one mix with a default one: trajectories, observables and the final
    field equal the single-replica path's bit for bit."""
    with ReplicaEngine(hip, max_slots=8) as e:
        ids = {}
        for name in PHYS_CASES:
            s, cfg, prec, rng, rescale = phys_system(name)
            ids[name] = e.submit(s, STEPS, precision=prec, cfg=cfg, seed=rng, rescale=rescale,
                                 keep_field=True)
        plain = e.submit(systems()[0], STEPS)
        st = e.run()
        assert st["max_running"] == len(PHYS_CASES) + 1
        for name, rid in ids.items():
            ref = phys_runs[name]
            rs = e.request_stats(rid)
            assert rs["steps"] == STEPS and rs["failures"] == 0, name
            assert rs["arena_overflow"] == 0, name
            assert same(e.result(rid), ref), name
            ob = e.observables(rid)
            for k in ("kinetic", "temperature", "energy_elec", "potential_tf"):
                assert ob[k] == ref[k], (name, k, ob[k], ref[k])
            assert np.array_equal(ob["momentum"], ref["momentum"]), name
            assert np.array_equal(e.field(rid), ref["field"]), name
        assert e.request_stats(plain)["failures"] == 0
        with pytest.raises(RuntimeError):
            e.field(plain)  # not kept


def test_engine_physics_rejects_invalid(hip):
    s = systems()[0]
    with ReplicaEngine(hip, max_slots=1) as e:
        rid = e.submit(s, 1)
        with pytest.raises(ValueError):  # the multigrid methods need an even n >= 32
            e.submit(W.nacl_system(64, 16, seed=1), 1, cfg=driver.RunConfig(method="MAZE-MULTIGRID"))
        with pytest.raises(ValueError):  # SSOR is not provided
            e.submit(s, 1, cfg=driver.RunConfig(precond="SSOR"))
        st = e.run()  # the refused requests were withdrawn
        assert st["replica_steps_per_s"] > 0 and e.request_stats(rid)["steps"] == 1
        for other in (rid + 1, rid + 2):
            assert e.request_stats(other)["steps"] == 0
            with pytest.raises(RuntimeError):
                e.result(other)

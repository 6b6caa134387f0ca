"""The CPU oracle against golden vectors produced by the reference C (compiled from its
own sources with strict IEEE flags, single thread; tests/golden/make_golden.py).

The oracle restates the reference with the same floating-point operation order, so
the comparison is bit for bit: positions, velocities, forces, energies, the charge
grid, the field and the CG / V-cycle iteration counts."""
import json
import os

import numpy as np
import pytest

from maze_poisson_amd import driver, workloads as W

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), "golden")
CASES = json.load(open(os.path.join(GOLDEN, "ref_md_cases.json")))


@pytest.mark.parametrize("name", sorted(CASES))
def test_md_golden(oracle, name):
    kw = dict(CASES[name])
    n_ions, n_grid, seed = kw.pop("n_ions"), kw.pop("n_grid"), kw.pop("seed")
    method, steps = kw.pop("method"), kw.pop("steps")
    ref = np.load(os.path.join(GOLDEN, f"ref_md_{name}.npz"))
    s = W.nacl_system(n_ions, n_grid, seed=seed, density=kw.pop("density", 1.379))
    cfg = driver.RunConfig(method=method, tol=W.poisson_tol(s), **kw)
    driver.initialize(oracle, s, cfg)
    driver.bootstrap(oracle, cfg)
    its, vs = [], []
    for _ in range(steps):
        it, v = driver.md_step(oracle, cfg)
        its.append(it)
        vs.append(v)
    snap = driver.snapshot(oracle, s)
    oracle.solver_finalize()
    assert its == ref["iters"].tolist()
    assert np.array_equal(np.array(vs), ref["v_tf"])
    keys = ["pos", "vel", "fcs_elec", "fcs_noel", "fcs_tot", "momentum"]
    if cfg.pb_radii:  # radius-based PB boundary forces
        keys += ["fcs_db", "fcs_ib", "fcs_np"]
        assert np.abs(ref["fcs_db"]).max() > 0 and np.abs(ref["fcs_ib"]).max() > 0
        assert (np.abs(ref["fcs_np"]).max() > 0) == cfg.nonpolar
    for k in keys:
        assert np.array_equal(snap[k], ref[k]), k
    for k in ("kinetic", "temperature", "energy_elec"):
        assert snap[k] == float(ref[k]), k
    for k in ("field", "field_prev", "q"):
        a = snap[k]
        assert np.array_equal(a[0], ref[k + "_plane0"]), k
        assert np.array_equal(np.array([a.sum(), (a * a).sum(), np.abs(a).max()]),
                              ref[k + "_stats"]), k


@pytest.fixture(scope="module")
def kv():
    return np.load(os.path.join(GOLDEN, "ref_kernels.npz"))


def test_laplace_golden(oracle_raw, kv):
    u = np.ascontiguousarray(kv["lap_in"])
    o = np.zeros_like(u)
    oracle_raw.or_laplace(u, o, u.shape[0])
    assert np.array_equal(o, kv["lap_out"])


def test_laplace_pb_golden(oracle_raw, kv):
    u = np.ascontiguousarray(kv["lap_in"])
    e = kv["lappb_eps"]
    o = np.zeros_like(u)
    oracle_raw.or_laplace_pb(u, o, u.shape[0], np.ascontiguousarray(e[0]),
                             np.ascontiguousarray(e[1]), np.ascontiguousarray(e[2]),
                             np.ascontiguousarray(kv["lappb_k2"]))
    assert np.array_equal(o, kv["lappb_out"])


def test_conj_grad_golden(oracle_raw, kv):
    b = np.ascontiguousarray(kv["cg_b"])
    x = np.ascontiguousarray(kv["cg_x0"]).copy()
    it = oracle_raw.or_conj_grad(b, x, 1e-10, b.shape[0])
    assert it == int(kv["cg_iter"])
    assert np.array_equal(x, kv["cg_x"])


def test_multigrid_pieces_golden(oracle_raw, kv):
    rhs = np.ascontiguousarray(kv["rbgs_rhs"])
    x = np.ascontiguousarray(kv["rbgs_x0"]).copy()
    n = rhs.shape[0]
    oracle_raw.or_mg_smooth_rbgs(rhs, x, n, 3)
    assert np.array_equal(x, kv["rbgs_out"])
    c = np.zeros((n // 2,) * 3)
    oracle_raw.or_mg_restrict27(rhs, c, n)
    assert np.array_equal(c, kv["restr_out"])
    f = np.zeros((n,) * 3)
    oracle_raw.or_mg_prolong_trilinear(np.ascontiguousarray(kv["prol_in"]), f, n // 2)
    assert np.array_equal(f, kv["prol_out"])


# ---- reference Poisson-Boltzmann multigrid (library/src/multigrid_pb.c) ----------

def pb_coefficients(n, seed):
    """The seeded coefficient fields make_golden.py fed the reference."""
    g = np.random.default_rng(seed)
    eps = [g.uniform(2, 80, (n, n, n)) for _ in range(3)]
    return eps, g.uniform(0, 0.1, (n, n, n))


@pytest.fixture(scope="module")
def pbv():
    return np.load(os.path.join(GOLDEN, "ref_pb_mg.npz"))


def test_pb_rbgs_and_restrictions_golden(oracle_raw, pbv):
    n = 16
    eps, k2 = pb_coefficients(n, 201)
    rhs = np.random.default_rng(202).standard_normal((n, n, n))
    x = np.random.default_rng(203).standard_normal((n, n, n))
    oracle_raw.or_mg_smooth_pb_rbgs(rhs, x, n, 2, *eps, k2)
    assert np.array_equal(x, pbv["rbgspb_out"])
    for axis in range(3):
        c = np.zeros((n // 2,) * 3)
        oracle_raw.or_mg_restrict_eps(eps[axis], c, n, axis)
        assert np.array_equal(c, pbv[f"reps{axis}_out"]), axis
    c = np.zeros((n // 2,) * 3)
    oracle_raw.or_mg_restrict_k2(k2, c, n)
    assert np.array_equal(c, pbv["rk2_out"])


def test_conj_grad_pb_golden(oracle_raw, pbv):
    n = 16
    eps, k2 = pb_coefficients(n, 201)
    b = np.random.default_rng(204).standard_normal((n, n, n))
    x = np.zeros((n, n, n))
    it = oracle_raw.or_conj_grad_pb(b, x, 1e-9, n, *eps, k2)
    assert it == int(pbv["cgpb_iter"])
    assert np.array_equal(x, pbv["cgpb_x"])


def test_vcycle_pb_golden(oracle_raw, pbv):
    n = 32
    eps, k2 = pb_coefficients(n, 205)
    rhs = np.random.default_rng(206).standard_normal((n, n, n))
    x = np.zeros((n, n, n))
    oracle_raw.or_mg_vcycle_pb(rhs, x, n, 4, *eps, k2)
    assert np.array_equal(x, pbv["vcpb_out"])


def test_verlet_pb_multigrid_golden(oracle_raw, pbv):
    n = 32
    eps, k2 = pb_coefficients(n, 205)
    g = np.random.default_rng(207)
    phi, phip = g.standard_normal((n, n, n)), g.standard_normal((n, n, n))
    q = g.standard_normal((n, n, n)) * 1e-3
    y = g.standard_normal((n, n, n)) * 1e-2
    phi_in = phi.copy()
    it = oracle_raw.or_verlet_pb_multigrid(1e-6, 0.5, phi, phip, q, y, n, *eps, k2)
    assert it == int(pbv["vpbmg_iter"])
    assert np.array_equal(y, pbv["vpbmg_y"])
    assert np.array_equal(phi, pbv["vpbmg_phi"])
    assert np.array_equal(phip, phi_in)

"""Pins of the CPU oracle that do not depend on the reference: closed forms,
invariants and brute force, each chosen so that a plausible slip (dropped term,
wrong sign or index, transposed operand) fails."""
import ctypes
import math

import numpy as np
import pytest

from maze_poisson_amd import driver, workloads as W


def lam(n):
    """Eigenvalues of the periodic 7-point operator: sum_d (2 cos(2 pi k_d / n) - 2)."""
    k = 2 * np.cos(2 * np.pi * np.arange(n) / n) - 2
    return k[:, None, None] + k[None, :, None] + k[None, None, :]


@pytest.mark.parametrize("n,kv", [(8, (1, 0, 0)), (12, (1, 2, 3)), (9, (4, 1, 2))])
def test_laplace_plane_wave_eigen(oracle_raw, n, kv):
    i = np.arange(n)
    I, J, K = np.meshgrid(i, i, i, indexing="ij")
    u = np.ascontiguousarray(np.cos(2 * np.pi * (kv[0] * I + 2 * kv[1] * J + 3 * kv[2] * K) / n))
    o = np.zeros_like(u)
    oracle_raw.or_laplace(u, o, n)
    l = sum(2 * math.cos(2 * math.pi * c * m / n) - 2 for c, m in zip(kv, (1, 2, 3)))
    assert np.abs(o - l * u).max() < 1e-12


def test_laplace_asymmetric_brute_force(oracle_raw):
    """Anisotropic input: each of the 6 neighbour directions checked separately."""
    n = 5
    u = np.ascontiguousarray(np.random.default_rng(0).standard_normal((n, n, n)))
    o = np.zeros_like(u)
    oracle_raw.or_laplace(u, o, n)
    ref = sum(np.roll(u, s, axis=a) for a in range(3) for s in (1, -1)) - 6 * u
    assert np.abs(o - ref).max() < 1e-13


def test_laplace_pb_reduces_and_scales(oracle_raw):
    n = 6
    rng = np.random.default_rng(1)
    u = np.ascontiguousarray(rng.standard_normal((n, n, n)))
    one, zero = np.ones_like(u), np.zeros_like(u)
    a, b = np.zeros_like(u), np.zeros_like(u)
    oracle_raw.or_laplace_pb(u, a, n, one, one, one, zero)
    oracle_raw.or_laplace(u, b, n)
    assert np.abs(a - b).max() < 1e-13
    # face-centred coefficients: eps_x[c] couples c and c+x; brute force
    ex, ey, ez = (np.ascontiguousarray(rng.uniform(2, 80, (n, n, n))) for _ in range(3))
    k2 = np.ascontiguousarray(rng.uniform(0, 0.5, (n, n, n)))
    oracle_raw.or_laplace_pb(u, a, n, ex, ey, ez, k2)
    ref = -k2 * u
    for e, ax in ((ex, 0), (ey, 1), (ez, 2)):
        ref += e * (np.roll(u, -1, ax) - u) + np.roll(e, 1, ax) * (np.roll(u, 1, ax) - u)
    assert np.abs(a - ref).max() < 1e-11


@pytest.mark.parametrize("n", [8, 16])
def test_conj_grad_matches_fft_solve(oracle_raw, n):
    b = np.random.default_rng(n).standard_normal((n, n, n))
    b = np.ascontiguousarray(b - b.mean())
    x = np.zeros_like(b)
    tol = 1e-11
    it = oracle_raw.or_conj_grad(b, x, tol, n)
    assert it >= 0
    L = lam(n)
    L[0, 0, 0] = 1.0
    xh = np.fft.fftn(b) / L
    xh[0, 0, 0] = 0.0
    xf = np.real(np.fft.ifftn(xh))
    assert np.abs((x - x.mean()) - xf).max() < 1e3 * tol


def test_conj_grad_stops_on_infinity_norm(oracle_raw):
    """Returns the 0-based index of the first iteration whose residual is below tol."""
    n = 8
    b = np.random.default_rng(3).standard_normal((n, n, n))
    b = np.ascontiguousarray(b - b.mean())
    res = []
    for tol in (1e-2, 1e-6, 1e-10):
        x = np.zeros_like(b)
        it = oracle_raw.or_conj_grad(b, x, tol, n)
        r = np.zeros_like(b)
        oracle_raw.or_laplace(x, r, n)
        assert np.abs(r - b).max() <= tol * 1.01 + 1e-13
        res.append(it)
    assert res[0] < res[1] < res[2]


def test_maze_step_satisfies_constraint(oracle_raw):
    """After the MaZe projection, A phi = -(4 pi / h) q up to the CG tolerance."""
    n, h = 16, 0.7
    rng = np.random.default_rng(5)
    q = rng.standard_normal((n, n, n))
    q = np.ascontiguousarray(q - q.mean())
    phi = np.ascontiguousarray(rng.standard_normal((n, n, n)))
    phip = np.ascontiguousarray(phi + 0.01 * rng.standard_normal((n, n, n)))
    y = np.zeros_like(q)
    phi0 = phi.copy()
    tol = 1e-10
    it = oracle_raw.or_verlet_poisson(tol, h, phi, phip, q, y, n, 0)
    assert it >= 0
    assert np.array_equal(phip, phi0)           # phi_prev <- old phi
    a = np.zeros_like(q)
    oracle_raw.or_laplace(phi, a, n)
    assert np.abs(a + 4 * np.pi / h * q).max() < 20 * tol


def test_cic_weights_partition_of_unity(oracle_raw):
    rng = np.random.default_rng(7)
    n, h = 16, 0.5
    L = n * h
    pos = np.ascontiguousarray(rng.uniform(0, L, (200, 3)))
    nb = np.zeros((200, 8, 3), dtype=np.int64)
    oracle_raw.or_cic_neighbors(pos, 200, n, h, nb)
    for i in range(200):
        w = [np.prod([oracle_raw.or_cic_weight(pos[i, d] - nb[i, c, d] * h, L, h) for d in range(3)])
             for c in range(8)]
        assert abs(sum(w) - 1.0) < 1e-14
        assert np.array_equal(nb[i, 0], np.floor(pos[i] / h).astype(np.int64) % n)


def test_cic_node_and_conservation(oracle_raw):
    n, h = 8, 1.0
    pos = np.ascontiguousarray([[3.0, 5.0, 7.0], [0.5, 0.25, 7.9]])
    chg = np.ascontiguousarray([2.0, -0.5])
    nb = np.zeros((2, 8, 3), dtype=np.int64)
    oracle_raw.or_cic_neighbors(pos, 2, n, h, nb)
    q = np.zeros((n, n, n))
    tot = oracle_raw.or_spread_cic(n, 2, h, pos, nb, chg, q)
    assert q[3, 5, 7] == 2.0
    assert abs(tot - 1.5) < 1e-15 and abs(q.sum() - 1.5) < 1e-14
    # the second ion wraps across the k boundary: corners at k = 7 and 0
    assert abs(q[0, 0, 7] - (-0.5 * 0.5 * 0.75 * 0.1)) < 1e-15
    assert abs(q[0, 0, 0] - (-0.5 * 0.5 * 0.75 * 0.9)) < 1e-15


def test_force_fd_central_difference(oracle_raw):
    """An ion on a node of a sine potential feels q (phi[i-1] - phi[i+1]) / 2h."""
    n, h = 16, 0.5
    i = np.arange(n)
    phi = np.ascontiguousarray(np.broadcast_to(np.sin(2 * np.pi * i / n)[:, None, None],
                                               (n, n, n)).copy())
    pos = np.ascontiguousarray([[3 * h, 2 * h, 5 * h]])
    nb = np.zeros((1, 8, 3), dtype=np.int64)
    oracle_raw.or_cic_neighbors(pos, 1, n, h, nb)
    f = np.zeros((1, 3))
    oracle_raw.or_force_fd(n, 1, h, phi, nb, np.ascontiguousarray([1.5]), pos, f)
    e = (np.sin(2 * np.pi * 2 / n) - np.sin(2 * np.pi * 4 / n)) / (2 * h)
    assert abs(f[0, 0] - 1.5 * e) < 1e-14 and abs(f[0, 1]) < 1e-15 and abs(f[0, 2]) < 1e-15


def tf_pair(table, t, r):
    a, b, c, d, s, al, be = table[t * 7:(t + 1) * 7]
    v = a * np.exp(b * (s - r)) - c / r**6 - d / r**8 + al * r + be
    return v


def test_tf_shifted_force_and_derivative(oracle_raw):
    rc = 12.0
    pp = np.ascontiguousarray(W.tf_params_au())
    table = np.zeros(4 * 7)
    oracle_raw.or_tf_pair_table(2, pp, rc, table)
    for t in range(4):
        assert abs(tf_pair(table, t, rc)) < 1e-15                  # V(rc) = 0
        dv = (tf_pair(table, t, rc + 1e-5) - tf_pair(table, t, rc - 1e-5)) / 2e-5
        assert abs(dv) < 1e-12                                      # F(rc) = 0
    # two ions: force = -dV/dr along the separation, energy = V
    L = 40.0
    for r in (4.5, 6.0, 11.9):
        pos = np.ascontiguousarray([[1.0, 2.0, 3.0], [1.0 + r / math.sqrt(3)] * 1 + [2.0 + r / math.sqrt(3), 3.0 + r / math.sqrt(3)]])
        types = np.ascontiguousarray([0, 1], dtype=np.int32)
        f = np.zeros((2, 3))
        pe = oracle_raw.or_tf_forces(2, L, pos, types, 2, table, L / 2, f)
        rr = np.linalg.norm(pos[0] - pos[1])
        assert abs(pe - tf_pair(table, 1, rr)) < 1e-14
        dv = (tf_pair(table, 1, rr + 1e-6) - tf_pair(table, 1, rr - 1e-6)) / 2e-6
        u = (pos[0] - pos[1]) / rr
        assert np.abs(f[0] - (-dv) * u).max() < 1e-7 * max(1e-3, abs(dv))
        assert np.abs(f[0] + f[1]).max() < 1e-16                      # Newton 3


def test_tf_newton3_and_counts(oracle_raw):
    s = W.nacl_system(216, 32, seed=9)
    table = np.zeros(4 * 7)
    oracle_raw.or_tf_pair_table(2, np.ascontiguousarray(s.tf), s.L / 2, table)
    f = np.zeros((s.n_p, 3))
    oracle_raw.or_tf_forces(s.n_p, s.L, s.pos, s.types, 2, table, s.L / 2, f)
    assert np.abs(f.sum(axis=0)).max() < 1e-12 * np.abs(f).max() * s.n_p
    cnt = np.zeros(s.n_p, dtype=np.int32)
    oracle_raw.or_tf_neighbor_counts(s.n_p, s.L, s.pos, s.L / 2, cnt)
    d = s.pos[:, None, :] - s.pos[None, :, :]
    d -= s.L * np.round(d / s.L)
    r = np.sqrt((d ** 2).sum(-1))
    brute = ((r <= s.L / 2).sum(1) - 1)
    assert np.array_equal(cnt, brute)


PHILOX_KAT = [  # Random123 kat_vectors, philox4x32 10 rounds
    ((0, 0, 0, 0), (0, 0), (0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8)),
    ((0xffffffff,) * 4, (0xffffffff, 0xffffffff), (0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd)),
    ((0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344), (0xa4093822, 0x299f31d0),
     (0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1)),
]


@pytest.mark.parametrize("ctr,key,want", PHILOX_KAT)
def test_philox_known_answers(oracle_raw, ctr, key, want):
    c = (ctypes.c_uint32 * 4)(*ctr)
    k = (ctypes.c_uint32 * 2)(*key)
    o = (ctypes.c_uint32 * 4)()
    oracle_raw.or_philox4x32(ctypes.addressof(c), ctypes.addressof(k), ctypes.addressof(o))
    assert tuple(o) == want


def test_normal_deviates_statistics(oracle_raw):
    out = np.zeros(3)
    xs = []
    for ion in range(20000):
        oracle_raw.or_normal3(42, 3, ion, out)
        xs.append(out.copy())
    xs = np.array(xs)
    assert np.abs(xs.mean(axis=0)).max() < 0.03
    assert np.abs(xs.var(axis=0) - 1).max() < 0.04
    assert abs(np.corrcoef(xs[:, 0], xs[:, 1])[0, 1]) < 0.03


def test_ovrvo_without_thermostat_is_verlet(oracle):
    """gamma -> 0 limit: with the thermostat off, c1 = c2 = 1 and OVRVO is velocity
    Verlet bit for bit."""
    s = W.nacl_system(216, 32, seed=4)
    snaps = []
    for itg in ("OVRVO", "VERLET"):
        cfg = driver.RunConfig(tol=W.poisson_tol(s), integrator=itg)
        driver.initialize(oracle, s, cfg)
        driver.bootstrap(oracle)
        for _ in range(3):
            driver.md_step(oracle)
        snaps.append(driver.snapshot(oracle, s, fields=False))
        oracle.solver_finalize()
    for k in ("pos", "vel", "fcs_tot"):
        assert np.array_equal(snaps[0][k], snaps[1][k])


def test_ovrvo_zero_temperature_is_friction(oracle):
    """T = 0: the O block is v <- sqrt(c1) v exactly."""
    s = W.nacl_system(64, 32, seed=6)
    gamma = 0.01
    cfg = driver.RunConfig(tol=W.poisson_tol(s), integrator="OVRVO", thermostat=True, T=0.0,
                           gamma=gamma)
    driver.initialize(oracle, s, cfg)
    v0 = np.zeros((s.n_p, 3))
    oracle.get_vel(v0)
    oracle.solver_update_charges()
    oracle.solver_init_field()
    driver.compute_forces(oracle)
    f = np.zeros((s.n_p, 3))
    oracle.get_fcs_tot(f)
    oracle.integrator_part_1()
    v1 = np.zeros((s.n_p, 3))
    oracle.get_vel(v1)
    oracle.solver_finalize()
    dt = W.dt_au(0.25)
    c1 = math.exp(-gamma * dt)
    c2 = math.sqrt(2 / (gamma * dt) * math.tanh(0.5 * gamma * dt))
    want = v0 * math.sqrt(c1) + 0.5 * dt * c2 * f / s.mass[:, None]
    assert np.abs(v1 - want).max() < 1e-14 * np.abs(want).max()


def test_verlet_time_reversal(oracle):
    """Velocity Verlet + converged field solves is time reversible: N steps forward,
    flip the velocities, N steps back returns to the start."""
    s = W.nacl_system(64, 32, seed=8, T=300.0)
    cfg = driver.RunConfig(tol=1e-12)
    driver.initialize(oracle, s, cfg)
    driver.bootstrap(oracle)
    x0 = np.zeros((s.n_p, 3)); oracle.get_pos(x0)
    v = np.zeros((s.n_p, 3)); oracle.get_vel(v)
    oracle.solver_finalize()
    s2 = W.System(s.n_grid, s.L, s.types, x0.copy(), v.copy(), s.mass, s.charges)
    lcg = driver.RunConfig(method="LCG", tol=1e-12)
    driver.initialize(oracle, s2, lcg)
    oracle.solver_update_charges(); oracle.solver_init_field(); driver.compute_forces(oracle)
    for _ in range(5):
        driver.md_step(oracle)
    x1 = np.zeros((s.n_p, 3)); oracle.get_pos(x1)
    v1 = np.zeros((s.n_p, 3)); oracle.get_vel(v1)
    oracle.solver_finalize()
    s3 = W.System(s.n_grid, s.L, s.types, x1.copy(), -v1, s.mass, s.charges)
    driver.initialize(oracle, s3, lcg)
    oracle.solver_update_charges(); oracle.solver_init_field(); driver.compute_forces(oracle)
    for _ in range(5):
        driver.md_step(oracle)
    x2 = np.zeros((s.n_p, 3)); oracle.get_pos(x2)
    oracle.solver_finalize()
    d = x2 - x0
    d -= s.L * np.round(d / s.L)
    assert np.abs(d).max() < 1e-8 and np.abs(x1 - x0).max() > 1e-3


def test_multigrid_transfer_operators(oracle_raw):
    n = 16
    c = np.full((n, n, n), 2.5)
    out = np.zeros((n // 2,) * 3)
    oracle_raw.or_mg_restrict27(c, out, n)
    assert np.abs(out - 2.5).max() < 1e-15                  # weights sum to 1
    coarse = np.ascontiguousarray(np.random.default_rng(1).standard_normal((8, 8, 8)))
    fine = np.zeros((n, n, n))
    oracle_raw.or_mg_prolong_trilinear(coarse, fine, 8)
    assert np.array_equal(fine[::2, ::2, ::2], coarse)       # injection at coarse nodes
    mid = 0.5 * (coarse + np.roll(coarse, -1, axis=0))
    assert np.abs(fine[1::2, ::2, ::2] - mid).max() < 1e-15  # linear along i
    mid3 = sum(np.roll(np.roll(np.roll(coarse, -a, 0), -b, 1), -c_, 2)
               for a in (0, 1) for b in (0, 1) for c_ in (0, 1)) / 8
    assert np.abs(fine[1::2, 1::2, 1::2] - mid3).max() < 1e-15


def test_rbgs_fixed_point_and_vcycle_contracts(oracle_raw):
    n = 32
    x = np.random.default_rng(2).standard_normal((n, n, n))
    x = np.ascontiguousarray(x - x.mean())
    b = np.zeros_like(x)
    oracle_raw.or_laplace(x, b, n)
    y = x.copy()
    oracle_raw.or_mg_smooth_rbgs(b, y, n, 2)
    assert np.abs(y - x).max() < 1e-12
    rhs = np.random.default_rng(3).standard_normal((n, n, n))
    rhs = np.ascontiguousarray(rhs - rhs.mean())
    z = np.zeros_like(rhs)
    r0 = np.abs(rhs).max()
    oracle_raw.or_mg_vcycle(rhs, z, n, 3)
    r = np.zeros_like(rhs)
    oracle_raw.or_laplace(z, r, n)
    assert np.abs(r - rhs).max() < 0.2 * r0


def test_workload_geometry():
    assert abs(W.box_length_ang(512) - 26.215) < 2e-3
    assert abs(W.box_length_ang(4096) - 52.429) < 2e-3
    assert abs(W.box_length_ang(8000) - 65.537) < 2e-3
    s = W.nacl_system(2048, 102, seed=1)
    assert (s.types == 0).sum() == 1024 and s.charges.sum() == 0
    assert np.abs((s.mass[:, None] * s.vel).sum(0)).max() < 1e-9
    assert s.pos.min() >= 0 and s.pos.max() < s.L


@pytest.mark.parametrize("which", ["or_spline_quadr_weight", "or_spline_cubic_weight"])
def test_spline_weights_partition_of_unity(oracle_raw, which):
    """B-spline assignment weights: over the 4 nodes base-1..base+2 they sum to 1 and
    reproduce the position (first moment), are even, and take the textbook values at
    the node (3/4 quadratic, 2/3 cubic)."""
    import ctypes
    f = getattr(oracle_raw, which)
    f.restype, f.argtypes = ctypes.c_double, [ctypes.c_double] * 3
    L, h = 10.0, 0.5
    g = np.random.default_rng(9)
    for x in g.uniform(0, L, 50):
        base = np.floor(x / h)
        nodes = (base + np.arange(-1, 3)) * h
        w = np.array([f(x - xn, L, h) for xn in nodes])
        assert abs(w.sum() - 1.0) <= 1e-14
        assert abs((w * nodes).sum() - x) <= 1e-12
        assert abs(f(x - nodes[1], L, h) - f(nodes[1] - x, L, h)) <= 1e-16
    assert f(0.0, L, h) == (0.75 if "quadr" in which else 2.0 / 3.0)
    assert f(2.1 * h, L, h) == 0.0


def pb_step_expected(dist, r, w):
    """Solvent fraction of one ion at distance `dist`: 0 inside r - w, 1 beyond r + w,
    and between them the cubic s^2 (3 w - s) / (4 w^3), s = dist - r + w, which has
    zero slope at both ends and equals 1/2 at the radius."""
    s = dist - r + w
    mid = s * s * (3 * w - s) / (4 * w ** 3)
    return np.where(dist <= r - w, 0.0, np.where(dist >= r + w, 1.0, mid))


def test_pb_step_shape():
    r, w = 2.0, 0.5
    d = np.array([1.0, 1.5, 2.0, 2.5, 3.0])
    assert np.allclose(pb_step_expected(d, r, w), [0.0, 0.0, 0.5, 1.0, 1.0], rtol=0, atol=1e-15)


def test_pb_maps_geometry(oracle_raw):
    """Radius-based PB maps against their geometric definition: eps_x/y/z live on the
    +x/+y/+z face of each point ((i+1/2)h, jh, kh etc.), k2 on the point, each the
    solvent fraction (product over ions, nearest periodic image) scaled into
    [eps_int, eps_s] and [0, kbar2]; two overlapping ions multiply."""
    n, h = 16, 0.5
    L = n * h
    eps_s, eps_int, kbar2, w = 80.0, 2.0, 0.03, 0.4
    pos = np.array([[0.3, 7.9, 4.1], [1.4, 0.2, 4.6], [5.0, 5.0, 1.0]])
    radii = np.array([1.2, 0.9, 1.5])
    ex, ey, ez, k2 = (np.zeros((n, n, n)) for _ in range(4))
    oracle_raw.or_pb_maps(n, h, w, eps_s, eps_int, kbar2, len(pos), pos.reshape(-1), radii,
                          ex, ey, ez, k2)
    g = np.arange(n) * h
    for arr, off, lo, hi in ((ex, (0.5, 0, 0), eps_int, eps_s), (ey, (0, 0.5, 0), eps_int, eps_s),
                             (ez, (0, 0, 0.5), eps_int, eps_s), (k2, (0, 0, 0), 0.0, kbar2)):
        X = np.meshgrid(g + off[0] * h, g + off[1] * h, g + off[2] * h, indexing="ij")
        frac = np.ones((n, n, n))
        for p, r in zip(pos, radii):
            d2 = 0.0
            for a in range(3):
                d = p[a] - X[a]
                d -= L * np.round(d / L)
                d2 = d2 + d * d
            frac *= pb_step_expected(np.sqrt(d2), r, w)
        want = lo + (hi - lo) * frac
        assert np.abs(arr - want).max() <= 1e-13 * hi
    assert (k2 == 0).any() and (k2 == kbar2).any() and ((k2 > 0) & (k2 < kbar2)).any()


def test_pb_forces_symmetry_and_scaling(oracle_raw):
    """A lone ion on a grid point in a field and maps with the cubic symmetry of the
    grid feels no boundary force; a shifted copy of the whole system (ion, field, maps
    by whole cells) gives the same forces; the non-polar term scales with gamma_np
    and vanishes with it."""
    n, h = 24, 0.5
    L = n * h
    eps_s, eps_int, kbar2, w, r = 80.0, 1.0, 0.05, 0.5, 2.0
    g = np.arange(n) * h
    X = np.meshgrid(g, g, g, indexing="ij")

    def system(c):
        pos = np.array([c])
        ex, ey, ez, k2 = (np.zeros((n, n, n)) for _ in range(4))
        oracle_raw.or_pb_maps(n, h, w, eps_s, eps_int, kbar2, 1, pos.reshape(-1),
                              np.array([r]), ex, ey, ez, k2)
        d2 = 0.0
        for a in range(3):
            d = X[a] - c[a]
            d -= L * np.round(d / L)
            d2 = d2 + d * d
        phi = np.ascontiguousarray(1.0 / np.sqrt(d2 + 1.0))
        return pos, phi, ex, ey, ez, k2

    def forces(c, gamma=0.0, nonpolar=0):
        pos, phi, ex, ey, ez, k2 = system(c)
        out = [np.zeros(3) for _ in range(3)]
        oracle_raw.or_pb_forces(n, h, w, eps_s, eps_int, nonpolar, gamma, 1, pos.reshape(-1),
                                np.array([r]), phi, ex, ey, ez, k2, *out)
        return out

    db, ib, npf = forces([6.0, 6.0, 6.0], 0.3, 1)
    scale = np.abs(forces([6.2, 5.9, 6.1], 0.3, 1)[0]).max()
    assert scale > 0
    # Face terms cancel pairwise only up to the face offsets' rounding.
    assert np.abs(db).max() <= 1e-12 * scale and np.abs(ib).max() <= 1e-12 * scale
    a = forces([6.2, 5.9, 6.1], 0.3, 1)
    b = forces([6.2 + 3 * h, 5.9 - 2 * h, 6.1 + h], 0.3, 1)
    for u, v in zip(a, b):
        assert np.allclose(u, v, rtol=1e-9, atol=1e-15 * scale)
    c = forces([6.2, 5.9, 6.1], 0.6, 1)
    assert np.allclose(c[2], 2 * a[2], rtol=1e-12) and np.abs(a[2]).max() > 0
    assert np.array_equal(forces([6.2, 5.9, 6.1], 0.3, 0)[2], np.zeros(3))


def block_laplacian():
    """The 7-point Laplacian of one 10^3 block with zero values outside it, rows in
    the reference's in-block order ii * 100 + jj * 10 + kk
    (library/src/precond_blockjacobi.c:14-69), built here independently."""
    idx = np.arange(1000)
    ii, jj, kk = idx // 100, (idx // 10) % 10, idx % 10
    a = np.zeros((1000, 1000))
    a[idx, idx] = -6.0
    for c, stride in ((ii, 100), (jj, 10), (kk, 1)):
        m = c > 0
        a[idx[m], idx[m] - stride] = 1.0
        m = c < 9
        a[idx[m], idx[m] + stride] = 1.0
    return a


def test_blockjacobi_solves_each_block(oracle_raw):
    """The oracle's block-Jacobi preconditioner against numpy's dense solve on every
    10^3 block, and A_B v = r back on each block: a dropped neighbour, a wrong
    block offset or a transposed index fails both."""
    n = 20
    r = np.ascontiguousarray(np.random.default_rng(5).standard_normal((n, n, n)))
    v = np.zeros_like(r)
    oracle_raw.or_blockjacobi_apply(r, v, n)
    a = block_laplacian()
    for bi in range(2):
        for bj in range(2):
            for bk in range(2):
                sl = (slice(10 * bi, 10 * bi + 10), slice(10 * bj, 10 * bj + 10),
                      slice(10 * bk, 10 * bk + 10))
                want = np.linalg.solve(a, r[sl].reshape(-1))
                assert np.abs(v[sl].reshape(-1) - want).max() <= 1e-13 * np.abs(want).max()
                assert np.abs(a @ v[sl].reshape(-1) - r[sl].reshape(-1)).max() <= 1e-12


def test_blockjacobi_pcg_matches_fft_solve(oracle_raw):
    """The block-Jacobi preconditioned CG converges to the periodic Poisson solution."""
    n = 20
    rhs = np.random.default_rng(6).standard_normal((n, n, n))
    rhs = np.ascontiguousarray(rhs - rhs.mean())
    x = np.zeros_like(rhs)
    it = oracle_raw.or_conj_grad_bj(rhs, x, 1e-10, n)
    assert 0 <= it < 200
    l = lam(n)
    l[0, 0, 0] = 1.0
    ref = np.real(np.fft.ifftn(np.fft.fftn(rhs) / l))
    ref -= ref.mean()
    assert np.abs((x - x.mean()) - ref).max() <= 1e-8 * np.abs(ref).max()

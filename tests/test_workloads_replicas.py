"""Host side of the many-replica workloads (no GPU): the uneven-size ensemble and the
serving trace generators, and the engine wrapper's refusal of a library without
the engine."""
import numpy as np
import pytest

from maze_poisson_amd import workloads as W
from maze_poisson_amd.engine import ReplicaEngine


def test_ensemble_systems_seeded_and_sized():
    a = W.ensemble_systems(12)
    b = W.ensemble_systems(12)
    for x, y in zip(a, b):
        assert x.n_p == y.n_p and x.n_grid == y.n_grid and np.array_equal(x.pos, y.pos)
    for s in a:
        assert s.n_p in W.ENSEMBLE_SIZES and s.n_grid % 2 == 0
        rho = (s.n_p / 2) * (W.M_NA + W.M_CL) * W.AMU_G / (s.L * W.A0 * 1e-8) ** 3
        assert 1.30 <= rho <= 1.40
        assert abs(s.L * W.A0 / s.n_grid - 0.41) <= 0.41 / s.n_grid * 1.01 + 1e-3
        assert np.all((s.pos >= 0) & (s.pos < s.L))
        assert abs(s.charges.sum()) == 0


def test_serving_trace_ranges():
    tr = W.serving_trace(40, rate_hz=10.0, seed=5)
    t = np.array([r.arrival for r in tr])
    assert np.all(np.diff(t) > 0) and t[0] > 0
    for r in tr:
        assert 216 <= r.system.n_p <= 32768 and r.system.n_p % 2 == 0
        assert 100 <= r.steps <= 2000
        assert r.system.n_grid % 2 == 0
    # Mean inter-arrival close to 1 / rate (40 samples: within 50 %).
    assert 0.05 <= np.diff(np.concatenate([[0.0], t])).mean() <= 0.15


def test_engine_needs_the_engine_symbols(oracle):
    with pytest.raises(RuntimeError):
        ReplicaEngine(oracle)

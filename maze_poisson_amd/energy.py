"""Energy bookkeeping and NVE drift of an MD run over the C-ABI.

E_tot = K + V_TF + E_elec per step, with E_elec both ways:
  * ``work``: the reference's running work of the electrostatic forces,
    E_work(n+1) = E_work(n) - sum_i 1/2 (F_i(n) + F_i(n+1)) . (x_i(n+1) - x_i(n))
    (minimum-image displacement), ``maze_poisson/cli/plotters/plot_T_E_tot.py:91-101``;
  * ``qphi``: 1/2 sum_g q_g phi_g on the grid (extension ``get_energy_elec_qphi``; the
    reference's ``get_energy_elec`` returns 1/2 sum phi, ``library/src/grid_general.c:304-316``).
Drift statistics follow ``maze_poisson/cli/plotters/plot_T_E_tot.py:109-110`` (std/|mean|)
plus the least-squares slope per ps.

    python -m maze_poisson_amd.energy --ions 4096 --grid 128 --steps 1000
"""
from __future__ import annotations

import argparse
import json

import numpy as np

from . import driver, workloads as W


def run_energy(lib, s: W.System, cfg: driver.RunConfig, steps: int, qphi: bool = True,
               precision: int | None = None) -> dict:
    """Bootstrap and run `steps` MD steps, recording the energy terms after each."""
    driver.initialize(lib, s, cfg)
    if precision is not None:
        lib.lib.solver_set_precision(precision)
    driver.bootstrap(lib)
    n = s.n_p
    L = s.L
    pos0, f0 = np.zeros((n, 3)), np.zeros((n, 3))
    lib.get_pos(pos0)
    lib.get_fcs_elec(f0)
    K, V, U, Wk = [], [], [], []
    work = 0.0
    pos1, f1 = np.zeros((n, 3)), np.zeros((n, 3))
    for _ in range(steps):
        _, v = driver.md_step(lib)
        lib.get_pos(pos1)
        lib.get_fcs_elec(f1)
        dx = pos1 - pos0
        dx -= L * np.round(dx / L)
        work -= float(np.sum(0.5 * (f0 + f1) * dx))
        pos0[:], f0[:] = pos1, f1
        K.append(lib.get_kinetic_energy())
        V.append(v)
        Wk.append(work)
        if qphi:
            U.append(lib.lib.get_energy_elec_qphi())
    lib.solver_finalize()
    out = {"K": np.array(K), "V_TF": np.array(V), "E_work": np.array(Wk)}
    if qphi:
        out["U_qphi"] = np.array(U)
    return out


def drift_stats(e: np.ndarray, dt_fs: float) -> dict:
    """std/|mean| and the linear-fit slope per ps (relative to |mean|) of a series."""
    t_ps = np.arange(len(e)) * dt_fs * 1e-3
    slope = float(np.polyfit(t_ps, e, 1)[0]) if len(e) > 1 else 0.0
    m = float(np.abs(e.mean()))
    return {"std_over_mean": float(e.std() / m), "slope_per_ps_over_mean": slope / m,
            "mean": float(e.mean())}


def summarize(rec: dict, dt_fs: float) -> dict:
    out = {"work": drift_stats(rec["K"] + rec["V_TF"] + rec["E_work"], dt_fs)}
    if "U_qphi" in rec:
        out["qphi"] = drift_stats(rec["K"] + rec["V_TF"] + rec["U_qphi"], dt_fs)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--ions", type=int, default=4096)
    ap.add_argument("--grid", type=int, default=128)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--steps", type=int, default=1000)
    ap.add_argument("--precision", type=int, default=2)
    ap.add_argument("--lib", default=None, help="another library exporting the ABI")
    args = ap.parse_args()
    from . import load_library
    from .capi import MazeLib
    lib = MazeLib(args.lib, extensions=False) if args.lib else load_library()
    s = W.nacl_system(args.ions, args.grid, seed=args.seed)
    cfg = driver.RunConfig(tol=W.poisson_tol(s))
    rec = run_energy(lib, s, cfg, args.steps, qphi=args.lib is None,
                     precision=None if args.lib else args.precision)
    print(json.dumps({"ions": args.ions, "grid": args.grid, "steps": args.steps,
                      "precision": None if args.lib else args.precision,
                      "drift": summarize(rec, cfg.dt_fs)}))


if __name__ == "__main__":
    main()

"""MD driver over the MaZe C-ABI.

Makes the reference driver's call sequence (``maze_poisson/solver.py:97-510``): the
initialisation (``initialize`` :97-108), the two bootstrap steps with cold field
solves (``initialize_md`` :371-409), and the MD loop (``md_loop_iter`` :483-492),
with electrostatic and non-electrostatic forces enabled.  It works unchanged on any
library exporting the ABI (:class:`maze_poisson_amd.capi.MazeLib`).
"""
from __future__ import annotations

from dataclasses import dataclass

import numpy as np

from .capi import MazeLib
from .workloads import A0, System, dt_au, kbar2, pb_eps_field


@dataclass
class RunConfig:
    method: str = "MAZE-LCG"
    precond: str = "NONE"
    tol: float = 1e-7
    eps_s: float = 1.0
    eps_int: float = 1.0
    dt_fs: float = 0.25
    T: float = 1500.0
    gamma: float = 1e-3
    integrator: str = "VERLET"
    thermostat: bool = False
    cas: str = "CIC"          # charge assignment: CIC, SPL_QUADR or SPL_CUBIC
    # Linearized Poisson-Boltzmann: synthetic face dielectric in [2, 80] (seeded) and a
    # uniform screening term for `ionic_molar` at eps_s (= 80 in water).
    pb: bool = False
    ionic_molar: float = 0.1
    pb_eps_s: float = 80.0
    eps_seed: int = 0
    # Radius-based PB boundary (the reference's own maps, rebuilt every step, and the
    # dielectric / ionic / non-polar boundary forces): transition width pb_w (Bohr),
    # solvation radius per type plus the probe radius (A), non-polar surface tension
    # gamma_np (kcal/mol/A^2).  Replaces the synthetic dielectric of `pb`.
    pb_radii: bool = False
    pb_w: float = 1.0
    radii_ang: tuple = (0.95, 1.81)
    probe_ang: float = 1.4
    nonpolar: bool = False
    gamma_np: float = 0.0
    beta_np: float = 0.0


def initialize(lib: MazeLib, s: System, cfg: RunConfig) -> None:
    """``maze_poisson/solver.py:97-106,129-144,276-346,358-369``"""
    maps = lib.type_maps()
    lib.solver_initialize()
    lib.solver_initialize_grid(s.n_grid, s.L, s.h, cfg.tol, cfg.eps_s, cfg.eps_int,
                               maps["grid"][cfg.method.upper()],
                               maps["precond"][cfg.precond.upper()])
    if cfg.pb_radii:
        kb2 = kbar2(cfg.ionic_molar, cfg.eps_s, cfg.T, s.h)
        lib.solver_initialize_grid_pois_boltz(cfg.pb_w, kb2, 1 if cfg.nonpolar else 0)
    elif cfg.pb:
        kb2 = kbar2(cfg.ionic_molar, cfg.pb_eps_s, cfg.T, s.h)
        lib.solver_initialize_grid_pois_boltz(0.0, kb2, 0)
        ex, ey, ez = pb_eps_field(s.n_grid, s.L, cfg.eps_seed)
        lib.solver_set_eps_k2(ex, ey, ez, np.full((s.n_grid,) * 3, kb2))
    lib.solver_initialize_particles(s.n_grid, s.n_typ, s.L, s.h, s.n_p,
                                    maps["potential"]["TF"], maps["cas"][cfg.cas.upper()],
                                    s.types, s.pos, s.vel, s.mass, s.charges, s.tf)
    if cfg.pb_radii:
        lib.solver_initialize_particles_pois_boltz(cfg.gamma_np * GAMMA_NP_AU, cfg.beta_np,
                                                   pb_radii(s, cfg))
    lib.solver_initialize_integrator(s.n_p, dt_au(cfg.dt_fs), cfg.T, cfg.gamma,
                                     maps["integrator"][cfg.integrator.upper()],
                                     1 if cfg.thermostat else 0)


GAMMA_NP_AU = 0.0065934  # kcal/mol/A^2 -> a.u., maze_poisson/myio/input/base_file_input.py:147


def pb_radii(s: System, cfg: RunConfig) -> np.ndarray:
    """Per-ion PB radius in Bohr: the type's solvation radius plus the probe radius
    (``maze_poisson/solver.py:348-352``)."""
    r = np.array(cfg.radii_ang, dtype=np.float64)[s.types] / A0 + cfg.probe_ang / A0
    return np.ascontiguousarray(r)


def _radii(cfg) -> bool:
    return cfg is not None and cfg.pb_radii


def compute_forces(lib: MazeLib, cfg: RunConfig | None = None) -> float:
    """``maze_poisson/solver.py:421-431``; returns the TF potential energy."""
    lib.solver_compute_forces_elec()
    v = lib.solver_compute_forces_noel()
    if _radii(cfg):
        lib.solver_compute_forces_pb()
    lib.solver_compute_forces_tot()
    return v


def update_charges(lib: MazeLib, cfg: RunConfig | None = None) -> None:
    """The charge spread and, with radius-based PB maps, their rebuild
    (``maze_poisson/solver.py:378-380``, :411-415)."""
    lib.solver_update_charges()
    if _radii(cfg):
        lib.solver_update_eps_k2()


def bootstrap(lib: MazeLib, cfg: RunConfig | None = None) -> None:
    """The two Verlet start-up steps with cold solves, ``maze_poisson/solver.py:376-398``."""
    update_charges(lib, cfg)
    lib.solver_init_field()
    compute_forces(lib, cfg)
    lib.integrator_part_1()
    update_charges(lib, cfg)
    lib.solver_init_field()
    compute_forces(lib, cfg)
    lib.integrator_part_2()


def md_step(lib: MazeLib, cfg: RunConfig | None = None) -> tuple[int, float]:
    """One MD step, ``maze_poisson/solver.py:475-483``; returns (CG result, V_TF)."""
    lib.integrator_part_1()
    update_charges(lib, cfg)
    it = lib.solver_update_field()
    v = compute_forces(lib, cfg)
    lib.integrator_part_2()
    return it, v


def snapshot(lib: MazeLib, s: System, fields: bool = True) -> dict:
    """Copy the observable state out through the getters."""
    n, N = s.n_grid, s.n_p
    out = {}
    for name in ("pos", "vel", "fcs_elec", "fcs_noel", "fcs_tot", "fcs_db", "fcs_ib", "fcs_np"):
        a = np.zeros((N, 3))
        getattr(lib, "get_" + name)(a)
        out[name] = a
    if fields:
        for name in ("field", "field_prev", "q"):
            a = np.zeros((n, n, n))
            getattr(lib, "get_" + name)(a)
            out[name] = a
    out["kinetic"] = lib.get_kinetic_energy()
    out["temperature"] = lib.get_temperature()
    out["energy_elec"] = lib.get_energy_elec()
    m = np.zeros(3)
    lib.get_momentum(m)
    out["momentum"] = m
    return out

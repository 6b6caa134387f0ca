"""Many independent MD replicas on one GPU through the library's replica engine.

The engine (``csrc/engine.hip``, C-ABI ``maze_engine_*``) holds each replica as a
full simulation state on the device and advances all admitted replicas together
from one host thread: a replica's next step is queued as soon as its previous
field solve has published its state, so several replicas' kernels run
concurrently.  Requests are admitted in arrival order by free slot and
device-memory budget (continuous batching: a finished replica's memory admits the
next request while the others keep running).  Each replica runs the reference
driver's sequence (``maze_poisson/solver.py:376-398`` start-up, then
``md_loop_iter`` :483-492), by default with MAZE-LCG, velocity Verlet, CIC and
Tosi-Fumi forces; a :class:`~maze_poisson_amd.driver.RunConfig` given to
:meth:`ReplicaEngine.submit` selects the method and preconditioner, the charge
assignment, the OVRVO integrator and thermostat, and linear Poisson-Boltzmann on
synthetic maps or on maps rebuilt from the ion radii, as the single-replica path
takes them.  A request's trajectory equals the single-replica C-ABI path's, alone or
in any mix.
"""
from __future__ import annotations

import numpy as np

from . import driver
from .capi import MazeLib
from .workloads import System, dt_au, kbar2, pb_eps_field, poisson_tol

PRECISION = {"FP64": 0, "MIXED-CG": 1, "MIXED-MG": 2}

REQUEST_FIELDS = ("arrival_s", "admit_s", "done_s", "steps", "mean_ms", "p50_ms", "p99_ms",
                  "est_bytes", "meas_bytes", "precision", "failures", "arena_used",
                  "arena_overflow", "stream_bytes")
OBSERVABLE_FIELDS = ("kinetic", "temperature", "energy_elec", "potential_tf")
STATS_FIELDS = ("wall_s", "replica_steps_per_s", "atom_steps_per_s", "p50_ms", "p99_ms",
                "peak_bytes", "max_running", "host_admit_s", "host_issue_s", "host_finish_s")


class ReplicaEngine:
    """Owns one ``maze_engine``; requests are :class:`~maze_poisson_amd.workloads.System`
    replicas with a step count and an arrival time."""

    def __init__(self, lib: MazeLib, budget_bytes: int = 0, max_slots: int = 64,
                 threads: int | None = None, graphs: bool | None = None,
                 mixed_init: bool | None = None):
        if not lib.has("maze_engine_create"):
            raise RuntimeError(f"{lib.path} has no replica engine")
        self.lib = lib
        self.h = lib.maze_engine_create(int(budget_bytes), int(max_slots))
        if not self.h:
            raise RuntimeError("maze_engine_create failed")
        if threads is not None:
            lib.maze_engine_set_threads(self.h, int(threads))
        if graphs is not None:
            lib.maze_engine_set_graphs(self.h, 1 if graphs else 0)
        if mixed_init is not None:
            lib.maze_engine_set_mixed_init(self.h, 1 if mixed_init else 0)
        self.systems = {}

    def close(self):
        if self.h:
            self.lib.maze_engine_destroy(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def submit(self, s: System, steps: int, arrival: float = 0.0, tol: float | None = None,
               precision: int = 2, dt_fs: float = 0.25, cfg: driver.RunConfig | None = None,
               seed: int | None = None, rescale: bool = False, keep_field: bool = False) -> int:
        """Queue a replica.  Without ``cfg`` it runs the defaults above; with one, its
        method, preconditioner, eps_s / eps_int, charge assignment, integrator, T, gamma,
        thermostat, time step and PB settings, as :func:`driver.initialize` applies them
        (``tol`` then defaults to ``cfg.tol``).  ``seed``: the OVRVO noise stream;
        ``rescale``: velocities rescaled after the start-up steps."""
        if cfg is not None:
            tol = cfg.tol if tol is None else tol
            dt_fs = cfg.dt_fs
        tol = poisson_tol(s) if tol is None else tol
        rid = self.lib.maze_engine_submit(
            self.h, float(arrival), s.n_grid, s.L, tol, int(precision), s.n_p, s.n_typ,
            s.types, s.pos, s.vel, s.mass, s.charges, s.tf, dt_au(dt_fs), int(steps))
        self._check(rid)
        self.systems[rid] = s
        try:
            if cfg is not None or seed is not None or rescale:
                self._physics(rid, s, cfg or driver.RunConfig(tol=tol), seed, rescale)
            if keep_field:
                self._check(self.lib.maze_engine_keep_field(self.h, rid, 1))
        except (ValueError, MemoryError, KeyError):
            self.lib.maze_engine_cancel(self.h, rid)
            raise
        return rid

    @staticmethod
    def _check(rc: int) -> int:
        if rc == -2:
            raise MemoryError("request exceeds the engine's device-memory budget")
        if rc < 0:
            raise ValueError("invalid request")
        return rc

    def _physics(self, rid: int, s: System, cfg: driver.RunConfig, seed: int | None,
                 rescale: bool) -> None:
        maps = self.lib.type_maps()
        self._check(self.lib.maze_engine_set_physics(
            self.h, rid, maps["grid"][cfg.method.upper()], maps["precond"][cfg.precond.upper()],
            maps["cas"][cfg.cas.upper()], cfg.eps_s, cfg.eps_int,
            maps["integrator"][cfg.integrator.upper()], cfg.T, cfg.gamma,
            1 if cfg.thermostat else 0, 0 if seed is None else 1,
            0 if seed is None else int(seed), 1 if rescale else 0))
        if cfg.pb_radii:
            kb2 = kbar2(cfg.ionic_molar, cfg.eps_s, cfg.T, s.h)
            self._check(self.lib.maze_engine_set_pb_radii(
                self.h, rid, cfg.pb_w, kb2, 1 if cfg.nonpolar else 0,
                cfg.gamma_np * driver.GAMMA_NP_AU, cfg.beta_np, driver.pb_radii(s, cfg)))
        elif cfg.pb:
            kb2 = kbar2(cfg.ionic_molar, cfg.pb_eps_s, cfg.T, s.h)
            ex, ey, ez = pb_eps_field(s.n_grid, s.L, cfg.eps_seed)
            self._check(self.lib.maze_engine_set_pb_maps(
                self.h, rid, kb2, ex, ey, ez, np.full((s.n_grid,) * 3, kb2)))

    def run(self) -> dict:
        self.lib.maze_engine_run(self.h)
        return self.stats()

    def result(self, rid: int) -> dict:
        n = self.systems[rid].n_p
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "fcs_tot")}
        if self.lib.maze_engine_result(self.h, rid, out["pos"], out["vel"], out["fcs_tot"]) != 0:
            raise RuntimeError(f"request {rid} has not finished")
        return out

    def observables(self, rid: int) -> dict:
        """Kinetic energy, temperature, electrostatic energy, Tosi-Fumi potential
        energy and momentum after the request's last step."""
        v = np.zeros(7)
        if self.lib.maze_engine_observables(self.h, rid, v, 7) != 7:
            raise RuntimeError(f"request {rid} has not finished")
        out = dict(zip(OBSERVABLE_FIELDS, v[:4].tolist()))
        out["momentum"] = v[4:7].copy()
        return out

    def field(self, rid: int) -> np.ndarray:
        """The final field of a request submitted with ``keep_field``."""
        n = self.systems[rid].n_grid
        phi = np.zeros((n, n, n))
        if self.lib.maze_engine_field(self.h, rid, phi) != 0:
            raise RuntimeError(f"request {rid} has no kept field")
        return phi

    def request_stats(self, rid: int) -> dict:
        v = np.zeros(len(REQUEST_FIELDS))
        k = self.lib.maze_engine_request_stats(self.h, rid, v, len(v))
        return dict(zip(REQUEST_FIELDS[:k], v[:k].tolist()))

    def stats(self) -> dict:
        v = np.zeros(len(STATS_FIELDS))
        k = self.lib.maze_engine_stats(self.h, v, len(v))
        return dict(zip(STATS_FIELDS[:k], v[:k].tolist()))

    @staticmethod
    def estimate_bytes(lib: MazeLib, n_grid: int, n_p: int, precision: int = 2) -> int:
        return int(lib.maze_engine_estimate_bytes(n_grid, n_p, precision))

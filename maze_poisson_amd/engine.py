"""Many independent MD replicas on one GPU through the library's replica engine.

The engine (``csrc/engine.hip``, C-ABI ``maze_engine_*``) holds each replica as a
full simulation state on the device and advances all admitted replicas together
from one host thread: a replica's next step is queued as soon as its previous
field solve has published its state, so several replicas' kernels run
concurrently.  Requests are admitted in arrival order by free slot and
device-memory budget (continuous batching: a finished replica's memory admits the
next request while the others keep running).  Each replica runs the reference
driver's sequence (``maze_poisson/solver.py:376-398`` start-up, then
``md_loop_iter`` :483-492) with velocity Verlet, CIC and Tosi-Fumi forces, so a
request's trajectory equals the single-replica C-ABI path's, alone or in any mix.
"""
from __future__ import annotations

import numpy as np

from .capi import MazeLib
from .workloads import System, dt_au, poisson_tol

PRECISION = {"FP64": 0, "MIXED-CG": 1, "MIXED-MG": 2}

REQUEST_FIELDS = ("arrival_s", "admit_s", "done_s", "steps", "mean_ms", "p50_ms", "p99_ms",
                  "est_bytes", "meas_bytes", "precision", "failures", "arena_used",
                  "arena_overflow", "stream_bytes")
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
               precision: int = 2, dt_fs: float = 0.25) -> int:
        tol = poisson_tol(s) if tol is None else tol
        rid = self.lib.maze_engine_submit(
            self.h, float(arrival), s.n_grid, s.L, tol, int(precision), s.n_p, s.n_typ,
            s.types, s.pos, s.vel, s.mass, s.charges, s.tf, dt_au(dt_fs), int(steps))
        if rid == -2:
            raise MemoryError("request exceeds the engine's device-memory budget")
        if rid < 0:
            raise ValueError("invalid request")
        self.systems[rid] = s
        return rid

    def run(self) -> dict:
        self.lib.maze_engine_run(self.h)
        return self.stats()

    def result(self, rid: int) -> dict:
        n = self.systems[rid].n_p
        out = {k: np.zeros((n, 3)) for k in ("pos", "vel", "fcs_tot")}
        if self.lib.maze_engine_result(self.h, rid, out["pos"], out["vel"], out["fcs_tot"]) != 0:
            raise RuntimeError(f"request {rid} has not finished")
        return out

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

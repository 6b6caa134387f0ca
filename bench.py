#!/usr/bin/env python3
"""MD steps/s of the MaZe electrostatics + MD hot path on one MI355X.

Workload: 4096 ions (2048 Na+, 2048 Cl-) in a cubic periodic box of 52.43 A
(NaCl melt at 1.379 g/cm^3), 128^3 grid, CIC, MAZE-LCG (Jacobi CG, tolerance
1e-8 of the right-hand-side scale), shifted-force Tosi-Fumi with cutoff L/2,
velocity Verlet NVE, dt = 0.25 fs, synthetic seeded input.

One "step" is one call of the reference's MD loop body (integrator part 1,
charge spreading, MaZe field solve, electrostatic + Tosi-Fumi forces,
integrator part 2), with all state resident in HBM.  The headline path is the
FP32-storage CG inside FP64 iterative refinement; the plain FP64 path is timed
beside it.  Prints one JSON line.
"""
from __future__ import annotations

import argparse
import ctypes
import json
import os
import subprocess
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

from maze_poisson_amd import driver, workloads as W  # noqa: E402

N_IONS, N_GRID, SEED = 4096, 128, 2
HBM_PEAK_GBS = 8000.0


def make_system():
    s = W.nacl_system(N_IONS, N_GRID, seed=SEED)
    return s, driver.RunConfig(method="MAZE-LCG", tol=W.poisson_tol(s), integrator="VERLET")


def bind_ext(lib):
    L = lib.lib
    L.solver_set_precision.argtypes = [ctypes.c_int]
    L.solver_set_precision.restype = ctypes.c_int
    L.maze_prof_enable.argtypes = [ctypes.c_int]
    L.maze_prof_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_long)]
    L.maze_prof_get.restype = ctypes.c_int
    L.solver_md_loop_iter.argtypes = []
    L.solver_sync.argtypes = []
    return L


def prof_get(L, name):
    t, c = ctypes.c_double(0), ctypes.c_long(0)
    if L.maze_prof_get(name.encode(), ctypes.byref(t), ctypes.byref(c)) != 0:
        return 0.0, 0
    return t.value, c.value


def run_path(lib, precision, steps, warmup, profile):
    """Bootstrap, warm up, then time `steps` MD steps one by one (device-synchronised
    per step for the latency distribution)."""
    L = bind_ext(lib)
    s, cfg = make_system()
    driver.initialize(lib, s, cfg)
    L.solver_set_precision(precision)
    driver.bootstrap(lib)
    for _ in range(warmup):
        L.solver_md_loop_iter()
    L.solver_sync()
    L.maze_prof_reset()
    L.maze_prof_enable(1 if profile else 0)
    lat = []
    t0 = time.perf_counter()
    for _ in range(steps):
        a = time.perf_counter()
        L.solver_md_loop_iter()
        L.solver_sync()
        lat.append(time.perf_counter() - a)
    L.solver_sync()
    elapsed = time.perf_counter() - t0
    L.maze_prof_enable(0)
    prof = {k: prof_get(L, k) for k in ("cg_dir", "cg_update", "cg_dir_f32", "cg_update_f32")}
    f = np.zeros((s.n_p, 3))
    lib.get_fcs_tot(f)
    lib.solver_finalize()
    lat = np.array(lat) * 1e3
    return {
        "steps_per_s": steps / elapsed,
        "ms_per_step": elapsed / steps * 1e3,
        "p50_ms": float(np.percentile(lat, 50)),
        "p99_ms": float(np.percentile(lat, 99)),
        "prof": prof,
        "iters_per_step": (prof["cg_update"][1] + prof["cg_update_f32"][1]) / max(steps, 1),
        "forces": f,
    }


def force_error(lib, nsteps=5):
    """Force RMS error of the mixed path relative to the FP64 path after the same
    bootstrap (FP64) and `nsteps` MD steps."""
    out = []
    for prec in (0, 1):
        L = bind_ext(lib)
        s, cfg = make_system()
        driver.initialize(lib, s, cfg)
        L.solver_set_precision(prec)
        driver.bootstrap(lib)
        for _ in range(nsteps):
            L.solver_md_loop_iter()
        f = np.zeros((s.n_p, 3))
        lib.get_fcs_tot(f)
        lib.solver_finalize()
        out.append(f)
    d = out[1] - out[0]
    return float(np.sqrt((d ** 2).mean()) / np.sqrt((out[0] ** 2).mean()))


def cpu_baseline_child(steps):
    """Time the reference C (compiled from its own sources with its own flags, OpenMP)
    on the host: bootstrap, then `steps` MD steps of the same workload."""
    from maze_poisson_amd.capi import MazeLib
    path = os.path.join(ROOT, "oracle", "_ref", "libmaze_ref_fast.so")
    lib = MazeLib(path, extensions=False)
    s, cfg = make_system()
    driver.initialize(lib, s, cfg)
    driver.bootstrap(lib)
    driver.md_step(lib)
    t0 = time.perf_counter()
    for _ in range(steps):
        driver.md_step(lib)
    el = time.perf_counter() - t0
    print(json.dumps({"steps_per_s": steps / el, "elapsed": el}))


def cpu_baseline(steps, cores):
    path = os.path.join(ROOT, "oracle", "_ref", "libmaze_ref_fast.so")
    if not os.path.exists(path):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child",
                              "--cpu-steps", str(steps)], env=env, capture_output=True,
                             timeout=240, text=True)
        line = [ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1]
        r = json.loads(line)
    except Exception as e:  # the baseline is reported, never required
        sys.stderr.write(f"cpu baseline failed: {e}\n")
        return None
    return {"value": r["steps_per_s"], "unit": "MD steps/s", "cores": cores, "kind": "reference",
            "sample": f"reference C (-O3 -ffast-math -mavx2, OpenMP x{cores}), same 4096-ion / "
                      f"128^3 workload, {steps} MD steps after the 2 bootstrap steps"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=100)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--fp64-steps", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-steps", type=int, default=15)
    ap.add_argument("--cpu-cores", type=int, default=16)
    ap.add_argument("--cpu-child", action="store_true")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_baseline_child(args.cpu_steps)
        return
    if args.gpus != 1:
        print("bench.py: this build runs one replica on one GPU; multi-GPU runs are not "
              "part of it (independent replicas would shard one per GPU).", file=sys.stderr)
        sys.exit(2)

    from maze_poisson_amd import load_library
    lib = load_library()
    fp64 = run_path(lib, 0, args.fp64_steps or max(10, args.steps // 2), args.warmup, True)
    mixed = run_path(lib, 1, args.steps, args.warmup, True)
    err = force_error(lib)

    G = N_GRID ** 3
    t_ms, cnt = mixed["prof"]["cg_update_f32"]
    upd_bytes = 5 * G * 4  # reads p, x, r; writes x, r (FP32)
    avg_ms = t_ms / max(cnt, 1)
    achieved = upd_bytes / (avg_ms * 1e-3) / 1e9 if cnt else 0.0
    roofline = {"bound": "hbm", "kernel": "k_cg_update<float>", "achieved": round(achieved, 1),
                "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                "traffic": None, "avg_us": round(avg_ms * 1e3, 2), "launches": cnt,
                "bytes_per_launch": upd_bytes}
    t64, c64 = fp64["prof"]["cg_update"]
    a64 = 5 * G * 8 / (t64 / max(c64, 1) * 1e-3) / 1e9 if c64 else 0.0

    cpu = None if args.no_cpu else cpu_baseline(args.cpu_steps, args.cpu_cores)
    line = {
        "metric": "MD steps/s at 4096 ions/128^3 grid, FP32+refine vs FP64; p50/p99 step latency",
        "value": round(mixed["steps_per_s"], 3),
        "unit": "MD steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(mixed["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "4096 ions / 128^3 grid, MAZE-LCG CG tol 1e-8 rel, TF L/2, "
                               "velocity Verlet NVE dt 0.25 fs",
                   "ions": N_IONS, "grid": N_GRID, "seed": SEED},
        "p50_ms": round(mixed["p50_ms"], 4),
        "p99_ms": round(mixed["p99_ms"], 4),
        "fp64_steps_per_s": round(fp64["steps_per_s"], 3),
        "fp64_p50_ms": round(fp64["p50_ms"], 4),
        "fp64_p99_ms": round(fp64["p99_ms"], 4),
        "speedup_vs_fp64": round(mixed["steps_per_s"] / fp64["steps_per_s"], 3),
        "force_rms_rel_err_vs_fp64": err,
        "cg_iters_per_step": {"mixed_inner": mixed["iters_per_step"],
                              "fp64": fp64["iters_per_step"]},
        "fp64_cg_update_gbs": round(a64, 1),
        "roofline": roofline,
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(mixed["steps_per_s"] / cpu["value"], 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

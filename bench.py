#!/usr/bin/env python3
"""MD steps/s of the MaZe electrostatics + MD hot path on one MI355X.

Workload: 4096 ions (2048 Na+, 2048 Cl-) in a cubic periodic box of 52.43 A
(NaCl melt at 1.379 g/cm^3, 1500 K), 128^3 grid, CIC, MAZE-LCG with the
reference's infinity-norm stop at 1e-8 of the right-hand-side scale,
shifted-force Tosi-Fumi with cutoff L/2, velocity Verlet NVE, dt = 0.25 fs,
synthetic seeded input.

One "step" is one call of the reference's MD loop body (integrator part 1,
charge spreading, MaZe field solve, electrostatic + Tosi-Fumi forces,
integrator part 2), all state resident in HBM.  Three field-solve paths:
  fp64      the reference algorithm (Jacobi CG) in FP64,
  mixed_mg  FP32 multigrid-preconditioned CG inside FP64 refinement (headline),
  mixed_cg  FP32 Jacobi CG inside FP64 refinement,
  fp64_mg   FP64 CG preconditioned by the headline's V-cycle carried out in FP64
            (the headline's algorithm without its FP32 storage: mixed_mg against
            fp64_mg is the precision's gain, fp64_mg against fp64 the algorithm's).
Prints one JSON line.
"""
from __future__ import annotations

import argparse
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
PATHS = {"fp64": 0, "mixed_cg": 1, "mixed_mg": 2, "fp64_mg": 3}

# Algorithmic HBM bytes per grid point of each timed kernel (arrays read + written
# once each; FP64 = 8 B, FP32 = 4 B).
BYTES_PER_POINT = {
    "cg_dir": 3 * 8, "cg_update": 5 * 8,                  # FP64 Jacobi CG
    "cg_dir_f32": 3 * 4, "cg_update_f32": 5 * 4,          # FP32 CG / PCG direction, update
    "mg_down_l0": 4 + 4 + 0.5,                            # finest multigrid level: b -> d2, b_H
    "mg_up_l0": 4 + 4 + 0.5 + 4,                          # b, d2, e -> d
    "refine": 8 + 4 + 8 + 8 + 4,                          # y, d, sigma -> y', b32
    "pred_sigma": 5 * 8,                                  # phi, phi_p, q -> phi*, sigma
}


# The plain reference-method case (512 ions, 64^3 grid, seed 1): its FP64 GPU and CPU rows.
C1_IONS, C1_GRID, C1_SEED = 512, 64, 1


def make_system(n_ions=N_IONS, n_grid=N_GRID, seed=SEED):
    s = W.nacl_system(n_ions, n_grid, seed=seed)
    return s, W.poisson_tol(s)


def prof_get(L, name):
    import ctypes
    t, c = ctypes.c_double(0), ctypes.c_long(0)
    L.maze_prof_get.argtypes = [ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                ctypes.POINTER(ctypes.c_long)]
    L.maze_prof_get.restype = ctypes.c_int
    if L.maze_prof_get(name.encode(), ctypes.byref(t), ctypes.byref(c)) != 0:
        return 0.0, 0
    return t.value, c.value


def start(lib, precision, tol_factor=1.0, system=None):
    s, tol = make_system(*(system or ()))
    cfg = driver.RunConfig(method="MAZE-LCG", tol=tol * tol_factor, integrator="VERLET")
    driver.initialize(lib, s, cfg)
    lib.solver_set_precision(precision)
    driver.bootstrap(lib)
    return s


def run_path(lib, precision, steps, warmup, lat_steps, prof_steps, system=None):
    """Bootstrap and warm up, then three phases on the same trajectory:
    1. `steps` MD steps back to back between two device synchronisations (throughput);
    2. `lat_steps` steps each followed by a synchronisation (step latency p50/p99);
    3. `prof_steps` steps with HIP events around the timed kernels (per-kernel
       average durations; the events add gaps, so this phase is not timed)."""
    L = lib.lib
    start(lib, precision, system=system)
    for _ in range(warmup):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    elapsed = time.perf_counter() - t0
    lat = []
    for _ in range(lat_steps):
        a = time.perf_counter()
        lib.solver_md_loop_iter()
        lib.solver_sync()
        lat.append(time.perf_counter() - a)
    L.maze_prof_reset()
    L.maze_prof_enable(1 if prof_steps else 0)
    for _ in range(prof_steps):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    L.maze_prof_enable(0)
    prof = {k: prof_get(L, k) for k in BYTES_PER_POINT}
    lib.solver_finalize()
    lat = np.array(lat) * 1e3 if lat else np.zeros(1)
    return {"steps_per_s": steps / elapsed, "ms_per_step": elapsed / steps * 1e3,
            "p50_ms": float(np.percentile(lat, 50)), "p99_ms": float(np.percentile(lat, 99)),
            "prof": prof}


def driver_rate(lib, precision, steps, getters):
    """Steps/s through the reference driver's own per-step call sequence (about 8 C-ABI
    calls, maze_poisson/solver.py:483-492), optionally copying positions, velocities and
    total forces to the host after every step (PCIe-inclusive, as a trajectory writer
    would)."""
    s = start(lib, precision)
    buf = np.zeros((s.n_p, 3))
    for _ in range(5):
        driver.md_step(lib)
    lib.solver_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        driver.md_step(lib)
        if getters:
            lib.get_pos(buf)
            lib.get_vel(buf)
            lib.get_fcs_tot(buf)
    lib.solver_sync()
    el = time.perf_counter() - t0
    lib.solver_finalize()
    return steps / el


# Radius-based PB boundary on the 8000-ion / 160^3 geometry (dense melt, radii
# Na 0.95 A / Cl 1.81 A + 1.4 A probe, w = 1 Bohr, eps 80 / 1, 0.1 M).
PB_IONS, PB_GRID, PB_SEED = 8000, 160, 3


def pb_rate(lib, cfg, s, precision, steps):
    """MD steps/s of a PB configuration (two warm-up steps after the start-up)."""
    driver.initialize(lib, s, cfg)
    lib.solver_set_precision(precision)
    driver.bootstrap(lib, cfg)
    for _ in range(2):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    el = time.perf_counter() - t0
    lib.solver_finalize()
    return steps / el


def pb_synthetic(lib, steps):
    """OVRVO Langevin (1500 K, 1 / ps) with linearized PB on a synthetic dielectric in
    [2, 80] and 0.1 M screening, 8000 ions / 160^3: the FP32 PB multigrid-PCG in FP64
    refinement against FP64 CG."""
    s = W.nacl_system(PB_IONS, PB_GRID, seed=PB_SEED)
    cfg = driver.RunConfig(method="MAZE-LCG", tol=W.poisson_tol(s), integrator="OVRVO",
                           thermostat=True, gamma=2.4189e-5, pb=True)
    mg = pb_rate(lib, cfg, s, PATHS["mixed_mg"], steps)
    f64 = pb_rate(lib, cfg, s, PATHS["fp64"], max(3, min(10, steps // 10)))
    return {"workload": f"{PB_IONS}-ion NaCl melt, {PB_GRID}^3 grid, OVRVO 1500 K, linear PB "
                        "with a synthetic dielectric in [2, 80] (8 Gaussians) and 0.1 M screening",
            "mixed_mg_steps_per_s": round(mg, 3), "fp64_steps_per_s": round(f64, 3),
            "speedup_vs_fp64": round(mg / f64, 2)}


def pb_radii_rate(lib, steps, prof_steps):
    """MD steps/s with the PB maps rebuilt from the radii every step and the boundary
    forces in the total (field solve: FP32 PB multigrid-PCG inside FP64 refinement), and
    the average durations of the map and force kernels."""
    L = lib.lib
    s = W.nacl_system(PB_IONS, PB_GRID, seed=PB_SEED)
    cfg = driver.RunConfig(method="MAZE-LCG", tol=W.poisson_tol(s), integrator="VERLET",
                           eps_s=80.0, pb_radii=True, pb_w=1.0)
    driver.initialize(lib, s, cfg)
    lib.solver_set_precision(PATHS["mixed_mg"])
    driver.bootstrap(lib, cfg)
    for _ in range(3):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    t0 = time.perf_counter()
    for _ in range(steps):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    el = time.perf_counter() - t0
    L.maze_prof_reset()
    L.maze_prof_enable(1)
    for _ in range(prof_steps):
        lib.solver_md_loop_iter()
    lib.solver_sync()
    L.maze_prof_enable(0)
    prof = {k: prof_get(L, k) for k in ("pb_maps", "pb_forces")}
    lib.solver_finalize()
    out = {"workload": f"{PB_IONS}-ion NaCl melt, {PB_GRID}^3 grid, PB maps from radii every "
                       "step + DB/IB/NP forces, MAZE-LCG FP32 PB multigrid-PCG in FP64 "
                       "refinement",
           "steps_per_s": round(steps / el, 3)}
    G = PB_GRID ** 3
    for k, (tot, cnt) in prof.items():
        if cnt:
            avg = tot / cnt * 1e-3
            out[k + "_avg_us"] = round(avg * 1e6, 2)
            if k == "pb_maps":  # writes eps_x, eps_y, eps_z, k2 (FP64) once per point
                gbs = 32.0 * G / avg / 1e9
                out["pb_maps_gbs"] = round(gbs, 1)
                out["pb_maps_frac_hbm"] = round(gbs / HBM_PEAK_GBS, 4)
    return out


def runner_rate(lib, steps):
    """Steps/s of a whole run through the input-file runner (YAML + CSV inputs, outputs
    energy / momentum / temperature / performance every 50 steps), headline workload."""
    import tempfile
    from maze_poisson_amd.convert import write_species, write_tosi_fumi
    from maze_poisson_amd.solver import run_file
    s, tol = make_system()
    with tempfile.TemporaryDirectory() as d:
        write_species(os.path.join(d, "species.csv"), [("Na", 1.0, W.M_NA), ("Cl", -1.0, W.M_CL)])
        write_tosi_fumi(os.path.join(d, "tosi_fumi.csv"))
        with open(os.path.join(d, "input.csv"), "w") as f:
            f.write("type,x,y,z,vx,vy,vz\n")
            for i in range(s.n_p):
                f.write(",".join([("Na", "Cl")[s.types[i]]] + [repr(float(v)) for v in s.pos[i] * W.A0]
                                 + [repr(float(v)) for v in s.vel[i]]) + "\n")
        yml = os.path.join(d, "run.yml")
        with open(yml, "w") as f:
            f.write(f"""grid_setting:
  N: {s.n_grid}
  N_p: {s.n_p}
  N_typs: 2
  L: {s.L!r}
  input_file: "{d}/input.csv"
  particles_file: "{d}/species.csv"
md_variables:
  N_steps: {steps}
  T: 1500
  dt_fs: 0.25
  tol: {tol!r}
  method: MAZE-LCG
  integrator: VERLET
  potential: TF
  potential_params_file: "{d}/tosi_fumi.csv"
  precision: MIXED-MG
output_settings:
  path: "{d}/out"
  print_energy: true
  print_momentum: true
  print_temperature: true
  print_performance: true
  stride: 50
  flushstride: 5
""")
        solver = run_file(yml, lib=lib)
        lib.solver_finalize()
    return solver.timers["steps_per_s"]


def forces_after(lib, precision, tol_factor, nsteps):
    s = start(lib, precision, tol_factor)
    its = [driver.md_step(lib)[0] for _ in range(nsteps)]
    f = np.zeros((s.n_p, 3))
    lib.get_fcs_tot(f)
    lib.solver_finalize()
    return f, its


def rms_rel(a, b):
    return float(np.sqrt(((a - b) ** 2).mean()) / np.sqrt((b ** 2).mean()))


# The reference's phase clocks (maze_poisson/solver.py:411-481): integrator (both
# halves), charges, field, forces (field + non-electrostatic + total).
CPU_PHASES = ("integrator", "charges", "field", "forces_field", "forces_notelec", "forces")


def cpu_baseline_child(steps, n_ions, n_grid, seed):
    """The reference C (compiled from its own sources with its own flags, OpenMP)
    on the host: bootstrap, one untimed step, then `steps` timed MD steps through the
    reference driver's call sequence, each phase timed on its own."""
    from maze_poisson_amd.capi import MazeLib
    lib = MazeLib(os.path.join(ROOT, "oracle", "_ref", "libmaze_ref_fast.so"), extensions=False)
    s, tol = make_system(n_ions, n_grid, seed)
    driver.initialize(lib, s, driver.RunConfig(method="MAZE-LCG", tol=tol, integrator="VERLET"))
    driver.bootstrap(lib)
    driver.md_step(lib)
    pc = time.perf_counter
    ph = {k: 0.0 for k in CPU_PHASES}
    lat = []
    t_start = pc()
    for _ in range(steps):
        t0 = pc()
        lib.integrator_part_1()
        t1 = pc()
        lib.solver_update_charges()
        t2 = pc()
        lib.solver_update_field()
        t3 = pc()
        lib.solver_compute_forces_elec()
        t4 = pc()
        lib.solver_compute_forces_noel()
        t5 = pc()
        lib.solver_compute_forces_tot()
        t6 = pc()
        lib.integrator_part_2()
        t7 = pc()
        ph["integrator"] += (t1 - t0) + (t7 - t6)
        ph["charges"] += t2 - t1
        ph["field"] += t3 - t2
        ph["forces_field"] += t4 - t3
        ph["forces_notelec"] += t5 - t4
        ph["forces"] += t6 - t3
        lat.append(t7 - t0)
    el = pc() - t_start
    lat = np.array(lat) * 1e3
    print(json.dumps({"steps_per_s": steps / el, "elapsed": el,
                      "p50_ms": float(np.percentile(lat, 50)),
                      "p99_ms": float(np.percentile(lat, 99)),
                      "phases_ms_per_step": {k: v / steps * 1e3 for k, v in ph.items()}}))


def host_cores():
    """Cores this process may run on (its CPU affinity)."""
    try:
        return len(os.sched_getaffinity(0))
    except (AttributeError, OSError):
        return os.cpu_count() or 1


def cpu_run(steps, cores, system=(N_IONS, N_GRID, SEED)):
    """One reference CPU measurement in a child process (OMP_NUM_THREADS = cores)."""
    path = os.path.join(ROOT, "oracle", "_ref", "libmaze_ref_fast.so")
    if not os.path.exists(path):
        return None
    env = dict(os.environ, OMP_NUM_THREADS=str(cores), OMP_PROC_BIND="close", OMP_PLACES="cores")
    n_ions, n_grid, seed = system
    try:
        out = subprocess.run([sys.executable, os.path.abspath(__file__), "--cpu-child",
                              "--cpu-steps", str(steps), "--cpu-system", str(n_ions), str(n_grid),
                              str(seed)], env=env, capture_output=True, timeout=300, text=True)
        r = json.loads([ln for ln in out.stdout.splitlines() if ln.startswith("{")][-1])
    except Exception as e:  # a reported baseline, never required
        sys.stderr.write(f"cpu baseline failed: {e}\n")
        return None
    return {"steps_per_s": round(r["steps_per_s"], 4), "cores": cores, "steps": steps,
            "p50_ms": round(r["p50_ms"], 3), "p99_ms": round(r["p99_ms"], 3),
            "phases_ms_per_step": {k: round(v, 3) for k, v in r["phases_ms_per_step"].items()}}


def cpu_baseline(steps, cores):
    r = cpu_run(steps, cores)
    if r is None:
        return None, None
    line = {"value": r["steps_per_s"], "unit": "MD steps/s", "cores": cores,
            "kind": "reference",
            "sample": f"reference C compiled from its sources (-O3 -ffast-math -mavx2, OpenMP "
                      f"x{cores}, all cores of this process's affinity) on the same 4096-ion / "
                      f"128^3 workload: {steps} MD steps after the 2 bootstrap steps and 1 "
                      f"untimed step"}
    return line, r


def ensemble_rate(lib, replicas, steps, slots, graphs=True):
    """Uneven-size replicas advanced together by the replica engine: the rate of the
    whole run (set-up and start-up steps included) and the step rate from the
    difference of two runs of `steps` and 2 x `steps` steps per replica."""
    from maze_poisson_amd.engine import ReplicaEngine
    ss = W.ensemble_systems(replicas)
    runs = {}
    for k in (steps, 2 * steps):
        with ReplicaEngine(lib, max_slots=slots, graphs=graphs) as e:
            ids = [e.submit(s, k) for s in ss]
            t = time.perf_counter()
            st = e.run()
            runs[k] = (time.perf_counter() - t, st,
                       [e.request_stats(r) for r in ids])
    (t1, st1, rs1), (t2, _, _) = runs[steps], runs[2 * steps]
    atoms = sum(s.n_p for s in ss)
    return {"replicas": replicas, "slots": slots, "steps_each": steps,
            "sizes": sorted({(s.n_p, s.n_grid) for s in ss}),
            "fp32_multigrid_replicas": sum(1 for r in rs1 if r["precision"] == 2),
            "replica_steps_per_s": round(replicas * steps / (t2 - t1), 1),
            "atom_steps_per_s": round(atoms * steps / (t2 - t1), 0),
            "whole_run_replica_steps_per_s": round(replicas * steps / t1, 1),
            "p50_ms": round(st1["p50_ms"], 3), "p99_ms": round(st1["p99_ms"], 3),
            "max_running": int(st1["max_running"]),
            "host_issue_us_per_step": round(st1["host_issue_s"] / (replicas * steps) * 1e6, 1),
            "unconverged_solves": int(sum(r["failures"] for r in rs1))}


def serving_run(lib, count, rate_hz, steps_range, slots, budget_gb, threads=1):
    """Poisson-arrival MD-segment requests through the replica engine (continuous
    batching: admission by free slot and device-memory budget while others run)."""
    from maze_poisson_amd.engine import ReplicaEngine
    trace = W.serving_trace(count, rate_hz, steps_range=steps_range)
    with ReplicaEngine(lib, budget_bytes=int(budget_gb * 2**30), max_slots=slots,
                       threads=threads) as e:
        ids = [e.submit(r.system, r.steps, arrival=r.arrival) for r in trace]
        st = e.run()
        rs = [e.request_stats(i) for i in ids]
    waits = [r["admit_s"] - r["arrival_s"] for r in rs]
    return {"requests": count, "rate_hz": rate_hz, "steps_range": list(steps_range),
            "slots": slots, "budget_gb": budget_gb, "host_threads": threads,
            "ions_range": [min(r.system.n_p for r in trace), max(r.system.n_p for r in trace)],
            "grid_range": [min(r.system.n_grid for r in trace),
                           max(r.system.n_grid for r in trace)],
            "wall_s": round(st["wall_s"], 3),
            "atom_steps_per_s": round(st["atom_steps_per_s"], 0),
            "replica_steps_per_s": round(st["replica_steps_per_s"], 1),
            "p50_step_ms": round(st["p50_ms"], 3), "p99_step_ms": round(st["p99_ms"], 3),
            "max_running": int(st["max_running"]), "peak_gb": round(st["peak_bytes"] / 2**30, 2),
            "p50_admission_wait_s": round(float(np.median(waits)), 4),
            "max_admission_wait_s": round(float(np.max(waits)), 4),
            "unconverged_solves": int(sum(r["failures"] for r in rs))}


TRAFFIC_FILE = os.path.join(ROOT, "profiles", "hbm_traffic_4096ions_128grid.json")


def measured_traffic(name):
    """HBM bytes per launch of a timed kernel from the committed rocprofv3 FETCH_SIZE /
    WRITE_SIZE summary of this workload (copy-kernel calibrated), or None."""
    try:
        with open(TRAFFIC_FILE) as fh:
            return json.load(fh).get(name, {}).get("traffic_bytes")
    except (OSError, ValueError):
        return None


COPY_GBS = {}


def copy_rates(lib):
    """The copy kernel's measured rate on this box: over 26 MB (about the dominant
    kernel's working set, resident in the 256 MB Infinity Cache) and over 2 GB
    (HBM)."""
    if not COPY_GBS and lib.has("maze_copy_bandwidth"):
        COPY_GBS["cache_26mb"] = round(lib.maze_copy_bandwidth(26_214_400 // 16, 50), 1)
        COPY_GBS["hbm_2gb"] = round(lib.maze_copy_bandwidth(1 << 27, 10), 1)
    return COPY_GBS


def roofline(prof):
    """Dominant grid kernel (largest total event time) of the headline path."""
    G = N_GRID ** 3
    timed = [(k, v) for k, v in prof.items() if v[1] > 0]
    if not timed:
        return None
    best = max(timed, key=lambda kv: kv[1][0])
    name, (tot_ms, cnt) = best
    avg_s = tot_ms / cnt * 1e-3
    bpl = BYTES_PER_POINT[name] * G
    ach = bpl / avg_s / 1e9
    out = {"bound": "hbm", "kernel": name, "achieved": round(ach, 1), "peak": HBM_PEAK_GBS,
           "unit": "GB/s", "frac": round(ach / HBM_PEAK_GBS, 4), "traffic": measured_traffic(name),
           "avg_us": round(avg_s * 1e6, 2), "launches": cnt, "bytes_per_launch": bpl}
    if COPY_GBS:
        out["copy_gbs_measured"] = dict(COPY_GBS)
        out["frac_of_copy_cache"] = round(ach / COPY_GBS["cache_26mb"], 4)
        out["frac_of_copy_hbm"] = round(ach / COPY_GBS["hbm_2gb"], 4)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=200)
    ap.add_argument("--warmup", type=int, default=10)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="headline path only")
    ap.add_argument("--prof-steps", type=int, default=30,
                    help="steps run with per-kernel HIP events after the timed steps")
    ap.add_argument("--cpu-steps", type=int, default=15)
    ap.add_argument("--cpu-cores", type=int, default=0,
                    help="threads of the CPU baseline (default: the cores of this process)")
    ap.add_argument("--cpu-child", action="store_true")
    ap.add_argument("--cpu-system", type=int, nargs=3, default=[N_IONS, N_GRID, SEED],
                    metavar=("IONS", "GRID", "SEED"), help=argparse.SUPPRESS)
    ap.add_argument("--pb-radii-only", action="store_true",
                    help="only the radius-based PB step (8000 ions, 160^3); prints its JSON")
    ap.add_argument("--pb-synthetic-only", action="store_true",
                    help="only the synthetic-dielectric PB line (8000 ions, 160^3); prints its JSON")
    ap.add_argument("--ensemble", action="store_true",
                    help="only the 64-replica ensemble (replica engine); prints its JSON")
    ap.add_argument("--ensemble-steps", type=int, default=50)
    ap.add_argument("--no-graphs", action="store_true",
                    help="ensemble replicas launch every kernel separately (no HIP graphs)")
    ap.add_argument("--serving", action="store_true",
                    help="only the serving trace (replica engine); prints its JSON")
    ap.add_argument("--serve-requests", type=int, default=24)
    ap.add_argument("--serve-rate", type=float, default=4.0)
    ap.add_argument("--serve-steps", type=int, nargs=2, default=[100, 2000])
    ap.add_argument("--serve-slots", type=int, default=32)
    ap.add_argument("--serve-budget-gb", type=float, default=64.0)
    ap.add_argument("--serve-threads", type=int, default=1)
    ap.add_argument("--calib-copy", action="store_true",
                    help="launch one 256 MiB copy kernel first (byte-counter calibration)")
    args = ap.parse_args()
    if args.cpu_child:
        cpu_baseline_child(args.cpu_steps, *args.cpu_system)
        return
    if args.gpus != 1:
        print("bench.py: this build runs on one GPU; a multi-GPU run would shard independent "
              "replicas one per GPU and is not part of it.", file=sys.stderr)
        sys.exit(2)

    from maze_poisson_amd import load_library
    lib = load_library()
    if args.calib_copy:
        lib.lib.maze_calib_copy.argtypes = [__import__("ctypes").c_long]
        lib.lib.maze_calib_copy(32 << 20)

    if args.pb_radii_only:
        print(json.dumps(pb_radii_rate(lib, args.steps, args.prof_steps)))
        return
    if args.pb_synthetic_only:
        print(json.dumps(pb_synthetic(lib, args.steps)))
        return
    if args.ensemble:
        ens = ensemble_rate(lib, 64, args.ensemble_steps, 64, not args.no_graphs)
        ser = ensemble_rate(lib, 64, max(10, args.ensemble_steps // 2), 1, not args.no_graphs)
        ens["serial_replica_steps_per_s"] = ser["replica_steps_per_s"]
        print(json.dumps(ens))
        return
    if args.serving:
        print(json.dumps(serving_run(lib, args.serve_requests, args.serve_rate,
                                     tuple(args.serve_steps), args.serve_slots,
                                     args.serve_budget_gb, args.serve_threads)))
        return
    copy_rates(lib)
    head = run_path(lib, PATHS["mixed_mg"], args.steps, args.warmup, args.steps, args.prof_steps)
    line_extra = {}
    if not args.no_extra:
        k = max(10, args.steps // 2)
        fp64 = run_path(lib, PATHS["fp64"], k, args.warmup, k, args.prof_steps)
        mcg = run_path(lib, PATHS["mixed_cg"], max(5, args.steps // 10), 2, 0, 0)
        fmg = run_path(lib, PATHS["fp64_mg"], k, args.warmup, k, 0)
        # Force error of each path after 5 steps, against the FP64 solve converged
        # 1000x tighter from the same start.
        f_ref, _ = forces_after(lib, PATHS["fp64"], 1e-3, 5)
        errs, iters = {}, {}
        for name, prec in PATHS.items():
            f, its = forces_after(lib, prec, 1.0, 5)
            errs[name] = rms_rel(f, f_ref)
            iters[name] = float(np.mean(its))
        line_extra = {
            "runner_steps_per_s": round(runner_rate(lib, 200), 2),
            "pb_radii": pb_radii_rate(lib, max(10, args.steps // 10), 5),
            # The workloads' own step counts: 500 PB steps, 200 steps per replica,
            # 100-2000 steps per serving request.
            "pb_synthetic": pb_synthetic(lib, 500),
            "ensemble_64_replicas": ensemble_rate(lib, 64, 200, 64),
            "serving": serving_run(lib, 16, 4.0, (100, 2000), 32, 64.0),
            "python_driver_steps_per_s": round(driver_rate(lib, PATHS["mixed_mg"], 100, False), 2),
            "python_driver_with_getters_steps_per_s": round(
                driver_rate(lib, PATHS["mixed_mg"], 100, True), 2),
            "fp64_steps_per_s": round(fp64["steps_per_s"], 3),
            "fp64_p50_ms": round(fp64["p50_ms"], 4), "fp64_p99_ms": round(fp64["p99_ms"], 4),
            "mixed_cg_steps_per_s": round(mcg["steps_per_s"], 3),
            "fp64_mg_steps_per_s": round(fmg["steps_per_s"], 3),
            "fp64_mg_p50_ms": round(fmg["p50_ms"], 4),
            "speedup_vs_fp64_mg": round(head["steps_per_s"] / fmg["steps_per_s"], 3),
            "speedup_vs_fp64": round(head["steps_per_s"] / fp64["steps_per_s"], 3),
            "force_rms_rel_err_vs_fp64_tight": errs,
            "solver_iters_per_step": iters,
            "fp64_roofline": roofline(fp64["prof"]),
        }
        c1 = run_path(lib, PATHS["fp64"], 100, args.warmup, 100, 0,
                      system=(C1_IONS, C1_GRID, C1_SEED))
        line_extra["fp64_512_ions_64_grid"] = {
            "steps_per_s": round(c1["steps_per_s"], 2), "p50_ms": round(c1["p50_ms"], 4),
            "p99_ms": round(c1["p99_ms"], 4)}
    cores = args.cpu_cores or host_cores()
    cpu, cpu_detail = (None, None) if args.no_cpu else cpu_baseline(args.cpu_steps, cores)
    if cpu_detail is not None:
        # The plain reference-method case (512 ions / 64^3) on the host with all cores and
        # with one thread, beside its FP64 GPU row.
        c1 = (C1_IONS, C1_GRID, C1_SEED)
        cpu_detail = {"4096_ions_128_grid": cpu_detail,
                      "512_ions_64_grid": cpu_run(20, cores, c1),
                      "512_ions_64_grid_one_thread": cpu_run(10, 1, c1)}
        line_extra["cpu_baseline_detail"] = cpu_detail
    line = {
        "metric": "MD steps/s at 4096 ions/128^3 grid, FP32+refine vs FP64; p50/p99 step latency",
        "value": round(head["steps_per_s"], 3),
        "unit": "MD steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32+f64",
        "data": "synthetic",
        "config": {"workload": "4096-ion NaCl melt, 128^3 grid, MAZE-LCG (tol 1e-8 of the RHS "
                               "scale), TF cutoff L/2, velocity Verlet NVE dt 0.25 fs; field solve "
                               "FP32 multigrid-PCG inside FP64 refinement",
                   "ions": N_IONS, "grid": N_GRID, "seed": SEED},
        "p50_ms": round(head["p50_ms"], 4),
        "p99_ms": round(head["p99_ms"], 4),
        **line_extra,
        "roofline": roofline(head["prof"]),
        "kernel_avg_us": {k: round(v[0] / v[1] * 1e3, 2) for k, v in head["prof"].items() if v[1]},
        "cpu_baseline": cpu,
    }
    if cpu:
        line["speedup_vs_cpu"] = round(head["steps_per_s"] / cpu["value"], 2)
    print(json.dumps(line))


if __name__ == "__main__":
    main()

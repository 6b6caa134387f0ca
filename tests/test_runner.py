"""End-to-end runs from the reference's input files (YAML schema, species / potential /
input CSVs, output CSVs, restart files, time reversal), through any ABI library.

The runner's call sequence is checked against the scripted driver sequence (which the
golden vectors of the reference C pin), and, where the reference C library was built
from source here, against that library driven by the same runner.  Time reversal
(``examples/backward.yml``: a run with ``invert_time`` from the forward run's restart
file) must bring every ion back to its starting point to within the solver tolerance.
"""
import csv
import json
import os

import numpy as np
import pytest

from maze_poisson_amd import driver
from maze_poisson_amd import workloads as W
from maze_poisson_amd.config import load_file
from maze_poisson_amd.convert import convert_old_input, write_species, write_tosi_fumi
from maze_poisson_amd.solver import SolverMD, run_file

A0 = W.A0


PB_RADII_ANG = (0.95, 1.81)


def write_inputs(d, s, with_vel=True, names=("Na", "Cl")):
    """species.csv (with PB radii), tosi_fumi.csv and input.csv for System `s` under
    folder d."""
    write_species(os.path.join(d, "species.csv"),
                  [("Na", 1.0, W.M_NA, PB_RADII_ANG[0]), ("Cl", -1.0, W.M_CL, PB_RADII_ANG[1])])
    write_tosi_fumi(os.path.join(d, "tosi_fumi.csv"))
    with open(os.path.join(d, "input.csv"), "w") as f:
        f.write("type,x,y,z" + (",vx,vy,vz" if with_vel else "") + "\n")
        for i in range(s.n_p):
            row = [names[s.types[i]]] + [repr(float(v)) for v in s.pos[i] * A0]
            if with_vel:
                row += [repr(float(v)) for v in s.vel[i]]
            f.write(",".join(row) + "\n")


def yaml_text(s, path, *, n_steps=6, method="MAZE-LCG", tol=None, integrator="OVRVO",
              stride=1, flush=2, invert=False, input_file="input.csv", extra_grid="",
              extra_md="", extra_out="", restart=True, restart_field=False):
    tol = tol if tol is not None else W.poisson_tol(s)
    return f"""
grid_setting:
  N: {s.n_grid}
  N_p: {s.n_p}
  N_typs: 2
  L: {s.L!r}
  input_file: "{input_file}"
  particles_file: "species.csv"
  cas: CIC
{extra_grid}
md_variables:
  N_steps: {n_steps}
  T: 1550
  dt_fs: 0.25
  tol: {tol!r}
  elec: TRUE
  not_elec: TRUE
  invert_time: {str(invert).lower()}
  method: "{method}"
  integrator: {integrator}
  potential: TF
  potential_params_file: "tosi_fumi.csv"
{extra_md}
output_settings:
  path: '{path}'
  print_tot_force: true
  print_performance: true
  print_solute: TRUE
  print_energy: TRUE
  print_momentum: TRUE
  print_temperature: TRUE
  print_restart: {str(restart).lower()}
  print_restart_field: {str(restart_field).lower()}
  stride: {stride}
  flushstride: {flush}
{extra_out}
"""


def read_rows(path):
    with open(path) as f:
        r = csv.reader(f)
        head = next(r)
        return head, [row for row in r]


def restart_arrays(path):
    head, rows = read_rows(path)
    assert head == ["type", "x", "y", "z", "vx", "vy", "vz"]
    types = [r[0] for r in rows]
    x = np.array([[float(v) for v in r[1:4]] for r in rows])
    v = np.array([[float(v) for v in r[4:7]] for r in rows])
    return types, x, v


def run_yaml(tmp_path, text, lib, name="run.yml"):
    p = tmp_path / name
    p.write_text(text)
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        return run_file(str(p), lib=lib)
    finally:
        os.chdir(cwd)


def minimum_image(d, L):
    return d - L * np.round(d / L)


# ---- configuration ------------------------------------------------------------------

def test_schema_defaults_and_units(tmp_path):
    s = W.nacl_system(216, 24, seed=5)
    p = tmp_path / "in.yml"
    p.write_text(yaml_text(s, "out/", extra_md="  probe_radius_ang: 1.4"))
    gs, os_, md = load_file(str(p))
    assert gs.h == s.L / s.n_grid and gs.cas == "CIC" and gs.precond == "NONE"
    assert md.probe_radius == 1.4 / A0
    assert md.dt == 0.25 / W.T_AU and md.kBT == 1550 * W.KB
    assert md.gamma == 1e-3 and md.thermostat is False and md.elec is True
    assert os_.format == "csv" and os_.restart_step is None
    p.write_text(yaml_text(s, "out/", invert=True))
    assert load_file(str(p))[2].dt == -0.25 / W.T_AU
    # JSON carries the same schema.
    j = tmp_path / "in.json"
    j.write_text(json.dumps({"grid_setting": {"N": 8, "N_p": 2, "L": 8.0, "N_typs": 1,
                                              "input_file": "x.csv"},
                             "md_variables": {"N_steps": 1, "T": 300, "dt_fs": 1.0},
                             "output_settings": {}}))
    gs, os_, md = load_file(str(j))
    assert gs.h == 1.0 and md.method == "FFT" and os_.stride == 50


@pytest.mark.parametrize("bad, msg", [
    ("  h: 0.3\n", r"N \* h must equal L"),
    ("  bogus: 1\n", "unknown key"),
])
def test_schema_errors(tmp_path, bad, msg):
    s = W.nacl_system(216, 24, seed=5)
    p = tmp_path / "in.yml"
    p.write_text(yaml_text(s, "out/", extra_grid=bad))
    with pytest.raises(ValueError, match=msg):
        load_file(str(p))
    (tmp_path / "in.txt").write_text("x")
    with pytest.raises(ValueError, match="not supported"):
        load_file(str(tmp_path / "in.txt"))


def test_rejected_inputs(tmp_path, oracle):
    s = W.nacl_system(216, 24, seed=5)
    write_inputs(str(tmp_path), s)
    with pytest.raises(ValueError, match="only TF"):
        run_yaml(tmp_path, yaml_text(s, "o1/").replace("potential: TF", "potential: LJ"), oracle)
    with pytest.raises(ValueError, match="ionic strength I"):
        run_yaml(tmp_path, yaml_text(s, "o2/", extra_md="  poisson_boltzmann: true"), oracle)
    with pytest.raises(ValueError, match="not recognized"):
        run_yaml(tmp_path, yaml_text(s, "o3/", method="SOR"), oracle)
    # Old-schema input names the converter.
    (tmp_path / "old.csv").write_text("charge,mass,x,y,z\n1.0,22.99,1.0,1.0,1.0\n")
    with pytest.raises(ValueError, match="convert"):
        run_yaml(tmp_path, yaml_text(s, "o4/", input_file="old.csv"), oracle)
    # Wrong particle count.
    with open(tmp_path / "short.csv", "w") as f:
        f.write("type,x,y,z\nNa,1.0,1.0,1.0\n")
    with pytest.raises(ValueError, match="does not match N_p"):
        run_yaml(tmp_path, yaml_text(s, "o5/", input_file="short.csv"), oracle)
    # Missing TF pair.
    rows = (tmp_path / "tosi_fumi.csv").read_text().splitlines()
    (tmp_path / "tosi_fumi.csv").write_text("\n".join(rows[:-1]) + "\n")
    with pytest.raises(ValueError, match="unique pairs"):
        run_yaml(tmp_path, yaml_text(s, "o6/"), oracle)


# ---- outputs and call sequence --------------------------------------------------------

def test_outputs_and_call_sequence(tmp_path, oracle):
    """File set, headers, cadence (stride 2: iterations 0, 2, 4 and the forced final 6)
    and the state: the runner's final positions equal the scripted driver sequence
    (bootstrap + N_steps steps) on the same library, bit for bit."""
    s = W.nacl_system(216, 24, seed=6)
    write_inputs(str(tmp_path), s)
    solver = run_yaml(tmp_path, yaml_text(s, "out/", n_steps=6, stride=2, flush=1), oracle)
    out = tmp_path / "out"
    heads = {
        "energy": ["iter", "K", "V_notelec"],
        "momentum": ["iter", "Px", "Py", "Pz"],
        "temperature": ["iter", "T"],
        "tot_force": ["iter", "Fx", "Fy", "Fz"],
        "performance": ["iter", "time", "n_iters"],
    }
    for name, head in heads.items():
        h, rows = read_rows(out / f"{name}.csv")
        assert h == head, name
        assert [int(r[0]) for r in rows] == [0, 2, 4, 6], name
    h, rows = read_rows(out / "solute.csv")
    assert h == ["charge", "iter", "particle", "x", "y", "z", "vx", "vy", "vz",
                 "fx_elec", "fy_elec", "fz_elec"]
    assert len(rows) == 4 * s.n_p
    assert not (out / "forces_pb.csv").exists() and not (out / "restart_field.csv").exists()
    assert json.loads((out / "input.json").read_text())["md_variables"]["N_steps"] == 6

    # Re-run the scripted sequence on the arrays the files describe.
    s2 = W.System(n_grid=s.n_grid, L=s.L, types=s.types,
                  pos=np.ascontiguousarray(np.array([[float(repr(float(v))) for v in p]
                                                     for p in s.pos * A0]) / A0),
                  vel=s.vel, mass=s.mass, charges=s.charges)
    cfg = driver.RunConfig(method="MAZE-LCG", tol=W.poisson_tol(s), T=1550,
                           integrator="OVRVO")
    driver.initialize(oracle, s2, cfg)
    driver.bootstrap(oracle)
    its, vs = [], []
    for _ in range(6):
        it, v = driver.md_step(oracle)
        its.append(it)
        vs.append(v)
    snap = driver.snapshot(oracle, s2, fields=False)
    oracle.solver_finalize()
    _, x, v = restart_arrays(out / "restart.csv")
    assert np.array_equal(x, snap["pos"] * A0)
    assert np.array_equal(v, snap["vel"])
    # Iteration 4 is the state after 5 loop steps; its CG count is the 5th step's.
    _, perf = read_rows(out / "performance.csv")
    assert int(perf[2][2]) == its[4]
    _, en = read_rows(out / "energy.csv")
    assert float(en[3][2]) == vs[5] and float(en[3][1]) == snap["kinetic"]
    tot = np.array([float(c) for c in read_rows(out / "tot_force.csv")[1][3][1:]])
    assert np.array_equal(tot, snap["fcs_tot"].sum(axis=0))
    assert solver.types_num_to_str == {0: "Na", 1: "Cl"}


def test_generated_velocities(tmp_path, oracle):
    """Without velocities in the input they are N(0, kT/m) from a generator seeded 42."""
    s = W.nacl_system(216, 24, seed=7)
    write_inputs(str(tmp_path), s, with_vel=False)
    p = tmp_path / "in.yml"
    p.write_text(yaml_text(s, "out/", n_steps=0))
    gs, os_, md = load_file(str(p))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        solver = SolverMD(gs, md, os_, lib=oracle)
        oracle.solver_initialize()
        solver.maps = oracle.type_maps()
        solver.initialize_grid()
        solver.initialize_particles()
        v = np.zeros((s.n_p, 3))
        oracle.get_vel(v)
        oracle.solver_finalize()
    finally:
        os.chdir(cwd)
    ref = np.random.RandomState(42).normal(0.0, 1.0, size=(s.n_p, 3))
    scale = np.sqrt(1550 * W.KB / s.mass)[:, None]
    assert np.allclose(v / scale, ref, rtol=1e-13, atol=0)


@pytest.mark.parametrize("pb", [False, True])
def test_reference_library_through_runner(tmp_path, oracle, ref, pb):
    """The reference C (built from its sources) and the oracle, both driven by the
    runner from the same files, write the same trajectory (plain and with the
    radius-based PB boundary)."""
    if pb:
        s = W.nacl_system(64, 32, seed=25, density=0.1)
        kw = dict(extra_grid=PB_GRID, extra_md=PB_MD)
    else:
        s = W.nacl_system(216, 24, seed=8)
        kw = {}
    write_inputs(str(tmp_path), s)
    run_yaml(tmp_path, yaml_text(s, "o/", n_steps=4, **kw), oracle)
    run_yaml(tmp_path, yaml_text(s, "r/", n_steps=4, **kw), ref)
    _, xo, vo = restart_arrays(tmp_path / "o" / "restart.csv")
    _, xr, vr = restart_arrays(tmp_path / "r" / "restart.csv")
    assert np.abs(xo - xr).max() <= 1e-11 * np.abs(xr).max()
    assert np.abs(vo - vr).max() <= 1e-9 * np.abs(vr).max()
    eo = np.array(read_rows(tmp_path / "o" / "energy.csv")[1], dtype=float)
    er = np.array(read_rows(tmp_path / "r" / "energy.csv")[1], dtype=float)
    assert np.abs(eo - er).max() <= 1e-10 * np.abs(er).max()
    po = [int(r[2]) for r in read_rows(tmp_path / "o" / "performance.csv")[1]]
    pr = [int(r[2]) for r in read_rows(tmp_path / "r" / "performance.csv")[1]]
    assert all(abs(a - b) <= 1 for a, b in zip(po, pr)), (po, pr)


PB_GRID = "  eps_s: 80.0\n  I: 0.1\n  w: 1.0\n"
PB_MD = "  poisson_boltzmann: true\n  nonpolar_forces: true\n  gamma_np: 0.05\n"


def test_pb_run_matches_driver(tmp_path, oracle):
    """A PB run (radius-based maps every step, boundary forces, forces_pb.csv) equals
    the scripted driver sequence with the same radii, screening and surface tension."""
    s = W.nacl_system(64, 32, seed=24, density=0.1)
    write_inputs(str(tmp_path), s)
    run_yaml(tmp_path, yaml_text(s, "pb/", n_steps=3, extra_grid=PB_GRID, extra_md=PB_MD,
                                 extra_out="  print_forces_pb: true"), oracle)
    s2 = W.System(n_grid=s.n_grid, L=s.L, types=s.types,
                  pos=np.ascontiguousarray(np.array([[float(repr(float(v))) for v in p]
                                                     for p in s.pos * A0]) / A0),
                  vel=s.vel, mass=s.mass, charges=s.charges)
    cfg = driver.RunConfig(method="MAZE-LCG", tol=W.poisson_tol(s), T=1550, integrator="OVRVO",
                           eps_s=80.0, pb_radii=True, pb_w=1.0, radii_ang=PB_RADII_ANG,
                           nonpolar=True, gamma_np=0.05)
    driver.initialize(oracle, s2, cfg)
    driver.bootstrap(oracle, cfg)
    for _ in range(3):
        driver.md_step(oracle, cfg)
    snap = driver.snapshot(oracle, s2, fields=False)
    oracle.solver_finalize()
    _, x, v = restart_arrays(tmp_path / "pb" / "restart.csv")
    assert np.array_equal(x, snap["pos"] * A0) and np.array_equal(v, snap["vel"])
    h, rows = read_rows(tmp_path / "pb" / "forces_pb.csv")
    assert h == ["iter", "particle", "Fx_RF", "Fy_RF", "Fz_RF", "Fx_DB", "Fy_DB", "Fz_DB",
                 "Fx_IB", "Fy_IB", "Fz_IB", "Fx_NP", "Fy_NP", "Fz_NP"]
    last = np.array([[float(c) for c in r[2:]] for r in rows if r[0] == "3"])
    assert last.shape == (s.n_p, 12)
    for col, key in ((0, "fcs_elec"), (3, "fcs_db"), (6, "fcs_ib"), (9, "fcs_np")):
        assert np.array_equal(last[:, col:col + 3], snap[key]), key
        assert np.abs(snap[key]).max() > 0, key


# ---- restart and time reversal -----------------------------------------------------------

def reversal(tmp_path, lib, s, n_steps, method, tol, extra_md=""):
    """Forward run, then the backward run from its restart file; returns the largest
    minimum-image position error (A) and velocity error against the start."""
    write_inputs(str(tmp_path), s)
    run_yaml(tmp_path, yaml_text(s, "fwd/", n_steps=n_steps, method=method, tol=tol,
                                 stride=n_steps, extra_md=extra_md), lib, "fwd.yml")
    run_yaml(tmp_path, yaml_text(s, "bwd/", n_steps=n_steps, method=method, tol=tol,
                                 stride=n_steps, invert=True, input_file="fwd/restart.csv",
                                 extra_md=extra_md), lib, "bwd.yml")
    t0 = read_rows(tmp_path / "input.csv")[1]
    x0 = np.array([[float(v) for v in r[1:4]] for r in t0])
    v0 = np.array([[float(v) for v in r[4:7]] for r in t0])
    types, xf, _ = restart_arrays(tmp_path / "fwd" / "restart.csv")
    assert types == [r[0] for r in t0]
    types, xb, vb = restart_arrays(tmp_path / "bwd" / "restart.csv")
    moved = np.abs(minimum_image(xf - x0, s.L * A0)).max()
    return np.abs(minimum_image(xb - x0, s.L * A0)).max(), np.abs(vb - v0).max(), moved


@pytest.mark.parametrize("method", ["MAZE-LCG", "MAZE-MULTIGRID"])
def test_time_reversal(tmp_path, oracle, method):
    """``examples/forward.yml`` then ``examples/backward.yml``: velocity Verlet with
    the solver at tolerance 1e-10 (relative to the source scale) retraces its path."""
    s = W.nacl_system(216, 32, seed=9)  # the multigrid needs n >= 32
    dx, dv, moved = reversal(tmp_path, oracle, s, 12, method, W.poisson_tol(s, 1e-10))
    assert moved > 1e-3
    assert dx <= 1e-9, dx
    assert dv <= 1e-8 * np.abs(s.vel).max(), dv


def test_restart_field_file(tmp_path, oracle):
    """A time-reversed run that names ``restart_field_file`` starts from the fields of
    that file (``phi`` and ``phi_prev`` columns, C order) instead of the bootstrap."""
    s = W.nacl_system(216, 24, seed=10)
    write_inputs(str(tmp_path), s)
    run_yaml(tmp_path, yaml_text(s, "fwd/", n_steps=3, restart_field=True), oracle, "f.yml")
    head, rows = read_rows(tmp_path / "fwd" / "restart_field.csv")
    assert head == ["phi_prev", "phi"] and len(rows) == s.n_grid ** 3
    n = s.n_grid
    phip = np.array([float(r[0]) for r in rows]).reshape(n, n, n)
    phi = np.array([float(r[1]) for r in rows]).reshape(n, n, n)
    assert np.abs(phi).max() > 0 and not np.array_equal(phi, phip)
    p = tmp_path / "b.yml"
    p.write_text(yaml_text(s, "bwd/", n_steps=3, invert=True, input_file="fwd/restart.csv",
                           extra_grid='  restart_field_file: "fwd/restart_field.csv"'))
    gs, os_, md = load_file(str(p))
    cwd = os.getcwd()
    os.chdir(tmp_path)
    try:
        solver = SolverMD(gs, md, os_, lib=oracle)
        solver.initialize()
        a, b = np.zeros((n, n, n)), np.zeros((n, n, n))
        oracle.get_field(a)
        oracle.get_field_prev(b)
        assert np.array_equal(a, phi) and np.array_equal(b, phip)
        solver.md_loop()
    finally:
        os.chdir(cwd)
    assert solver.n_iters > 0


# ---- old-schema converter --------------------------------------------------------------

def test_convert_old_schema(tmp_path):
    src = tmp_path / "old.csv"
    src.write_text("charge,mass,radius,x,y,z,vx,vy,vz\n"
                   "1.0,22.99,1.0,8.3,8.2,8.1,1e-4,2e-4,3e-4\n"
                   "-1.0,35.453,1.0,10.3,10.2,10.1,-1e-4,0.0,5e-5\n"
                   "1.0,22.99,1.0,1.5,2.5,3.5,0.0,0.0,0.0\n"
                   "2.0,40.0,1.2,4.5,5.5,6.5,0.0,0.0,0.0\n")
    names = convert_old_input(str(src), str(tmp_path / "in.csv"), str(tmp_path / "sp.csv"))
    assert list(names) == ["Na", "Cl", "T0"]
    h, rows = read_rows(tmp_path / "in.csv")
    assert h == ["type", "x", "y", "z", "vx", "vy", "vz"]
    assert [r[0] for r in rows] == ["Na", "Cl", "Na", "T0"]
    assert rows[1][1:] == ["10.3", "10.2", "10.1", "-1e-4", "0.0", "5e-5"]
    h, rows = read_rows(tmp_path / "sp.csv")
    assert h == ["type", "charge", "mass", "radius"]
    assert rows[2] == ["T0", "2.0", "40.0", "1.2"]
    # Without velocities or radius.
    src.write_text("charge,mass,x,y,z\n-1.0,35.45,1.0,2.0,3.0\n")
    from maze_poisson_amd.convert import main
    assert main([str(src), str(tmp_path / "conv")]) == 0
    assert read_rows(tmp_path / "conv" / "input.csv") == (["type", "x", "y", "z"],
                                                          [["Cl", "1.0", "2.0", "3.0"]])
    assert read_rows(tmp_path / "conv" / "tosi_fumi.csv")[1][0][:2] == ["Cl", "Cl"]


@pytest.mark.skipif(not os.path.isdir("/root/reference/examples/input_files"),
                    reason="reference examples not present")
def test_convert_reference_examples(tmp_path):
    """Every old-schema example input of the reference converts to two NaCl species."""
    base = "/root/reference/examples"
    files = [os.path.join(base, "input_files", f) for f in
             sorted(os.listdir(os.path.join(base, "input_files")))]
    for f in files:
        names = convert_old_input(f, str(tmp_path / "in.csv"), str(tmp_path / "sp.csv"))
        assert set(names) == {"Na", "Cl"}, f
        n = len(read_rows(f)[1])
        assert len(read_rows(tmp_path / "in.csv")[1]) == n


# ---- the HIP library ----------------------------------------------------------------------

@pytest.mark.gpu
def test_runner_gpu_matches_oracle(tmp_path, hip, oracle):
    """The same input files through the HIP library and the oracle: trajectories,
    energies and per-ion electrostatic forces of every output step agree."""
    s = W.nacl_system(512, 32, seed=11)
    write_inputs(str(tmp_path), s)
    run_yaml(tmp_path, yaml_text(s, "g/", n_steps=4), hip)
    run_yaml(tmp_path, yaml_text(s, "o/", n_steps=4), oracle)
    # Total momentum cancels to ~1e-9 of the per-ion momenta: it is compared on their
    # scale, sum_i m_i |v_i|.
    p_scale = float((s.mass[:, None] * np.abs(s.vel)).sum())
    for name in ("energy", "momentum", "temperature"):
        g = np.array(read_rows(tmp_path / "g" / f"{name}.csv")[1], dtype=float)
        o = np.array(read_rows(tmp_path / "o" / f"{name}.csv")[1], dtype=float)
        scale = p_scale if name == "momentum" else np.abs(o[:, 1:]).max()
        assert np.array_equal(g[:, 0], o[:, 0]), name
        assert np.abs(g[:, 1:] - o[:, 1:]).max() <= 1e-9 * scale, name
    g = np.array(read_rows(tmp_path / "g" / "solute.csv")[1], dtype=float)
    o = np.array(read_rows(tmp_path / "o" / "solute.csv")[1], dtype=float)
    assert np.abs(g[:, 3:6] - o[:, 3:6]).max() <= 1e-11 * np.abs(o[:, 3:6]).max()
    assert np.abs(g[:, 9:12] - o[:, 9:12]).max() <= 1e-7 * np.abs(o[:, 9:12]).max()
    gp = [int(r[2]) for r in read_rows(tmp_path / "g" / "performance.csv")[1]]
    op = [int(r[2]) for r in read_rows(tmp_path / "o" / "performance.csv")[1]]
    assert all(abs(a - b) <= 1 for a, b in zip(gp, op)), (gp, op)


@pytest.mark.gpu
def test_runner_gpu_pb_matches_oracle(tmp_path, hip, oracle):
    """A PB run (maps from the radii every step, boundary forces) through the HIP library
    and the oracle: trajectories and every per-ion force column of forces_pb.csv."""
    s = W.nacl_system(64, 32, seed=26, density=0.1)
    write_inputs(str(tmp_path), s)
    for d, lib in (("g/", hip), ("o/", oracle)):
        run_yaml(tmp_path, yaml_text(s, d, n_steps=4, method="MAZE-MULTIGRID", extra_grid=PB_GRID,
                                     extra_md=PB_MD, extra_out="  print_forces_pb: true"), lib)
    _, xg, _ = restart_arrays(tmp_path / "g" / "restart.csv")
    _, xo, _ = restart_arrays(tmp_path / "o" / "restart.csv")
    assert np.abs(xg - xo).max() <= 1e-11 * np.abs(xo).max()
    g = np.array([[float(c) for c in r] for r in read_rows(tmp_path / "g" / "forces_pb.csv")[1]])
    o = np.array([[float(c) for c in r] for r in read_rows(tmp_path / "o" / "forces_pb.csv")[1]])
    assert np.array_equal(g[:, :2], o[:, :2])
    for c in range(2, 14, 3):
        scale = np.abs(o[:, c:c + 3]).max()
        assert scale > 0
        assert np.abs(g[:, c:c + 3] - o[:, c:c + 3]).max() <= 1e-7 * scale, c


@pytest.mark.gpu
@pytest.mark.parametrize("precision", ["FP64", "MIXED-MG"])
def test_time_reversal_gpu(tmp_path, hip, precision):
    """Forward then backward on the 4096-ion / 128^3 workload (MaZe multigrid, tolerance
    1e-10 relative to the source scale): every ion returns to its start."""
    s = W.nacl_system(4096, 128, seed=12)
    dx, dv, moved = reversal(tmp_path, hip, s, 20, "MAZE-MULTIGRID", W.poisson_tol(s, 1e-10),
                             extra_md=f"  precision: {precision}")
    assert moved > 1e-3
    assert dx <= 1e-8, dx
    assert dv <= 1e-8 * np.abs(s.vel).max(), dv

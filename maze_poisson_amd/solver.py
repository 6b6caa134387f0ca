"""End-to-end MD runs from the reference's input files: ``SolverMD``.

The counterpart of the reference driver (``maze_poisson/solver.py:57-517``), over
any library exporting the MaZe C-ABI (the HIP library by default):

* ``initialize`` -- enum discovery, grid, particles from ``species.csv`` + the input
  CSV + ``tosi_fumi.csv`` with the reference's unit conversions, integrator
  (``solver.py:97-369``);
* ``initialize_md`` -- the two-step bootstrap with cold field solves, or, for a
  time-reversed run that names ``restart_field_file``, the fields read back from
  that file; then the optional velocity rescale (``solver.py:371-409``);
* ``md_loop`` -- ``init_steps`` unrecorded steps, then ``N_steps`` steps with the
  outputs of :mod:`.outputs` after each (``solver.py:483-510``), and a final forced
  output labelled ``N_steps`` (``solver.py:512-517``).

File formats (all CSV with a header row):

* particles file: ``type,charge,mass[,radius]`` -- one row per species, masses in
  amu; the row order defines the type numbers;
* input file: ``type,x,y,z[,vx,vy,vz]`` -- positions in Angstrom, velocities in a.u.;
  without velocities they are drawn from N(0, kT/m) with a NumPy generator seeded 42
  (``solver.py:17,319-330``);
* potential file: ``type1,type2,A,B,C,D,sigma`` (kJ/mol, 1/A, kJ/mol A^6,
  kJ/mol A^8, A), one row per unordered pair of types;
* restart field file: ``phi_prev,phi`` with n^3 rows in C order.

With ``poisson_boltzmann: true`` the grid gets the linear PB operator with the
screening of ionic strength ``I`` (mol/L) and transition width ``w`` (Bohr), each
ion the radius of its species (particles-file column ``radius``, A) plus
``probe_radius``, the dielectric and screening maps are rebuilt from the radii every
step and the boundary forces join the total (``solver.py:146-162``, :348-356,
:411-415, :421-431).

Old-schema inputs (``charge,mass[,radius],x,y,z[,...]``) are converted by
:mod:`.convert`.  Only the Tosi-Fumi potential is provided by the library.
"""
from __future__ import annotations

import csv
import json
import logging
import math
import os
import time

import numpy as np

from .capi import MazeLib
from .config import GridSetting, MDVariables, OutputSettings, load_file
from .outputs import OutputFiles
from .workloads import A0, CONV_MASS, KJMOL_TO_HARTREE, kbar2

log = logging.getLogger("maze_poisson_amd")

# solver_set_precision modes of the HIP library.
PRECISION_MODES = {"FP64": 0, "MIXED-CG": 1, "MIXED-MG": 2}


def read_csv_columns(path: str) -> dict[str, list[str]]:
    """Columns of a CSV file with a header row, as lists of strings."""
    with open(path, newline="", encoding="utf-8") as f:
        reader = csv.reader(f)
        try:
            header = [h.strip() for h in next(reader)]
        except StopIteration:
            raise ValueError(f"Empty or incorrect input file `{path}`.") from None
        cols: dict[str, list[str]] = {h: [] for h in header}
        for row in reader:
            if not row:
                continue
            if len(row) != len(header):
                raise ValueError(f"{path}: row {row} does not match header {header}")
            for h, v in zip(header, row):
                cols[h].append(v.strip())
    return cols


def _floats(col: list[str]) -> np.ndarray:
    return np.array([float(v) for v in col], dtype=np.float64)


class SolverMD:
    """One MD run (``maze_poisson/solver.py:57-95``)."""

    def __init__(self, gset: GridSetting, mdv: MDVariables, outset: OutputSettings,
                 lib: MazeLib | None = None):
        if lib is None:
            from . import load_library
            lib = load_library()
        self.lib = lib
        self.gset, self.mdv, self.outset = gset, mdv, outset
        self.L, self.h, self.N = gset.L, gset.h, gset.N
        self.N_p, self.N_typs = gset.N_p, gset.N_typs
        self.thermostat = mdv.thermostat
        self.n_iters = 0
        self.t_iters = 0.0
        self.potential_notelec = 0.0
        self.energy_nonpolar = 0.0
        self.types_str_to_num: dict[str, int] = {}
        self.types_num_to_str: dict[int, str] = {}
        self.timers: dict[str, float] = {}
        if outset.print_restart:
            outset.restart_step = outset.restart_step or mdv.N_steps
        self.ofiles = OutputFiles(outset)
        self.save_input()

    # ---- set-up --------------------------------------------------------------
    def save_input(self) -> None:
        """``solver.py:519-528``: the parsed input, as ``input.json``."""
        dct = {"grid_setting": self.gset.to_dict(), "md_variables": self.mdv.to_dict(),
               "output_settings": self.outset.to_dict()}
        with open(os.path.join(self.outset.path, "input.json"), "w") as f:
            json.dump(dct, f, indent=2)

    def initialize(self) -> None:
        """``solver.py:97-108``"""
        self.lib.solver_initialize()
        self.maps = self.lib.type_maps()
        self.initialize_grid()
        self.initialize_particles()
        self.initialize_integrator()
        self.initialize_md()

    def initialize_grid(self) -> None:
        """``solver.py:129-162``"""
        method = self.mdv.method.upper()
        if method not in self.maps["grid"]:
            raise ValueError(f"Method {method} not recognized.")
        precond = self.gset.precond.upper()
        if precond not in self.maps["precond"]:
            raise ValueError(f"Preconditioner {precond} not recognized.")
        self.lib.solver_initialize_grid(self.N, self.L, self.h, self.mdv.tol, self.gset.eps_s,
                                        self.gset.eps_int, self.maps["grid"][method],
                                        self.maps["precond"][precond])
        if self.mdv.poisson_boltzmann:
            if self.gset.I is None or self.gset.w is None:
                raise ValueError("poisson_boltzmann needs the ionic strength I and the "
                                 "transition width w in grid_setting.")
            kb2 = kbar2(self.gset.I, self.gset.eps_s, self.mdv.T, self.h)
            self.lib.solver_initialize_grid_pois_boltz(self.gset.w, kb2,
                                                       1 if self.mdv.nonpolar_forces else 0)
        prec = self.mdv.precision.upper()
        if prec not in PRECISION_MODES:
            raise ValueError(f"precision {prec} not recognized among "
                             f"{','.join(PRECISION_MODES)}.")
        if prec != "FP64" or self.lib.has("solver_set_precision"):
            if self.lib.solver_set_precision(PRECISION_MODES[prec]) != 0:
                raise ValueError(f"precision {prec} rejected by the library")

    def get_tosi_fumi_params(self, species: dict[str, dict]) -> np.ndarray:
        """``solver.py:164-205``: the (n_typ, n_typ, 5) symmetric table in a.u."""
        path = self.mdv.potential_params_file
        if path is None:
            raise ValueError("Potential parameters file must be provided for TF potential.")
        cols = read_csv_columns(path)
        need = ["type1", "type2", "A", "B", "C", "D", "sigma"]
        missing = [c for c in need if c not in cols]
        if missing:
            raise ValueError(f"{path}: missing column(s) {', '.join(missing)}")
        nt = self.N_typs
        expected = nt * (nt + 1) // 2
        if len(cols["type1"]) != expected:
            raise ValueError(
                f"Potential parameters file must have {expected} unique pairs of types.")
        table = np.full((nt, nt, 5), np.nan)
        seen = set()
        for r, (t1, t2) in enumerate(zip(cols["type1"], cols["type2"])):
            for t in (t1, t2):
                if t not in species:
                    raise ValueError(f"Particle type not found in particles file: {t}")
            if (t1, t2) in seen:
                raise ValueError(
                    "Potential parameters file must have unique pairs of types (type1, type2).")
            seen.add((t1, t2))
            a, b = species[t1]["enum"], species[t2]["enum"]
            if not np.isnan(table[a, b, 0]):  # (t1, t2) after (t2, t1)
                raise ValueError(f"Duplicate potential parameters for types {a} and {b}.")
            table[a, b] = [float(cols[c][r]) for c in ("A", "B", "C", "D", "sigma")]
            table[b, a] = table[a, b]
        if np.any(np.isnan(table)):
            raise ValueError("Potential parameters for some particle types are missing.")
        table[:, :, 0] *= KJMOL_TO_HARTREE
        table[:, :, 1] *= A0
        table[:, :, 2] *= KJMOL_TO_HARTREE / A0**6
        table[:, :, 3] *= KJMOL_TO_HARTREE / A0**8
        table[:, :, 4] /= A0
        return np.ascontiguousarray(table.reshape(-1), dtype=np.float64)

    def read_species(self) -> dict[str, dict]:
        """``solver.py:288-301``: species by name, numbered in file order."""
        cols = read_csv_columns(self.gset.particles_file)
        for c in ("type", "charge", "mass"):
            if c not in cols:
                raise ValueError(f"{self.gset.particles_file}: missing column {c}")
        names = cols["type"]
        if len(names) != self.N_typs:
            raise ValueError(f"Number of particle types in file ({len(names)}) does not "
                             f"match N_typs ({self.N_typs}).")
        if len(set(names)) != self.N_typs:
            raise ValueError("Particle types in file must be unique.")
        species = {}
        for i, name in enumerate(names):
            self.types_str_to_num[name] = i
            self.types_num_to_str[i] = name
            species[name] = {"enum": i, "charge": float(cols["charge"][i]),
                             "mass": float(cols["mass"][i])}
            if "radius" in cols:
                species[name]["radius"] = float(cols["radius"][i])
        return species

    def initialize_particles(self) -> None:
        """``solver.py:288-346``"""
        species = self.read_species()
        potential = self.mdv.potential.upper()
        if potential not in self.maps["potential"]:
            raise ValueError(f"Potential {potential} not recognized among "
                             f"{','.join(self.maps['potential'])}.")
        if potential != "TF":
            raise ValueError(f"Potential {potential}: only TF is provided by this library.")
        cas = self.gset.cas.upper()
        if cas not in self.maps["cas"]:
            raise ValueError(f"Charge assignment scheme {cas} not recognized.")

        path = self.gset.input_file
        cols = read_csv_columns(path)
        if "type" not in cols:
            if "charge" in cols:
                raise ValueError(f"`{path}` has the old input schema (charge, mass, ...); "
                                 "convert it with `python -m maze_poisson_amd convert`.")
            raise ValueError(f"Empty or incorrect input file `{path}`.")
        for t in cols["type"]:
            if t not in species:
                raise ValueError(f"Particle type {t} of `{path}` not in the particles file.")
        sp = [species[t] for t in cols["type"]]
        types = np.ascontiguousarray([s["enum"] for s in sp], dtype=np.int32)
        pos = np.stack([_floats(cols[c]) for c in ("x", "y", "z")], axis=1) / A0
        pos = np.ascontiguousarray(pos, dtype=np.float64)
        charges = np.ascontiguousarray([s["charge"] for s in sp], dtype=np.float64)
        mass = np.ascontiguousarray([s["mass"] for s in sp], dtype=np.float64) * CONV_MASS
        if not pos.size:
            raise ValueError(f"Empty or incorrect input file `{path}`.")
        if len(pos) != self.N_p:
            raise ValueError(f"Number of particles in file ({len(pos)}) does not match "
                             f"N_p ({self.N_p}).")
        if "vx" in cols:
            vel = np.stack([_floats(cols[c]) for c in ("vx", "vy", "vz")], axis=1)
        else:
            rng = np.random.RandomState(42)
            vel = rng.normal(loc=0.0, scale=np.sqrt(self.mdv.kBT / mass[:, None]),
                             size=(self.N_p, 3))
        vel = np.ascontiguousarray(vel, dtype=np.float64)
        tf = self.get_tosi_fumi_params(species)
        self.lib.solver_initialize_particles(self.N, self.N_typs, self.L, self.h, self.N_p,
                                             self.maps["potential"][potential],
                                             self.maps["cas"][cas], types, pos, vel, mass,
                                             charges, tf)
        if self.mdv.poisson_boltzmann:
            if any("radius" not in s for s in sp):
                raise ValueError("Probe radius must be provided in the input file for "
                                 "Poisson-Boltzmann.")
            radius = np.array([s["radius"] for s in sp], dtype=np.float64) / A0
            radius = np.ascontiguousarray(radius + self.mdv.probe_radius)
            self.lib.solver_initialize_particles_pois_boltz(self.mdv.gamma_np_au,
                                                            self.mdv.beta_np, radius)

    def initialize_integrator(self) -> None:
        """``solver.py:358-369``"""
        name = self.mdv.integrator.upper()
        if name not in self.maps["integrator"]:
            raise ValueError(f"Integrator {name} not recognized.")
        self.lib.solver_initialize_integrator(self.N_p, self.mdv.dt, self.mdv.T,
                                              self.mdv.gamma, self.maps["integrator"][name],
                                              1 if self.mdv.thermostat else 0)

    def initialize_md(self) -> None:
        """``solver.py:371-409``.  With a restart field (time-reversed runs only) the
        bootstrap is skipped, so the first half-kick uses the forces the library
        holds after initialisation (zero), exactly as in the reference."""
        ffile = self.gset.restart_field_file
        if ffile is None or not self.mdv.invert_time:
            self.update_charges()
            self.update_eps_k2()
            self.lib.solver_init_field()
            self.compute_forces()
            self.lib.integrator_part_1()
            self.update_charges()
            self.update_eps_k2()
            self.lib.solver_init_field()
            self.compute_forces()
            self.lib.integrator_part_2()
        else:
            cols = read_csv_columns(ffile)
            n = self.N
            for c in ("phi", "phi_prev"):
                if c not in cols or len(cols[c]) != n ** 3:
                    raise ValueError(f"{ffile}: column {c} must hold {n}^3 values")
            self.lib.solver_set_field(np.ascontiguousarray(_floats(cols["phi"]).reshape(n, n, n)))
            self.lib.solver_set_field_prev(
                np.ascontiguousarray(_floats(cols["phi_prev"]).reshape(n, n, n)))
            log.info("Initialization step skipped due to field loaded from file.")
        if self.mdv.rescale:
            self.lib.solver_rescale_velocities()

    # ---- one step ----------------------------------------------------------------
    def update_charges(self) -> None:
        """``solver.py:458-463``: a charge-conservation failure ends the run."""
        if self.lib.solver_update_charges() != 0:
            raise RuntimeError("Error: change initial position, charge is not preserved.")

    def update_eps_k2(self) -> None:
        """``solver.py:411-415``"""
        if self.mdv.poisson_boltzmann:
            self.lib.solver_update_eps_k2()

    def update_field(self) -> int:
        res = self.lib.solver_update_field()
        if res == -1:
            log.warning("Warning: CG did not converge.")
        return res

    def compute_forces(self) -> None:
        """``solver.py:421-431``"""
        if self.mdv.elec:
            self.lib.solver_compute_forces_elec()
        if self.mdv.not_elec:
            self.potential_notelec = self.lib.solver_compute_forces_noel()
        if self.mdv.poisson_boltzmann:
            self.energy_nonpolar = self.lib.solver_compute_forces_pb()
        self.lib.solver_compute_forces_tot()

    def md_loop_iter(self) -> None:
        """``solver.py:475-483``"""
        self.lib.integrator_part_1()
        if self.mdv.elec:
            self.update_charges()
            self.update_eps_k2()
            t0 = time.perf_counter()
            self.n_iters = self.update_field()
            self.t_iters = time.perf_counter() - t0
        self.compute_forces()
        self.lib.integrator_part_2()

    def md_loop(self) -> None:
        """``solver.py:485-503``"""
        for _ in range(self.mdv.init_steps):
            self.md_loop_iter()
        log.info("Temperature: %.2f K", self.lib.get_temperature())
        for i in range(self.mdv.N_steps):
            self.md_loop_iter()
            self.ofiles.output(i, self)

    def run(self) -> None:
        """``solver.py:505-510``"""
        t0 = time.perf_counter()
        self.initialize()
        t1 = time.perf_counter()
        self.md_loop()
        t2 = time.perf_counter()
        self.ofiles.output(self.mdv.N_steps, self, force=True)
        self.ofiles.flush()
        self.timers = {"initialize": t1 - t0, "md_loop": t2 - t1,
                       "steps_per_s": (self.mdv.N_steps + self.mdv.init_steps) / (t2 - t1)
                       if t2 > t1 else math.inf}
        log.info("timers: %s", self.timers)


def run_file(path: str, lib: MazeLib | None = None, verbose: bool = False) -> SolverMD:
    """``maze run md FILE`` (``maze_poisson/cli/run.py:13-36``).  Relative file names
    in the input are taken relative to the current directory, as in the reference."""
    gs, os_, ms = load_file(path)
    if verbose:
        os_.debug = True
        logging.basicConfig(level=logging.INFO)
    solver = SolverMD(gs, ms, os_, lib=lib)
    solver.run()
    return solver

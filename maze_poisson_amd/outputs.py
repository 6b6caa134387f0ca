"""CSV outputs of an MD run, in the reference's file names, columns and cadence.

Files (``<path>/<name>.csv``), each written when ``print_<name>`` is set
(``maze_poisson/myio/output/csv.py:45-199``):

=============  ==================================================================
energy         iter, K, V_notelec                     (one row per output step)
momentum       iter, Px, Py, Pz
temperature    iter, T
tot_force      iter, Fx, Fy, Fz                       (sum of the total forces)
performance    iter, time, n_iters                    (field update: s, CG result)
solute         charge, iter, particle, x, y, z, vx, vy, vz, fx_elec, fy_elec, fz_elec
               (one row per particle; positions in Bohr)
forces_pb      iter, particle, Fx_RF..Fz_RF, Fx_DB.., Fx_IB.., Fx_NP..
restart        type, x, y, z, vx, vy, vz              (positions in Angstrom)
restart_field  phi_prev, phi                          (n^3 rows, C order)
=============  ==================================================================

Cadence (``maze_poisson/myio/output/base_out.py:142-160``): the per-step files are
written when ``itr % stride == 0`` (or when forced at the end of the run) and at
most once per ``itr``; buffers are flushed to disk every ``flushstride * stride``
steps and when forced.  The restart files are rewritten (header included) when
``itr == restart_step``.  Numbers are written in Python's shortest round-trip
form, so a restart file read back gives the same doubles.
"""
from __future__ import annotations

import io
import os

import numpy as np

from .workloads import A0


def _fmt(v) -> str:
    if isinstance(v, (float, np.floating)):
        return repr(float(v))
    if isinstance(v, (int, np.integer)):
        return str(int(v))
    return str(v)


class CSVOutputFile:
    """One output file: header at creation, rows appended to a buffer, the buffer
    appended to disk on flush (``csv.py:10-33``, ``base_out.py:29-70``)."""
    name = ""
    headers: list[str] = []

    def __init__(self, path: str, enabled: bool):
        self.path = os.path.abspath(path)
        self.enabled = enabled
        self.buffer = io.StringIO()
        if not enabled:
            return
        if os.path.exists(self.path):
            os.remove(self.path)
        self.buffer.write(",".join(self.headers) + "\n")

    def rows(self, itr: int, solver) -> list[list]:
        raise NotImplementedError

    def write_data(self, itr: int, solver, mode: str = "a") -> None:
        if not self.enabled:
            return
        if mode == "w":
            open(self.path, "w").close()
            self.buffer.seek(0)
            self.buffer.truncate(0)
            self.buffer.write(",".join(self.headers) + "\n")
        for row in self.rows(itr, solver):
            self.buffer.write(",".join(_fmt(v) for v in row) + "\n")

    def flush(self) -> None:
        if self.enabled:
            with open(self.path, "a", encoding="utf-8") as f:
                f.write(self.buffer.getvalue())
            self.buffer.seek(0)
            self.buffer.truncate(0)

    close = flush


class EnergyFile(CSVOutputFile):
    name = "energy"
    headers = ["iter", "K", "V_notelec"]

    def rows(self, itr, s):
        return [[itr, s.lib.get_kinetic_energy(), s.potential_notelec]]


class MomentumFile(CSVOutputFile):
    name = "momentum"
    headers = ["iter", "Px", "Py", "Pz"]

    def rows(self, itr, s):
        m = np.empty(3)
        s.lib.get_momentum(m)
        return [[itr, m[0], m[1], m[2]]]


class TemperatureFile(CSVOutputFile):
    name = "temperature"
    headers = ["iter", "T"]

    def rows(self, itr, s):
        return [[itr, s.lib.get_temperature()]]


class TotForceFile(CSVOutputFile):
    name = "tot_force"
    headers = ["iter", "Fx", "Fy", "Fz"]

    def rows(self, itr, s):
        f = np.empty((s.N_p, 3))
        s.lib.get_fcs_tot(f)
        t = f.sum(axis=0)
        return [[itr, t[0], t[1], t[2]]]


class PerformanceFile(CSVOutputFile):
    name = "performance"
    headers = ["iter", "time", "n_iters"]

    def rows(self, itr, s):
        return [[itr, s.t_iters, s.n_iters]]


class SoluteFile(CSVOutputFile):
    name = "solute"
    headers = ["charge", "iter", "particle", "x", "y", "z", "vx", "vy", "vz",
               "fx_elec", "fy_elec", "fz_elec"]

    def rows(self, itr, s):
        n = s.N_p
        pos, vel, fel = np.empty((n, 3)), np.empty((n, 3)), np.empty((n, 3))
        q = np.empty(n)
        s.lib.get_pos(pos)
        s.lib.get_vel(vel)
        s.lib.get_fcs_elec(fel)
        s.lib.get_charges(q)
        return [[q[i], itr, i, *pos[i], *vel[i], *fel[i]] for i in range(n)]


class ForcesPBFile(CSVOutputFile):
    name = "forces_pb"
    headers = ["iter", "particle", "Fx_RF", "Fy_RF", "Fz_RF", "Fx_DB", "Fy_DB", "Fz_DB",
               "Fx_IB", "Fy_IB", "Fz_IB", "Fx_NP", "Fy_NP", "Fz_NP"]

    def rows(self, itr, s):
        n = s.N_p
        parts = []
        for getter in ("get_fcs_elec", "get_fcs_db", "get_fcs_ib", "get_fcs_np"):
            a = np.empty((n, 3))
            getattr(s.lib, getter)(a)
            parts.append(a)
        return [[itr, i, *parts[0][i], *parts[1][i], *parts[2][i], *parts[3][i]]
                for i in range(n)]


class RestartFile(CSVOutputFile):
    name = "restart"
    headers = ["type", "x", "y", "z", "vx", "vy", "vz"]

    def rows(self, itr, s):
        n = s.N_p
        pos, vel = np.empty((n, 3)), np.empty((n, 3))
        types = np.empty(n, dtype=np.int32)
        s.lib.get_pos(pos)
        s.lib.get_vel(vel)
        s.lib.get_types(types)
        pos = pos * A0
        return [[s.types_num_to_str[int(types[i])], *pos[i], *vel[i]] for i in range(n)]


class RestartFieldFile(CSVOutputFile):
    name = "restart_field"
    headers = ["phi_prev", "phi"]

    def rows(self, itr, s):
        n = s.N
        phi, phip = np.empty((n, n, n)), np.empty((n, n, n))
        s.lib.get_field(phi)
        s.lib.get_field_prev(phip)
        return list(zip(phip.reshape(-1).tolist(), phi.reshape(-1).tolist()))


class OutputFiles:
    """The set of output files of one run (``base_out.py:72-160``)."""
    files = [PerformanceFile, EnergyFile, MomentumFile, TemperatureFile, SoluteFile,
             TotForceFile, ForcesPBFile]
    files_restart = [RestartFile, RestartFieldFile]

    def __init__(self, oset):
        if oset.format != "csv":
            raise ValueError(f"Output format {oset.format!r} not supported (csv only)")
        self.oset = oset
        self.base_path = oset.path
        os.makedirs(self.base_path, exist_ok=True)
        self.out_stride = oset.stride
        self.out_flushstride = (oset.flushstride or 0) * oset.stride
        self.restart_step = oset.restart_step
        self.last = -1
        self.step_files = []
        self.restart_files = []
        for group, classes in ((self.step_files, self.files),
                               (self.restart_files, self.files_restart)):
            for cls in classes:
                f = cls(os.path.join(self.base_path, f"{cls.name}.csv"),
                        bool(getattr(oset, f"print_{cls.name}")))
                setattr(self, cls.name, f)
                group.append(f)

    def flush(self) -> None:
        for f in self.step_files + self.restart_files:
            f.flush()

    def output(self, itr: int, solver, force: bool = False) -> None:
        if force or itr % self.out_stride == 0:
            if self.last != itr:
                self.last = itr
                for f in self.step_files:
                    f.write_data(itr, solver)
                if force or (self.out_flushstride and itr % self.out_flushstride == 0):
                    self.flush()
        if self.restart_step == itr:
            for f in self.restart_files:
                if f.enabled:
                    f.write_data(itr, solver, mode="w")
                    f.flush()

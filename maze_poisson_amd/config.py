"""Run configuration: the reference's YAML / JSON input schema.

Three sections, each a dataclass with the reference's field names, defaults and
units (``maze_poisson/myio/input/base_file_input.py:48-147``):

* ``grid_setting`` -> :class:`GridSetting` (grid, box, input files, charge assignment,
  preconditioner, PB constants);
* ``md_variables`` -> :class:`MDVariables` (steps, temperature, time step, method,
  integrator, thermostat, potential, time reversal, PB switches);
* ``output_settings`` -> :class:`OutputSettings` (which CSV files to write, stride,
  flush stride, restart step).

Lengths are in Bohr.  A key ending in ``_ang`` is given in Angstrom and stored under
the name without the suffix, divided by the Bohr radius
(``base_file_input.py:10-18``).  An unknown key is an error, as in the reference
(``base_file_input.py:21-29``, where it logs and exits; here it raises).
"""
from __future__ import annotations

import json
import os
from dataclasses import asdict, dataclass, fields
from typing import Optional

from .workloads import A0, KB, T_AU


def normalize_ang_to_au(dct: dict) -> dict:
    """``key_ang: v`` (Angstrom) -> ``key: v / a0`` (Bohr); other keys unchanged."""
    out = {}
    for key, value in dct.items():
        if key.endswith("_ang"):
            out[key[: -len("_ang")]] = value / A0
        else:
            out[key] = value
    return out


class _Section:
    @classmethod
    def from_dict(cls, dct: Optional[dict]):
        dct = normalize_ang_to_au(dict(dct or {}))
        known = {f.name for f in fields(cls)}
        unknown = sorted(set(dct) - known)
        if unknown:
            raise ValueError(f"{cls.__name__}: unknown key(s) {', '.join(unknown)}")
        try:
            return cls(**dct)
        except TypeError as e:  # a required key is missing
            raise ValueError(f"{cls.__name__}: {e}") from None

    def to_dict(self) -> dict:
        return asdict(self)


@dataclass(kw_only=True)
class OutputSettings(_Section):
    """``base_file_input.py:48-67``"""
    print_solute: bool = False
    print_performance: bool = False
    print_momentum: bool = False
    print_energy: bool = False
    print_temperature: bool = False
    print_tot_force: bool = False
    print_forces_pb: bool = False
    print_restart: bool = False
    print_restart_field: bool = False

    path: str = "Outputs/"
    format: str = "csv"

    stride: int = 50
    flushstride: int = 5

    debug: bool = False
    restart_step: Optional[int] = None


@dataclass(kw_only=True)
class GridSetting(_Section):
    """``base_file_input.py:69-98``.  ``h`` defaults to L / N; when given, N * h must
    equal L exactly."""
    N: int
    N_p: int
    L: float
    h: Optional[float] = None
    eps_s: float = 1.0
    eps_int: float = 1.0

    N_typs: int

    particles_file: str = "species.csv"
    input_file: str

    restart_field_file: Optional[str] = None
    cas: str = "CIC"

    precond: str = "NONE"
    smoother: str = "LCG"

    I: Optional[float] = None   # ionic strength, mol/L (PB)
    w: Optional[float] = None   # width of the dielectric transition (PB)

    def __post_init__(self):
        if self.h is None:
            self.h = self.L / self.N
        elif self.N * self.h != self.L:
            raise ValueError("N * h must equal L. Check your values.")


@dataclass(kw_only=True)
class MDVariables(_Section):
    """``base_file_input.py:100-147``"""
    N_steps: int
    T: float        # K
    dt_fs: float    # fs

    init_steps: int = 0
    elec: bool = True
    not_elec: bool = True

    potential: str = "TF"
    potential_params_file: Optional[str] = None

    integrator: str = "OVRVO"
    method: str = "FFT"
    tol: float = 1e-7

    thermostat: bool = False
    gamma: float = 1e-3

    rescale: bool = False
    invert_time: bool = False

    poisson_boltzmann: bool = False
    nonpolar_forces: bool = False
    gamma_np: float = 0.0           # kcal/mol/A^2
    beta_np: float = 0.0            # kcal/mol
    probe_radius: float = 1.4 / A0  # Bohr

    # Not in the reference schema: arithmetic of the field solve, FP64 (the
    # reference's), MIXED-CG or MIXED-MG (FP32 inner solves refined to the same FP64
    # tolerance; MIXED-MG is the fast MaZe path).
    precision: str = "FP64"

    @property
    def kBT(self) -> float:
        return self.T * KB

    @property
    def dt(self) -> float:
        """Time step in a.u., negative when the run goes backwards in time."""
        return self.dt_fs / T_AU * (-1 if self.invert_time else 1)

    @property
    def gamma_np_au(self) -> float:
        return self.gamma_np * 0.0065934


def from_dict(data: dict) -> tuple[GridSetting, OutputSettings, MDVariables]:
    """``maze_poisson/myio/input/dct.py:6-15`` (same section order)."""
    if not isinstance(data, dict):
        raise ValueError("input must be a mapping with grid_setting, output_settings "
                         "and md_variables")
    return (GridSetting.from_dict(data.get("grid_setting")),
            OutputSettings.from_dict(data.get("output_settings")),
            MDVariables.from_dict(data.get("md_variables")))


def load_file(path: str) -> tuple[GridSetting, OutputSettings, MDVariables]:
    """YAML (``.yml`` / ``.yaml``) or JSON input, chosen by extension
    (``maze_poisson/myio/input/__init__.py:10-25``)."""
    ext = os.path.splitext(path)[1]
    with open(path, "r", encoding="utf-8") as f:
        if ext in (".yml", ".yaml"):
            import yaml
            data = yaml.safe_load(f)
        elif ext == ".json":
            data = json.load(f)
        else:
            raise ValueError(f"File extension {ext} not supported")
    return from_dict(data)

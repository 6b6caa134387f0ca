"""Converter for the old input schema and writers of the companion input files.

The reference's example inputs under ``examples/input_files`` and most of
``examples/restart_files`` use an older schema, ``charge,mass[,radius],x,y,z[,vx,vy,vz]``
(written by ``maze_poisson/restart.py:8-52``), which the current driver no longer reads
(it needs a ``type`` column, ``maze_poisson/solver.py:304-313``).  :func:`convert_old_input`
turns such a file into the current pair: an input file ``type,x,y,z[,vx,vy,vz]`` and a
particles file ``type,charge,mass[,radius]``.  Species are the distinct
(charge, mass[, radius]) rows in order of first appearance; they are named Na / Cl
for the (+1, 22.99) / (-1, 35.45x) rows of molten NaCl and T0, T1, ... otherwise.

:func:`write_tosi_fumi` writes the potential file for NaCl from the Fumi-Tosi table of
:mod:`.workloads` (the reference ships none).
"""
from __future__ import annotations

import os

from .outputs import _fmt
from .solver import read_csv_columns
from .workloads import TOSI_FUMI_NACL


def _species_name(charge: float, mass: float, taken: set) -> str:
    if charge == 1.0 and abs(mass - 22.99) < 0.01 and "Na" not in taken:
        return "Na"
    if charge == -1.0 and abs(mass - 35.45) < 0.01 and "Cl" not in taken:
        return "Cl"
    i = 0
    while f"T{i}" in taken:
        i += 1
    return f"T{i}"


def convert_old_input(src: str, input_out: str, species_out: str) -> dict[str, tuple]:
    """Old-schema ``src`` -> ``input_out`` (type, x, y, z[, v]) and ``species_out``
    (type, charge, mass[, radius]).  Returns {name: (charge, mass[, radius])}.
    Values are copied as text, so positions and velocities are unchanged."""
    cols = read_csv_columns(src)
    for c in ("charge", "mass", "x", "y", "z"):
        if c not in cols:
            raise ValueError(f"{src}: not an old-schema input (missing column {c})")
    has_r = "radius" in cols
    has_v = all(c in cols for c in ("vx", "vy", "vz"))
    keys = []
    for i in range(len(cols["charge"])):
        k = (float(cols["charge"][i]), float(cols["mass"][i]))
        if has_r:
            k += (float(cols["radius"][i]),)
        keys.append(k)
    names: dict[tuple, str] = {}
    for k in keys:
        if k not in names:
            names[k] = _species_name(k[0], k[1], set(names.values()))
    vcols = ["vx", "vy", "vz"] if has_v else []
    with open(input_out, "w", encoding="utf-8") as f:
        f.write(",".join(["type", "x", "y", "z"] + vcols) + "\n")
        for i, k in enumerate(keys):
            f.write(",".join([names[k]] + [cols[c][i] for c in ["x", "y", "z"] + vcols]) + "\n")
    with open(species_out, "w", encoding="utf-8") as f:
        f.write(",".join(["type", "charge", "mass"] + (["radius"] if has_r else [])) + "\n")
        for k, name in names.items():
            f.write(",".join([name] + [_fmt(v) for v in k]) + "\n")
    return {name: k for k, name in names.items()}


def write_tosi_fumi(path: str, species=("Na", "Cl")) -> None:
    """Potential file (type1, type2, A, B, C, D, sigma) for the NaCl pairs."""
    with open(path, "w", encoding="utf-8") as f:
        f.write("type1,type2,A,B,C,D,sigma\n")
        for (a, b), vals in TOSI_FUMI_NACL.items():
            if a in species and b in species:
                f.write(",".join([a, b] + [_fmt(v) for v in vals]) + "\n")


def write_species(path: str, rows: list[tuple]) -> None:
    """Particles file from (type, charge, mass[, radius]) rows."""
    has_r = any(len(r) > 3 for r in rows)
    with open(path, "w", encoding="utf-8") as f:
        f.write("type,charge,mass" + (",radius" if has_r else "") + "\n")
        for r in rows:
            f.write(",".join(_fmt(v) for v in r) + "\n")


def main(argv=None) -> int:
    import argparse
    ap = argparse.ArgumentParser(prog="python -m maze_poisson_amd convert",
                                 description=__doc__.split("\n\n")[0])
    ap.add_argument("src", help="old-schema CSV (charge,mass[,radius],x,y,z[,vx,vy,vz])")
    ap.add_argument("outdir", help="folder for input.csv, species.csv and tosi_fumi.csv")
    a = ap.parse_args(argv)
    os.makedirs(a.outdir, exist_ok=True)
    names = convert_old_input(a.src, os.path.join(a.outdir, "input.csv"),
                              os.path.join(a.outdir, "species.csv"))
    if set(names) <= {"Na", "Cl"}:
        write_tosi_fumi(os.path.join(a.outdir, "tosi_fumi.csv"), tuple(names))
    print(f"species: {names}")
    return 0

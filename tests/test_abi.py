"""C-ABI conformance (no GPU needed): the HIP library loads, exports every function
include/maze_poisson.h declares, and covers the reference's ctypes registry
(tests/golden/reference_abi.json, parsed from maze_poisson/c_api/*.py) with the same
arity; the oracle exports the same reference symbols."""
import ctypes
import json
import os
import re

import pytest

from maze_poisson_amd.capi import EXTENSION_ABI, REFERENCE_ABI

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "maze_poisson.h")
REF_ABI = json.load(open(os.path.join(ROOT, "tests", "golden", "reference_abi.json")))


def header_decls():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    decls = {}
    for m in re.finditer(r"^\s*[a-z_0-9 ]+?\**\s*\b([a-z_0-9]+)\s*\(([^;]*?)\)\s*;", text, re.M):
        args = m.group(2).strip()
        n = 0 if args in ("void", "") else args.count(",") + 1
        decls[m.group(1)] = n
    return decls


def test_header_parses():
    d = header_decls()
    assert "solver_initialize_particles" in d and d["solver_initialize_particles"] == 13
    assert d["solver_initialize_grid"] == 8 and d["solver_update_field"] == 0


def test_library_exports_header(hip):
    lib = hip.lib
    missing = [name for name in header_decls() if not hasattr(lib, name)]
    assert not missing, missing


def test_reference_abi_covered():
    d = header_decls()
    for e in REF_ABI:
        assert e["name"] in d, e["name"]
        assert d[e["name"]] == len(e["argtypes"]), (e["name"], d[e["name"]], e["argtypes"])


def test_binding_matches_reference_registry():
    mine = {name: (len(args)) for name, _, args in REFERENCE_ABI}
    for e in REF_ABI:
        assert mine[e["name"]] == len(e["argtypes"]), e["name"]
    assert len(mine) == len(REF_ABI)


def test_oracle_exports_reference_abi(oracle):
    for e in REF_ABI:
        assert oracle.has(e["name"]), e["name"]


def test_extensions_declared():
    d = header_decls()
    for name, _, args in EXTENSION_ABI:
        assert name in d and d[name] == len(args), name


def test_enum_tables_without_gpu(hip):
    maps = hip.type_maps()
    assert maps["grid"] == {"LCG": 0, "FFT": 1, "MULTIGRID": 2, "MAZE-LCG": 3, "MAZE-MULTIGRID": 4}
    assert maps["precond"]["JACOBI"] == 1 and maps["integrator"] == {"OVRVO": 0, "VERLET": 1}
    assert hip.init_mpi() == 0 and hip.get_rank() == 0


@pytest.mark.skipif(not os.path.isdir("/root/reference/maze_poisson/c_api"),
                    reason="reference tree not present")
def test_reference_abi_fixture_is_current():
    import importlib.util
    spec = importlib.util.spec_from_file_location(
        "make_golden", os.path.join(ROOT, "tests", "golden", "make_golden.py"))
    mg = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mg)
    assert mg.abi_fixture() == REF_ABI

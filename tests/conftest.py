"""Shared fixtures: the HIP library (product), the CPU oracle and, where it was built
from the reference sources, the reference C library (both test infrastructure)."""
import os
import sys

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

from maze_poisson_amd.capi import MazeLib  # noqa: E402

ORACLE_LIB = os.path.join(ROOT, "oracle", "liboracle.so")
REF_LIB = os.path.join(ROOT, "oracle", "_ref", "libmaze_ref.so")
HIP_LIB = os.path.join(ROOT, "maze_poisson_amd", "libmaze_poisson.so")
GOLDEN = os.path.join(ROOT, "tests", "golden")


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (runs the HIP library)")


@pytest.fixture(scope="session")
def oracle():
    if not os.path.exists(ORACLE_LIB):
        import subprocess
        subprocess.check_call(["make", "-s", "-C", os.path.join(ROOT, "oracle")])
    return MazeLib(ORACLE_LIB)


@pytest.fixture(scope="session")
def oracle_raw(oracle):
    """The oracle's kernel-level C functions, with argtypes."""
    import ctypes
    lib = oracle.lib
    D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    L = np.ctypeslib.ndpointer(dtype=np.int64, flags="C_CONTIGUOUS")
    I = np.ctypeslib.ndpointer(dtype=np.int32, flags="C_CONTIGUOUS")
    ci, cd = ctypes.c_int, ctypes.c_double
    sig = {
        "or_laplace": (None, [D, D, ci]),
        "or_laplace_pb": (None, [D, D, ci, D, D, D, D]),
        "or_conj_grad": (ci, [D, D, cd, ci]),
        "or_conj_grad_jacobi": (ci, [D, D, cd, ci]),
        "or_conj_grad_pb": (ci, [D, D, cd, ci, D, D, D, D]),
        "or_verlet_poisson": (ci, [cd, cd, D, D, D, D, ci, ci]),
        "or_mg_smooth_rbgs": (None, [D, D, ci, ci]),
        "or_mg_restrict27": (None, [D, D, ci]),
        "or_mg_prolong_trilinear": (None, [D, D, ci]),
        "or_mg_cg_coarse": (ci, [D, D, ci, ci, cd]),
        "or_mg_vcycle": (None, [D, D, ci, ci]),
        "or_blockjacobi_apply": (None, [D, D, ci]),
        "or_conj_grad_bj": (ci, [D, D, cd, ci]),
        "or_mg_smooth_pb_rbgs": (None, [D, D, ci, ci, D, D, D, D]),
        "or_mg_restrict_eps": (None, [D, D, ci, ci]),
        "or_mg_restrict_k2": (None, [D, D, ci]),
        "or_mg_cg_coarse_pb": (ci, [D, D, ci, ci, cd, D, D, D, D]),
        "or_mg_vcycle_pb": (None, [D, D, ci, ci, D, D, D, D]),
        "or_verlet_pb_multigrid": (ci, [cd, cd, D, D, D, D, ci, D, D, D, D]),
        "or_mg_pb_update": (ci, [cd, D, D, ci, D, D, D, D]),
        "or_cic_weight": (cd, [cd, cd, cd]),
        "or_cic_neighbors": (None, [D, ci, ci, cd, L]),
        "or_spread_cic": (cd, [ci, ci, cd, D, L, D, D]),
        "or_force_fd": (cd, [ci, ci, cd, D, L, D, D, D]),
        "or_tf_pair_table": (None, [ci, D, cd, D]),
        "or_tf_forces": (cd, [ci, cd, D, I, ci, D, cd, D]),
        "or_tf_neighbor_counts": (None, [ci, cd, D, cd, I]),
        "or_kinetic": (cd, [ci, D, D]),
        "or_pb_maps": (None, [ci, cd, cd, cd, cd, cd, ci, D, D, D, D, D, D]),
        "or_pb_forces": (cd, [ci, cd, cd, cd, cd, ci, cd, ci, D, D, D, D, D, D, D, D, D, D]),
        "or_normal3": (None, [ctypes.c_uint64, ctypes.c_uint64, ctypes.c_uint32, D]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    p4 = ctypes.POINTER(ctypes.c_uint32 * 4)
    lib.or_philox4x32.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_void_p]
    lib.or_philox4x32.restype = None
    return lib


@pytest.fixture(scope="session")
def ref():
    if not os.path.exists(REF_LIB):
        pytest.skip("reference C library not built (needs /root/reference; `make -C oracle ref`)")
    return MazeLib(REF_LIB, extensions=False)


@pytest.fixture(scope="session")
def hip():
    """The product library; fails (never skips) when it has not been built."""
    from maze_poisson_amd import load_library
    return load_library()


@pytest.fixture(scope="session")
def hip_raw(hip):
    import ctypes
    lib = hip.lib
    D = np.ctypeslib.ndpointer(dtype=np.float64, flags="C_CONTIGUOUS")
    F = np.ctypeslib.ndpointer(dtype=np.float32, flags="C_CONTIGUOUS")
    ci, cd = ctypes.c_int, ctypes.c_double
    sig = {
        "maze_test_laplace": (None, [D, D, ci]),
        "maze_test_laplace_pb": (None, [D, D, ci, D, D, D, D]),
        "maze_test_conj_grad": (ci, [D, D, cd, ci]),
        "maze_test_conj_grad_pb": (ci, [D, D, cd, ci, D, D, D, D]),
        "maze_test_rbgs": (None, [D, D, ci, ci]),
        "maze_test_restrict27": (ci, [D, D, ci]),
        "maze_test_prolong_add": (ci, [D, D, ci]),
        "maze_test_mg_vcycle": (ci, [D, D, ci]),
        "maze_test_blockjacobi": (ci, [D, D, ci]),
        "maze_test_rbgs_pb": (None, [D, D, ci, ci, D, D, D, D]),
        "maze_test_restrict_pb_coef": (ci, [D, D, D, D, D, D, D, D, ci]),
        "maze_test_mg_vcycle_pb": (ci, [D, D, ci, D, D, D, D]),
        "maze_test_mg_pcg": (ci, [D, D, cd, ci]),
        "maze_test_mg_down": (ci, [F, F, F, ci, cd, ci]),
        "maze_test_mg_up": (ci, [F, F, F, F, ci, cd, ci]),
        "maze_test_mg_down2": (ci, [F, F, F, ci, cd, cd, ci]),
        "maze_test_mg_up2": (ci, [F, F, F, F, ci, cd, cd, ci]),
        "maze_test_mg_tail": (ci, [F, F, ci]),
        "maze_test_mg_coarse": (ci, [F, F, ci]),
        "maze_test_pb_mg_pcg": (ci, [D, D, cd, ci, D, D, D, D]),
        "solver_set_mg_omega": (None, [cd]),
        "solver_mg_levels": (ci, []),
        "solver_set_precision": (ci, [ci]),
        "solver_set_inner_rtol": (None, [cd]),
        "maze_prof_enable": (None, [ci]),
        "maze_prof_reset": (None, []),
        "maze_prof_get": (ci, [ctypes.c_char_p, ctypes.POINTER(cd), ctypes.POINTER(ctypes.c_long)]),
    }
    for name, (res, args) in sig.items():
        f = getattr(lib, name)
        f.restype = res
        f.argtypes = args
    return lib

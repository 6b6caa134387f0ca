"""ctypes binding of the MaZe C-ABI.

Mirrors the reference's registry (``maze_poisson/c_api/__init__.py:24-101``,
``maze_poisson/c_api/solver.py:8-176``, ``maze_poisson/c_api/get_functions.py:9-194``,
``maze_poisson/c_api/mpi_base.py:8-9``): every symbol is registered with an explicit
restype and argtypes.  One :class:`MazeLib` can load any library that exports this
ABI -- the HIP library ``libmaze_poisson.so`` of this package, and, for testing,
the CPU oracle or the reference C compiled from source -- so the same driver code
runs against each.

Unlike the reference's ``SafeCallable`` (which logs and calls ``exit()``), a missing
library or symbol raises immediately: the product path must never fall back
silently.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np
import numpy.ctypeslib as npct

_D1 = npct.ndpointer(dtype=np.float64, ndim=1, flags="C_CONTIGUOUS")
_D2 = npct.ndpointer(dtype=np.float64, ndim=2, flags="C_CONTIGUOUS")
_D3 = npct.ndpointer(dtype=np.float64, ndim=3, flags="C_CONTIGUOUS")
_I1 = npct.ndpointer(dtype=np.int32, ndim=1, flags="C_CONTIGUOUS")
_L3 = npct.ndpointer(dtype=np.int64, ndim=3, flags="C_CONTIGUOUS")
_c_int, _c_double = ctypes.c_int, ctypes.c_double

# (name, restype, argtypes): the reference's ABI, in the order of its binding files.
REFERENCE_ABI = [
    ("solver_initialize", None, []),
    ("solver_initialize_grid", None, [_c_int, _c_double, _c_double, _c_double, _c_double,
                                      _c_double, _c_int, _c_int]),
    ("solver_initialize_grid_pois_boltz", None, [_c_double, _c_double, _c_int]),
    ("solver_initialize_particles", None, [_c_int, _c_int, _c_double, _c_double, _c_int, _c_int,
                                           _c_int, _I1, _D2, _D2, _D1, _D1, _D1]),
    ("solver_initialize_particles_pois_boltz", None, [_c_double, _c_double, _D1]),
    ("solver_initialize_integrator", None, [_c_int, _c_double, _c_double, _c_double, _c_int,
                                            _c_int]),
    ("solver_update_charges", _c_int, []),
    ("solver_init_field", None, []),
    ("solver_set_field", None, [_D3]),
    ("solver_set_field_prev", None, [_D3]),
    ("solver_update_field", _c_int, []),
    ("solver_update_eps_k2", None, []),
    ("solver_compute_forces_elec", None, []),
    ("solver_compute_forces_noel", _c_double, []),
    ("solver_compute_forces_pb", _c_double, []),
    ("solver_compute_forces_tot", None, []),
    ("integrator_part_1", None, []),
    ("integrator_part_2", None, []),
    ("solver_check_thermostat", _c_int, []),
    ("solver_rescale_velocities", None, []),
    ("solver_finalize", None, []),
    ("get_grid_type_num", _c_int, []),
    ("get_grid_type_str", ctypes.c_char_p, [_c_int]),
    ("get_potential_type_num", _c_int, []),
    ("get_potential_type_str", ctypes.c_char_p, [_c_int]),
    ("get_ca_scheme_type_num", _c_int, []),
    ("get_ca_scheme_type_str", ctypes.c_char_p, [_c_int]),
    ("get_integrator_type_num", _c_int, []),
    ("get_integrator_type_str", ctypes.c_char_p, [_c_int]),
    ("get_precond_type_num", _c_int, []),
    ("get_precond_type_str", ctypes.c_char_p, [_c_int]),
    ("get_pos", None, [_D2]),
    ("get_vel", None, [_D2]),
    ("get_fcs_elec", None, [_D2]),
    ("get_fcs_noel", None, [_D2]),
    ("get_fcs_db", None, [_D2]),
    ("get_fcs_ib", None, [_D2]),
    ("get_fcs_np", None, [_D2]),
    ("get_fcs_tot", None, [_D2]),
    ("get_charges", None, [_D1]),
    ("get_types", None, [_I1]),
    ("get_masses", None, [_D1]),
    ("get_radii", None, [_D1]),
    ("get_field", None, [_D3]),
    ("get_field_prev", None, [_D3]),
    ("get_q", None, [_D3]),
    ("get_kinetic_energy", _c_double, []),
    ("get_energy_elec", _c_double, []),
    ("get_momentum", None, [_D1]),
    ("get_temperature", _c_double, []),
    ("init_mpi", _c_int, []),
    ("get_rank", _c_int, []),
]

# Extensions this build adds beside the reference ABI (declared in include/maze_poisson.h).
EXTENSION_ABI = [
    ("solver_md_loop_iter", None, []),
    ("solver_run_n_steps", None, [_c_int]),
    ("solver_set_rng_seed", None, [ctypes.c_uint64]),
    ("solver_set_eps_k2", None, [_D3, _D3, _D3, _D3]),
    ("get_eps_k2", None, [_D3, _D3, _D3, _D3]),
    ("get_y", None, [_D3]),
    ("get_neighbors", None, [_L3]),
    ("get_tf_neighbor_counts", None, [_I1]),
    ("solver_sync", None, []),
    ("solver_set_precision", _c_int, [_c_int]),
    ("solver_get_last_passes", _c_int, [_I1, _c_int]),
    ("solver_get_precision", _c_int, []),
    ("solver_set_inner_rtol", None, [_c_double]),
    ("solver_set_y_extrapolation", None, [_c_int]),
    ("solver_set_mg_omegas", None, [_c_double, _c_double]),
    ("get_energy_elec_qphi", _c_double, []),
    ("solver_set_mg_omega", None, [_c_double]),
    ("solver_mg_levels", _c_int, []),
    ("solver_set_charge_check", None, [_c_int]),
    # the replica engine (many independent replicas on one GPU)
    ("maze_engine_create", ctypes.c_void_p, [ctypes.c_longlong, _c_int]),
    ("maze_engine_destroy", None, [ctypes.c_void_p]),
    ("maze_engine_set_threads", None, [ctypes.c_void_p, _c_int]),
    ("maze_engine_set_graphs", None, [ctypes.c_void_p, _c_int]),
    ("maze_engine_set_mixed_init", None, [ctypes.c_void_p, _c_int]),
    ("solver_set_mixed_init", None, [_c_int]),
    ("maze_copy_bandwidth", _c_double, [ctypes.c_long, _c_int]),
    ("maze_engine_estimate_bytes", ctypes.c_longlong, [_c_int, _c_int, _c_int]),
    ("maze_engine_precision_for", _c_int, [_c_int, _c_int]),
    ("maze_engine_submit", _c_int, [ctypes.c_void_p, _c_double, _c_int, _c_double, _c_double,
                                    _c_int, _c_int, _c_int, _I1, _D2, _D2, _D1, _D1, _D1,
                                    _c_double, _c_int]),
    ("maze_engine_set_physics", _c_int, [ctypes.c_void_p, _c_int, _c_int, _c_int, _c_int,
                                         _c_double, _c_double, _c_int, _c_double, _c_double,
                                         _c_int, _c_int, ctypes.c_ulonglong, _c_int]),
    ("maze_engine_set_pb_maps", _c_int, [ctypes.c_void_p, _c_int, _c_double, _D3, _D3, _D3,
                                         _D3]),
    ("maze_engine_set_pb_radii", _c_int, [ctypes.c_void_p, _c_int, _c_double, _c_double, _c_int,
                                          _c_double, _c_double, _D1]),
    ("maze_engine_keep_field", _c_int, [ctypes.c_void_p, _c_int, _c_int]),
    ("maze_engine_cancel", _c_int, [ctypes.c_void_p, _c_int]),
    ("maze_engine_run", _c_int, [ctypes.c_void_p]),
    ("maze_engine_observables", _c_int, [ctypes.c_void_p, _c_int, _D1, _c_int]),
    ("maze_engine_field", _c_int, [ctypes.c_void_p, _c_int, _D3]),
    ("maze_engine_result", _c_int, [ctypes.c_void_p, _c_int, _D2, _D2, _D2]),
    ("maze_engine_request_stats", _c_int, [ctypes.c_void_p, _c_int, _D1, _c_int]),
    ("maze_engine_stats", _c_int, [ctypes.c_void_p, _D1, _c_int]),
]


class MazeLib:
    """A loaded library exporting the MaZe ABI.  Attribute access returns the bound
    ctypes function; unknown or missing symbols raise ``AttributeError``."""

    def __init__(self, path: str, extensions: bool = True):
        if not os.path.exists(path):
            raise FileNotFoundError(f"MaZe library not found: {path}")
        self.path = path
        self.lib = ctypes.CDLL(os.path.abspath(path))
        self.functions = {}
        table = REFERENCE_ABI + (EXTENSION_ABI if extensions else [])
        for name, restype, argtypes in table:
            try:
                f = getattr(self.lib, name)
            except AttributeError:
                continue
            f.restype = restype
            f.argtypes = argtypes
            self.functions[name] = f

    def has(self, name: str) -> bool:
        return name in self.functions

    def __getattr__(self, name):
        funcs = self.__dict__.get("functions", {})
        if name in funcs:
            return funcs[name]
        raise AttributeError(f"{self.__dict__.get('path')}: symbol {name} not exported")

    def type_maps(self):
        """Enum discovery as in ``maze_poisson/solver.py:115-127``."""
        maps = {}
        for key, fnum, fstr in [
            ("grid", "get_grid_type_num", "get_grid_type_str"),
            ("potential", "get_potential_type_num", "get_potential_type_str"),
            ("cas", "get_ca_scheme_type_num", "get_ca_scheme_type_str"),
            ("integrator", "get_integrator_type_num", "get_integrator_type_str"),
            ("precond", "get_precond_type_num", "get_precond_type_str"),
        ]:
            n = getattr(self, fnum)()
            maps[key] = {getattr(self, fstr)(i).decode().upper(): i for i in range(n)}
        return maps

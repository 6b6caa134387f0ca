"""MI355X-native MaZe electrostatics + MD hot path.

The product is the HIP shared library ``libmaze_poisson.so`` (sources in ``csrc/``),
which exports the reference's C-ABI; this package holds its ctypes binding
(:mod:`.capi`), the reference driver's call sequence (:mod:`.driver`), synthetic
workloads (:mod:`.workloads`) and the build recipe (:mod:`.build`).
"""
import os

PKG_DIR = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(PKG_DIR, "libmaze_poisson.so")


def load_library():
    """Load the HIP library; raises if it has not been built (no CPU fallback)."""
    from .capi import MazeLib
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(
            f"libmaze_poisson.so is missing at {LIB_PATH}; build it with "
            "`python -c 'import __graft_entry__ as g; g.build()'`")
    return MazeLib(LIB_PATH)

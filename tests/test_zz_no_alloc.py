"""The allocation-free CG entry point (maze_conj_grad_noalloc: FP64 implicit-Jacobi CG on
caller-owned device buffers, work space in a static device array) against the
allocating entry point maze_test_conj_grad on the same inputs: same iteration index and
bit-identical solutions, on two calls in a row (the second on a smaller grid, so it
reuses the array the first one left dirty); sizes beyond the maximum are refused.

Device buffers come from the HIP runtime the library itself links (through ctypes), so
the test holds no other GPU runtime."""
import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

HIP_RUNTIME = "/opt/rocm/lib/libamdhip64.so.7"
H2D, D2H = 1, 2


class DeviceBuffer:
    def __init__(self, rt, nbytes):
        self.rt, self.nbytes = rt, nbytes
        self.ptr = ctypes.c_void_p()
        assert rt.hipMalloc(ctypes.byref(self.ptr), ctypes.c_size_t(nbytes)) == 0

    def put(self, a):
        a = np.ascontiguousarray(a, dtype=np.float64)
        assert a.nbytes == self.nbytes
        assert self.rt.hipMemcpy(self.ptr, a.ctypes.data_as(ctypes.c_void_p),
                                 ctypes.c_size_t(a.nbytes), H2D) == 0

    def get(self, shape):
        a = np.zeros(shape)
        assert self.rt.hipMemcpy(a.ctypes.data_as(ctypes.c_void_p), self.ptr,
                                 ctypes.c_size_t(a.nbytes), D2H) == 0
        return a

    def free(self):
        assert self.rt.hipFree(self.ptr) == 0


@pytest.fixture(scope="module")
def noalloc(hip):
    f = hip.lib.maze_conj_grad_noalloc
    f.restype = ctypes.c_int
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_double, ctypes.c_int,
                  ctypes.c_void_p]
    rt = ctypes.CDLL(HIP_RUNTIME)
    for name in ("hipMalloc", "hipMemcpy", "hipFree", "hipDeviceSynchronize"):
        getattr(rt, name).restype = ctypes.c_int
    rt.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
    rt.hipFree.argtypes = [ctypes.c_void_p]
    return f, rt


def solve_both(hip_raw, noalloc, n, seed):
    f, rt = noalloc
    g = np.random.default_rng(seed)
    b = g.standard_normal((n, n, n))
    b -= b.mean()
    x0 = 0.1 * g.standard_normal((n, n, n))
    tol = 1e-9 * np.abs(b).max()
    xa = x0.copy()
    ia = hip_raw.maze_test_conj_grad(b, xa, tol, n)
    db, dx = DeviceBuffer(rt, b.nbytes), DeviceBuffer(rt, b.nbytes)
    try:
        db.put(b)
        dx.put(x0)
        ib = f(db.ptr, dx.ptr, tol, n, None)
        assert rt.hipDeviceSynchronize() == 0
        xb = dx.get(b.shape)
    finally:
        db.free()
        dx.free()
    return (ia, xa), (ib, xb)


def test_noalloc_cg_matches_allocating_entry_point(hip_raw, noalloc):
    for n, seed in ((48, 1), (24, 2)):  # the second call is smaller and reuses the array
        (ia, xa), (ib, xb) = solve_both(hip_raw, noalloc, n, seed)
        assert ia >= 0 and ib == ia, (n, ia, ib)
        assert np.array_equal(xa, xb), n


def test_noalloc_cg_refuses_sizes_beyond_its_array(noalloc):
    f, _ = noalloc
    # Refused before any device access, so the pointers are never dereferenced.
    assert f(ctypes.c_void_p(16), ctypes.c_void_p(16), 1e-9, 65, None) == -2
    assert f(ctypes.c_void_p(16), ctypes.c_void_p(16), 1e-9, 1, None) == -2

"""The allocation-free exact repulsion (fa2_gpu_repulse_exact_noalloc of
include/fa2_hip.h) against the session entry point fa2_gpu_repulse_exact,
through the C ABI.

Gates: on two calls in a row, the second with fewer nodes (the scratch keeps
the first call's contents), in single and double precision, with and
without node sizes, the forces added into the caller's device arrays equal
bit for bit the forces fa2_gpu_repulse_exact adds into host arrays (same
kernels, same summation order); they are within 1e-5 (single) and 1e-12
(double) normwise of the oracle's exact FP64 forces; sizes beyond the
maximum are refused with a message. Device memory for the inputs and
outputs is allocated here, by the caller, through the HIP runtime the
library itself uses.
"""
import ctypes

import numpy as np
import pytest

from forceatlas2_rc_amd import graphs
from forceatlas2_rc_amd._lib import last_error, lib
from oracle import oracle as O

pytestmark = pytest.mark.gpu

hip = ctypes.CDLL("libamdhip64.so.7")  # the runtime libfa2hip.so is linked with
hip.hipMalloc.argtypes = [ctypes.POINTER(ctypes.c_void_p), ctypes.c_size_t]
hip.hipFree.argtypes = [ctypes.c_void_p]
hip.hipMemcpy.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t, ctypes.c_int]
hip.hipDeviceSynchronize.argtypes = []
H2D, D2H = 1, 2


class DeviceArrays:
    """Device copies of host float64 arrays, freed on exit."""

    def __init__(self, **arrays):
        self.host = {k: np.ascontiguousarray(v, np.float64) for k, v in arrays.items()}
        self.dev = {}

    def __enter__(self):
        for k, a in self.host.items():
            p = ctypes.c_void_p()
            assert hip.hipMalloc(ctypes.byref(p), max(8, a.nbytes)) == 0
            self.dev[k] = p
            assert hip.hipMemcpy(p, a.ctypes.data, a.nbytes, H2D) == 0
        return self

    def get(self, k):
        out = np.empty_like(self.host[k])
        assert hip.hipDeviceSynchronize() == 0
        assert hip.hipMemcpy(out.ctypes.data, self.dev[k], out.nbytes, D2H) == 0
        return out

    def __exit__(self, *exc):
        hip.hipDeviceSynchronize()
        for p in self.dev.values():
            hip.hipFree(p)


def _session(n, use_double, adjust, x, y, m, s, scaling, dx, dy):
    ptr = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    g = lib.fa2_gpu_open(0, use_double, n, adjust, 0, 1)
    assert g, last_error()
    try:
        assert lib.fa2_gpu_repulse_exact(g, ptr(x), ptr(y), ptr(m), ptr(s), scaling,
                                         ptr(dx), ptr(dy)) == 0, last_error()
    finally:
        lib.fa2_gpu_close(g)


def _case(n, seed, adjust):
    n, e = graphs.barabasi_albert(n, 3, seed=seed)
    rng = np.random.default_rng(seed)
    pos = rng.uniform(-10 * np.sqrt(n), 10 * np.sqrt(n), (n, 2))
    mass = 1.0 + np.bincount(e[:, 0], minlength=n) + np.bincount(e[:, 1], minlength=n)
    size = rng.uniform(0.5, 3.0, n) if adjust else None
    # the forces are added to what the arrays hold
    dx0, dy0 = rng.normal(0, 100, n), rng.normal(0, 100, n)
    return n, pos[:, 0].copy(), pos[:, 1].copy(), mass.astype(np.float64), size, dx0, dy0


@pytest.mark.parametrize("use_double", [0, 1])
@pytest.mark.parametrize("adjust", [0, 1])
def test_noalloc_equals_session_entry_point(use_double, adjust):
    for n0, seed in ((20000, 3), (4100, 4)):  # the second call is smaller
        n, x, y, m, s, dx0, dy0 = _case(n0, seed, adjust)
        dx_ref, dy_ref = dx0.copy(), dy0.copy()
        _session(n, use_double, adjust, x, y, m, s, 2.0, dx_ref, dy_ref)
        arrays = dict(x=x, y=y, m=m, dx=dx0, dy=dy0)
        if adjust:
            arrays["s"] = s
        with DeviceArrays(**arrays) as d:
            st = lib.fa2_gpu_repulse_exact_noalloc(
                0, use_double, n, adjust, d.dev["x"], d.dev["y"], d.dev["m"],
                d.dev["s"] if adjust else None, 2.0, d.dev["dx"], d.dev["dy"], None)
            assert st == 0, last_error()
            dx, dy = d.get("dx"), d.get("dy")
        assert np.array_equal(dx, dx_ref), n
        assert np.array_equal(dy, dy_ref), n
        if not adjust:
            xs = x.astype(np.float32).astype(np.float64) if not use_double else x
            ys = y.astype(np.float32).astype(np.float64) if not use_double else y
            want = O.repulse_exact(xs, ys, m, 2.0)
            got = np.stack([dx - dx0, dy - dy0], 1)
            tol = 1e-12 if use_double else 1e-5
            assert np.abs(got - want).max() / np.abs(want).max() <= tol


def test_noalloc_refuses_sizes_beyond_the_maximum():
    one = np.zeros(1)
    with DeviceArrays(a=one) as d:
        p = d.dev["a"]
        assert lib.fa2_gpu_repulse_exact_noalloc(0, 0, (1 << 18) + 1, 0, p, p, p, None, 2.0,
                                                 p, p, None) == 1
        assert "262144" in last_error()
        assert lib.fa2_gpu_repulse_exact_noalloc(0, 0, 0, 0, p, p, p, None, 2.0, p, p,
                                                 None) == 1
        assert lib.fa2_gpu_repulse_exact_noalloc(0, 0, 10, 1, p, p, p, None, 2.0, p, p,
                                                 None) == 1
        assert "size" in last_error()

"""ORACLE — test infrastructure only.

ctypes wrapper of liboracle.so, the C restatement of the reference
ForceAtlas2 algorithms (fa2_oracle.c). Built by ``make -C oracle``.
"""
import ctypes
import os
from ctypes import POINTER, Structure, c_double, c_int, c_uint32, c_void_p

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liboracle.so")


class LargeParams(Structure):
    """orc_large_params (fa2_oracle.h)."""
    _fields_ = [("n", c_int), ("inc_off", POINTER(c_int)),
                ("inc_node", POINTER(c_int)), ("inc_coef", POINTER(c_double)),
                ("mass", POINTER(c_double)), ("size", POINTER(c_double)),
                ("fixed", POINTER(c_int)), ("any_fixed", c_int),
                ("scaling", c_double), ("gravity", c_double),
                ("theta", c_double), ("jitter_tolerance", c_double),
                ("strong", c_int), ("linlog", c_int), ("adjust", c_int)]


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(f"{LIB_PATH} is missing: run make -C oracle")
    lib = ctypes.CDLL(LIB_PATH)
    V = c_void_p
    protos = {
        "orc_now": (c_double, []),
        "orc_has_openmp": (c_int, []),
        "orc_update_speed": (None, [c_int, c_double, c_double, c_double,
                                    POINTER(c_double), POINTER(c_double)]),
        "orc_gephi_tree": (c_int, [c_int, V, V, V, V, POINTER(c_int), V, V, V,
                                   V, V, V, V]),
        "orc_layout": (c_int, [c_int, c_int, V, V, V, V, V, V, c_int, V, V,
                               c_int, V, V, V]),
        "orc_repulse_exact": (None, [c_int, V, V, V, V, c_double, c_int, V, V]),
        "orc_repulse_bh": (c_int, [c_int, V, V, V, V, c_double, c_double,
                                   c_int, V, V]),
        "orc_repulse_bh_single": (c_int, [c_int, V, V, V, V, c_double, c_double, V, V]),
        "orc_spread16": (c_uint32, [c_uint32]),
        "orc_radix_sort": (None, [c_int, V, V, V]),
        "orc_karras": (None, [c_int, V, V, V, V]),
        "orc_large_prepare": (c_int, [c_int, c_int, V, V, V, c_int, V, V, V,
                                      V]),
    }
    for s in "fdm":
        protos[f"orc_morton_{s}"] = (None, [c_int, V, V, V])
        protos[f"orc_large_tree_{s}"] = (None, [c_int, V, V, V, V, V, V, V, V])
        protos[f"orc_large_regions_{s}"] = (c_int, [c_int, V, V, V, V])
        protos[f"orc_large_run_{s}"] = (c_int, [POINTER(LargeParams), V, c_int,
                                                V, V, V, V])
    for name, (res, args) in protos.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def _p(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def update_speed(n, sw, tr, jt, speed, eff):
    s, e = c_double(speed), c_double(eff)
    lib.orc_update_speed(n, sw, tr, jt, ctypes.byref(s), ctypes.byref(e))
    return s.value, e.value


def layout(n, edges, pos, iterations, scaling=2.0, gravity=1.0, jt=1.0,
           theta=1.2, strong=False, linlog=False, outbound=False,
           adjust=False, barnes_hut=False, w=None, size=None, fixed=None,
           threads=1, trace=False):
    """CPU layout of C_forceatlas2 in FP64. pos is (n, 2); returns the final
    (n, 2) positions, plus (positions, forces) per iteration with trace."""
    e = np.asarray(edges, np.int32).reshape(-1, 2)
    frm = np.ascontiguousarray(e[:, 0])
    to = np.ascontiguousarray(e[:, 1])
    m = frm.size
    w = np.ones(m) if w is None else np.ascontiguousarray(w, np.float64)
    size = np.ones(n) if size is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(size, np.float64), (n,)))
    fixed_a = None if fixed is None else np.ascontiguousarray(fixed, np.int32)
    xy = np.ascontiguousarray(np.asarray(pos, np.float64).T).reshape(-1).copy()
    dpar = np.array([scaling, gravity, jt, theta], np.float64)
    flags = np.array([strong, linlog, outbound, adjust, barnes_hut], np.int32)
    tp = np.zeros(iterations * 2 * n) if trace else None
    tf = np.zeros(iterations * 2 * n) if trace else None
    times = np.zeros(7)
    st = lib.orc_layout(n, m, _p(frm), _p(to), _p(w), _p(xy), _p(size),
                        _p(fixed_a), iterations, _p(dpar), _p(flags), threads,
                        _p(tp), _p(tf), _p(times))
    if st != 0:
        raise RuntimeError("orc_layout failed")
    out = xy.reshape(2, n).T.copy()
    if trace:
        tp = tp.reshape(iterations, 2, n).transpose(0, 2, 1).copy()
        tf = tf.reshape(iterations, 2, n).transpose(0, 2, 1).copy()
        return out, tp, tf, times
    return out


def gephi_tree(x, y, mass, size=None):
    n = len(x)
    x, y, mass = (np.ascontiguousarray(a, np.float64) for a in (x, y, mass))
    size = None if size is None else np.ascontiguousarray(size, np.float64)
    nreg = c_int(0)
    reg_d = np.zeros(4 * 2 * n)
    reg_i = np.zeros(5 * 2 * n, np.int32)
    perm = np.zeros(n, np.int32)
    px, py, pm = np.zeros(n), np.zeros(n), np.zeros(n)
    ps = np.zeros(n) if size is not None else None
    if lib.orc_gephi_tree(n, _p(x), _p(y), _p(mass), _p(size), ctypes.byref(nreg),
                          _p(reg_d), _p(reg_i), _p(perm), _p(px), _p(py),
                          _p(pm), _p(ps)) != 0:
        raise RuntimeError("orc_gephi_tree failed")
    k = nreg.value
    return dict(nreg=k, reg_d=reg_d[:4 * k].reshape(k, 4),
                reg_i=reg_i[:5 * k].reshape(k, 5), perm=perm, px=px, py=py,
                pm=pm, ps=ps)


def repulse_exact(x, y, mass, scaling, size=None, threads=1):
    n = len(x)
    x, y, mass = (np.ascontiguousarray(a, np.float64) for a in (x, y, mass))
    size = None if size is None else np.ascontiguousarray(size, np.float64)
    dx, dy = np.zeros(n), np.zeros(n)
    lib.orc_repulse_exact(n, _p(x), _p(y), _p(mass), _p(size), scaling,
                          threads, _p(dx), _p(dy))
    return np.stack([dx, dy], 1)


def repulse_bh(x, y, mass, scaling, theta, size=None, threads=1):
    n = len(x)
    x, y, mass = (np.ascontiguousarray(a, np.float64) for a in (x, y, mass))
    size = None if size is None else np.ascontiguousarray(size, np.float64)
    dx, dy = np.zeros(n), np.zeros(n)
    if lib.orc_repulse_bh(n, _p(x), _p(y), _p(mass), _p(size), scaling, theta,
                          threads, _p(dx), _p(dy)) != 0:
        raise RuntimeError("orc_repulse_bh failed")
    return np.stack([dx, dy], 1)


def repulse_bh_single(x, y, mass, scaling, theta, size=None):
    """The FP64 model of the single-precision offload over the same Gephi
    tree: float32 inputs and records, float32 acceptance test, FP64 terms."""
    n = len(x)
    x, y, mass = (np.ascontiguousarray(a, np.float64) for a in (x, y, mass))
    size = None if size is None else np.ascontiguousarray(size, np.float64)
    dx, dy = np.zeros(n), np.zeros(n)
    if lib.orc_repulse_bh_single(n, _p(x), _p(y), _p(mass), _p(size), scaling, theta,
                                 _p(dx), _p(dy)) != 0:
        raise RuntimeError("orc_repulse_bh_single failed")
    return np.stack([dx, dy], 1)


def morton(x, y, mode="f"):
    """Morton codes as the reference computes them: mode "f" on float32
    positions in float arithmetic, "d" in double."""
    n = len(x)
    dt = np.float32 if mode == "f" else np.float64
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    key = np.zeros(n, np.uint32)
    getattr(lib, f"orc_morton_{mode}")(n, _p(x), _p(y), _p(key))
    return key


def radix_sort(keys):
    keys = np.ascontiguousarray(keys, np.uint32)
    n = keys.size
    ks = np.zeros(n, np.uint32)
    perm = np.zeros(n, np.uint32)
    lib.orc_radix_sort(n, _p(keys), _p(ks), _p(perm))
    return ks, perm


def karras(sorted_keys):
    ks = np.ascontiguousarray(sorted_keys, np.uint32)
    n = ks.size
    child = np.zeros(2 * (n - 1), np.int32)
    rng_ = np.zeros(2 * (n - 1), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    lib.orc_karras(n, _p(ks), _p(child), _p(rng_), _p(parent))
    return child, rng_, parent


def large_tree(x, y, mode="f"):
    n = len(x)
    dt = np.float32 if mode == "f" else np.float64
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    key = np.zeros(n, np.uint32)
    ks = np.zeros(n, np.uint32)
    order = np.zeros(n, np.uint32)
    child = np.zeros(2 * (n - 1), np.int32)
    rng_ = np.zeros(2 * (n - 1), np.int32)
    parent = np.zeros(2 * n - 1, np.int32)
    getattr(lib, f"orc_large_tree_{mode}")(n, _p(x), _p(y), _p(key), _p(ks),
                                           _p(order), _p(child), _p(rng_),
                                           _p(parent))
    return dict(keys_unsorted=key, keys=ks, order=order, child=child,
                range=rng_, parent=parent)


def large_regions(x, y, mass, mode="f"):
    """Traversal records of the regions of the tree of (x, y): an (n-1, 4)
    array of (centre x, centre y, mass, size^2) in the reference numbering
    (pipeline.cl:276-323); mode as for large_run."""
    n = len(x)
    dt = np.float32 if mode == "f" else np.float64
    x = np.ascontiguousarray(x, dt)
    y = np.ascontiguousarray(y, dt)
    m = np.ascontiguousarray(mass, np.float64)
    ni = np.zeros(4 * (n - 1))
    if getattr(lib, f"orc_large_regions_{mode}")(n, _p(x), _p(y), _p(m), _p(ni)) != 0:
        raise ValueError("orc_large_regions failed")
    return ni.reshape(n - 1, 4)


def large_prepare(n, edges, w=None, outbound=False):
    e = np.asarray(edges, np.int32).reshape(-1, 2)
    frm = np.ascontiguousarray(e[:, 0])
    to = np.ascontiguousarray(e[:, 1])
    m = frm.size
    w = np.ones(m) if w is None else np.ascontiguousarray(w, np.float64)
    mass = np.zeros(n)
    off = np.zeros(n + 1, np.int32)
    node = np.zeros(2 * m + 1, np.int32)
    coef = np.zeros(2 * m + 1)
    ni = lib.orc_large_prepare(n, m, _p(frm), _p(to), _p(w), int(outbound),
                               _p(mass), _p(off), _p(node), _p(coef))
    return mass, off, node[:ni].copy(), coef[:ni].copy()


def large_run(n, edges, pos, iterations, mode="m", scaling=2.0, gravity=1.0,
              jt=1.0, theta=1.2, strong=False, linlog=False, outbound=False,
              adjust=False, w=None, size=None, fixed=None, trace=False):
    """The all-GPU layout restated on the CPU (mode f / d / m)."""
    mass, off, node, coef = large_prepare(n, edges, w, outbound)
    size = np.ones(n) if size is None else np.ascontiguousarray(
        np.broadcast_to(np.asarray(size, np.float64), (n,)))
    fixed_a = np.zeros(n, np.int32) if fixed is None else \
        np.ascontiguousarray(fixed, np.int32)
    p = LargeParams()
    p.n = n
    p.inc_off = off.ctypes.data_as(POINTER(c_int))
    p.inc_node = node.ctypes.data_as(POINTER(c_int))
    p.inc_coef = coef.ctypes.data_as(POINTER(c_double))
    p.mass = mass.ctypes.data_as(POINTER(c_double))
    p.size = size.ctypes.data_as(POINTER(c_double))
    p.fixed = fixed_a.ctypes.data_as(POINTER(c_int))
    p.any_fixed = int(fixed_a.any())
    p.scaling, p.gravity, p.theta, p.jitter_tolerance = scaling, gravity, theta, jt
    p.strong, p.linlog, p.adjust = int(strong), int(linlog), int(adjust)
    xy = np.ascontiguousarray(np.asarray(pos, np.float64).T).reshape(-1).copy()
    tp = np.zeros(iterations * 2 * n) if trace else None
    tf = np.zeros(iterations * 2 * n) if trace else None
    ts = np.zeros(iterations * 4) if trace else None
    tr = np.zeros(iterations * 2 * n) if trace else None
    st = getattr(lib, f"orc_large_run_{mode}")(ctypes.byref(p), _p(xy),
                                               iterations, _p(tp), _p(tf),
                                               _p(ts), _p(tr))
    if st != 0:
        raise RuntimeError(f"orc_large_run_{mode} failed ({st})")
    out = xy.reshape(2, n).T.copy()
    if trace:
        f = lambda a: a.reshape(iterations, 2, n).transpose(0, 2, 1).copy()
        return out, dict(pos=f(tp), force=f(tf), rep=f(tr),
                         speed=ts.reshape(iterations, 4))
    return out

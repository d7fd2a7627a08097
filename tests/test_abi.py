"""The C-ABI library loads and exports every entry point its headers
declare; host-side pieces that need no device (the Gephi quadtree of the
repulsion offload, the adaptive speed, argument checks) match the oracle."""
import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from forceatlas2_rc_amd import _lib, graphs
from oracle import oracle as O

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def declared_functions():
    names = set()
    for h in ("fa2.h", "fa2_hip.h"):
        text = open(os.path.join(ROOT, "include", h)).read()
        text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
        names |= set(re.findall(r"\b(fa2_\w+)\s*\(", text))
    return names


def test_every_declared_symbol_is_exported():
    names = declared_functions()
    assert "fa2_gpu_repulse_exact" in names and "fa2_full_run" in names
    lib = ctypes.CDLL(_lib.LIB_PATH)
    missing = [n for n in sorted(names) if not hasattr(lib, n)]
    assert not missing, missing
    # and the Python binding knows each of them
    assert names <= set(_lib.PROTOTYPES)


def test_struct_layouts_match_the_c_header(tmp_path):
    """The ctypes mirrors have the layout the C compiler gives include/fa2.h."""
    src = tmp_path / "lay.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fa2.h"\n'
        "int main(void){printf(\"%zu %zu %zu %zu %zu %zu %zu\\n\","
        "sizeof(fa2_region), sizeof(fa2_node), sizeof(fa2_tree),"
        "sizeof(fa2_gpu_devinfo), sizeof(fa2_full_params),"
        "offsetof(fa2_full_params, jitter_tolerance),"
        "offsetof(fa2_full_params, interrupted));return 0;}\n")
    exe = tmp_path / "lay"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True,
                                          capture_output=True, text=True).stdout.split()]
    want = [ctypes.sizeof(_lib.Region), ctypes.sizeof(_lib.Node),
            ctypes.sizeof(_lib.Tree), ctypes.sizeof(_lib.DevInfo),
            ctypes.sizeof(_lib.FullParams),
            _lib.FullParams.jitter_tolerance.offset,
            _lib.FullParams.interrupted.offset]
    assert got == want


def test_tuning_layout_matches_the_c_header(tmp_path):
    """The ctypes mirror of fa2_tuning has the layout of include/fa2_hip.h."""
    src = tmp_path / "tun.c"
    src.write_text(
        '#include <stdio.h>\n#include <stddef.h>\n#include "fa2_hip.h"\n'
        "int main(void){printf(\"%zu %zu %zu %d\\n\", sizeof(fa2_tuning),"
        "offsetof(fa2_tuning, repulse_xcd_run), offsetof(fa2_tuning, attract_fork),"
        "FA2_TUNE_NTIMES);return 0;}\n")
    exe = tmp_path / "tun"
    subprocess.run(["gcc", "-I", os.path.join(ROOT, "include"), str(src), "-o",
                    str(exe)], check=True)
    got = [int(v) for v in subprocess.run([str(exe)], check=True,
                                          capture_output=True, text=True).stdout.split()]
    assert got == [ctypes.sizeof(_lib.Tuning), _lib.Tuning.repulse_xcd_run.offset,
                   _lib.Tuning.attract_fork.offset, _lib.FA2_TUNE_NTIMES]


def test_clock_is_monotonic():
    a = _lib.lib.fa2_now()
    b = _lib.lib.fa2_now()
    assert b >= a > 0


def host_tree(x, y, m, size=None):
    n = len(x)
    reg = (_lib.Region * (2 * n))()
    perm = np.zeros(n, np.int32)
    cnt = np.zeros(8 * n, np.int32)
    b0 = (_lib.Node * n)()
    b1 = (_lib.Node * n)()
    px, py, pm = np.zeros(n), np.zeros(n), np.zeros(n)
    ps = np.zeros(n) if size is not None else None
    t = _lib.Tree()
    t.reg = ctypes.cast(reg, ctypes.POINTER(_lib.Region))
    t.perm = perm.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    t.cnt = cnt.ctypes.data_as(ctypes.POINTER(ctypes.c_int))
    t.buf[0] = ctypes.cast(b0, ctypes.POINTER(_lib.Node))
    t.buf[1] = ctypes.cast(b1, ctypes.POINTER(_lib.Node))
    P = lambda a: a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))
    t.px, t.py, t.pm = P(px), P(py), P(pm)
    t.ps = P(ps) if ps is not None else None
    V = lambda a: None if a is None else a.ctypes.data_as(ctypes.c_void_p)
    x, y, m = (np.ascontiguousarray(a, np.float64) for a in (x, y, m))
    size = None if size is None else np.ascontiguousarray(size, np.float64)
    _lib.lib.fa2_tree_build(ctypes.byref(t), n, V(x), V(y), V(m), V(size), 1)
    keep = (reg, b0, b1, perm, cnt, px, py, pm, ps)
    return t, keep


@pytest.mark.parametrize("n,dup", [(1, 0), (2, 0), (9, 0), (500, 0), (3000, 60)])
def test_host_gephi_tree_is_the_reference_tree(n, dup):
    """The product's tree_build (host side of fa2_gpu_repulse_bh) against the
    oracle's restatement of src/forceatlas2.c:64-239, bit for bit."""
    pos = graphs.uniform_positions(n, -100, 100, n)
    if dup:
        pos[:dup] = pos[0]
    m = 1.0 + np.random.default_rng(n).integers(0, 9, n)
    size = np.random.default_rng(1).uniform(0.5, 2, n)
    t, keep = host_tree(pos[:, 0], pos[:, 1], m, size)
    ref = O.gephi_tree(pos[:, 0], pos[:, 1], m, size)
    assert t.nreg == ref["nreg"]
    regs = np.ctypeslib.as_array(keep[0])[: t.nreg]
    got_d = np.stack([regs["mass"], regs["cx"], regs["cy"], regs["size"]], 1)
    got_i = np.stack([regs["start"], regs["count"], regs["child"],
                      regs["nchild"], regs["skip"]], 1)
    assert np.array_equal(got_d, ref["reg_d"])
    assert np.array_equal(got_i, ref["reg_i"])
    assert np.array_equal(keep[3], ref["perm"])
    for a, b in zip(keep[5:9], (ref["px"], ref["py"], ref["pm"], ref["ps"])):
        assert np.array_equal(a, b)


def test_host_update_speed_matches_oracle():
    s, e = ctypes.c_double(), ctypes.c_double()
    rng = np.random.default_rng(0)
    for _ in range(300):
        n = int(rng.integers(1, 10 ** 6))
        sw, tr = rng.uniform(0, 10, 2) ** 3
        jt = rng.uniform(0.05, 3)
        sp, ef = rng.uniform(0.01, 2000), rng.uniform(0.01, 2)
        s.value, e.value = sp, ef
        _lib.lib.fa2_update_speed(n, sw, tr, jt, ctypes.byref(s), ctypes.byref(e))
        assert (s.value, e.value) == O.update_speed(n, sw, tr, jt, sp, ef)


def test_argument_errors_need_no_device():
    import forceatlas2_rc_amd as F
    with pytest.raises(ValueError, match="theta"):
        F.layout_forceatlas2_large(10, [[0, 1]], theta=-1)
    with pytest.raises(ValueError, match="does not exist"):
        F.layout_forceatlas2_large(10, [[0, 10]])
    with pytest.raises(ValueError, match="init"):
        F.layout_forceatlas2_gpu(5, [[0, 1]], fixed=np.ones(5, bool))
    with pytest.raises(ValueError, match="arg"):
        F.layout_forceatlas2_gpu(5, [[0, 1]], precision="half")
    # the drop-in entry points refuse what the reference refuses
    lib = _lib.lib
    assert lib.fa2_gpu_open(0, 0, 0, 0, 0, 1) is None
    p = _lib.FullParams()
    p.n = 1
    t = np.zeros(7)
    assert lib.fa2_full_run(ctypes.byref(p), None, t.ctypes.data_as(ctypes.c_void_p)) == 1
    assert "between 2 and 67 million" in _lib.last_error()


def test_weight_transform_identities():
    """test-interface.R:80-90: invert -> normalize -> power."""
    from forceatlas2_rc_amd import prepare_weights
    w = np.array([1.0, 2.0, 4.0, 0.0])
    assert np.array_equal(prepare_weights(4, w, 1.0, False, True),
                          [1.0, 0.5, 0.25, 0.0])
    assert np.allclose(prepare_weights(4, w, 1.0, True, False),
                       [0.25, 0.5, 1.0, 0.0])
    assert np.allclose(prepare_weights(4, w, 0.5), np.sqrt(w))
    assert np.array_equal(prepare_weights(4, w, 0.0), np.ones(4))
    assert np.array_equal(prepare_weights(3, [2.0, 2.0, 2.0], 1.0, True), np.ones(3))


def test_graph_arguments_are_checked_before_the_device():
    """fa2_batch_open and fa2_full_run refuse a graph without masses or
    incidence lists, or whose lists are inconsistent, with a message and
    before any device call (so this runs without a GPU)."""
    lib = _lib.lib
    n = 4
    mass = np.full(n, 2.0)
    off = np.array([0, 1, 2, 3, 4], np.int32)
    node = np.array([1, 0, 3, 2], np.int32)
    coef = np.ones(4)
    xy = np.zeros(2 * n)
    pos = (ctypes.c_void_p * 1)(xy.ctypes.data)
    P = ctypes.POINTER

    def params(**kw):
        p = _lib.FullParams()
        p.n = n
        p.iterations = 1
        arrs = dict(inc_off=off, inc_node=node, inc_coef=coef, mass=mass)
        arrs.update(kw)
        for k, ct in (("inc_off", ctypes.c_int), ("inc_node", ctypes.c_int),
                      ("inc_coef", ctypes.c_double), ("mass", ctypes.c_double)):
            a = arrs[k]
            setattr(p, k, None if a is None else a.ctypes.data_as(P(ct)))
        return p

    for kw in (dict(inc_off=None), dict(mass=None), dict(inc_node=None), dict(inc_coef=None),
               dict(inc_off=np.array([1, 1, 2, 3, 4], np.int32)),
               dict(inc_off=np.array([0, 0, 0, 0, -1], np.int32))):
        p = params(**kw)
        assert lib.fa2_batch_open(1, ctypes.byref(p), pos) is None, kw
        assert "incidence lists" in _lib.last_error(), kw
        t = np.zeros(7)
        assert lib.fa2_full_run(ctypes.byref(p), xy.ctypes.data_as(ctypes.c_void_p),
                                t.ctypes.data_as(ctypes.c_void_p)) == 1
        assert "incidence lists" in _lib.last_error(), kw
    # a graph without incidences needs no lists
    p = params(inc_node=None, inc_coef=None, inc_off=np.zeros(n + 1, np.int32))
    e = _lib.Edges()
    e.m = -1
    assert lib.fa2_batch_open_edges(1, ctypes.byref(p), ctypes.byref(e), pos) is None
    assert "edge list" in _lib.last_error()


def test_prepared_lists_are_coerced_and_checked():
    """Prepared mass and incidence lists of a graph dict are converted to the
    C types, and lists of the wrong length are refused, before any device
    call."""
    from forceatlas2_rc_amd.api import _fill_params, prepare_large
    n = 5
    e = np.array([[0, 1], [1, 2], [3, 4], [0, 4]])
    mass, off, node, coef = prepare_large(n, e[:, 0], e[:, 1], np.ones(4), False)
    g = dict(n=n, frm=e[:, 0], to=e[:, 1], mass=mass.astype(np.float32), inc_off=off.astype(np.int64),
             inc_node=node.astype(np.int64), inc_coef=coef.astype(np.float32), init=np.zeros((n, 2)))
    p = _lib.FullParams()
    keep, _ = _fill_params(p, g, 0, "single", False)
    assert keep[1].dtype == np.int32 and keep[3].dtype == np.float64
    assert [p.inc_off[i] for i in range(n + 1)] == off.tolist()
    assert [p.inc_coef[i] for i in range(len(coef))] == coef.tolist()
    with pytest.raises(ValueError, match="inc_off"):
        _fill_params(p, dict(g, inc_node=node[:-1]), 0, "single", False)
    with pytest.raises(ValueError, match="outside"):
        _fill_params(p, dict(g, inc_node=node + n), 0, "single", False)

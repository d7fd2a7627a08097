"""Python mirror of the GPU interface of forceatlas2r 0.3.0, over libfa2hip.so.

``layout_forceatlas2_gpu`` and ``layout_forceatlas2_large`` take the
arguments of their R namesakes (R/gpu.R:115-139, R/large.R:107-131) with an
edge list (``n``, ``edges`` of shape (m, 2), 0-based) in place of an igraph
object, and return an (n, 2) array. The argument checks and defaults are
those of ``fa2_layout`` (R/forceatlas2.R:205-316): scaling 10 below 100 nodes
and 2 from there, Barnes-Hut from 10,000 nodes on the GPU, the weight
transform invert -> normalize -> power, node sizes, fixed nodes and the
random start U(-10 sqrt(n), 10 sqrt(n)). ``device`` is 1-based as in R.

``Batch`` lays out many graphs together on one device (fa2_batch_* of
include/fa2_hip.h).
"""
import ctypes
from ctypes import POINTER, byref, c_double, c_int, c_uint32, c_void_p

import numpy as np

from ._lib import (FA2_FULL_NTIMES, FA2_TUNE_NTIMES, DevInfo, Edges, FullParams, GPUError,
                   INTERRUPT_FN, Tuning, last_error, lib)

__all__ = ["gpu_devices", "layout_forceatlas2_gpu", "layout_forceatlas2_large",
           "prepare_weights", "prepare_large", "Batch", "GPUError"]


def _ptr(a):
    return None if a is None else a.ctypes.data_as(c_void_p)


def gpu_devices():
    """One dict per device: name, vendor, platform, version, type, memory_gb,
    double (R/gpu.R:143-148, src/forceatlas2.c:950-984)."""
    out = []
    for i in range(lib.fa2_gpu_ndevices()):
        info = DevInfo()
        if lib.fa2_gpu_device_info(i, byref(info)) != 0:
            continue
        out.append(dict(name=info.name.decode().strip(),
                        vendor=info.vendor.decode().strip(),
                        platform=info.platform.decode().strip(),
                        version=info.version.decode().strip(),
                        type="GPU" if info.is_gpu else "other",
                        memory_gb=info.memory / 1073741824.0,
                        double=bool(info.has_double)))
    return out


def _check_number(x, name, lo=-np.inf, strict=False, whole=False):
    ok = (np.isscalar(x) and np.isfinite(x) and (x > lo if strict else x >= lo)
          and (not whole or float(x) == round(float(x))))
    if not ok:
        raise ValueError(
            f"`{name}` must be a single {'whole ' if whole else ''}number "
            f"{'greater than' if strict else 'greater than or equal to'} {lo}.")
    return x


def _check_flag(x, name):
    if not isinstance(x, (bool, np.bool_)):
        raise ValueError(f"`{name}` must be TRUE or FALSE.")
    return bool(x)


def prepare_weights(m, weights=None, edge_weight_influence=1.0,
                    normalize_weights=False, invert_weights=False):
    """Edge weights as fa2_layout transforms them (R/forceatlas2.R:254-264)."""
    if weights is None:
        w = np.ones(m)
    else:
        w = np.asarray(weights, dtype=np.float64)
        if w.shape != (m,):
            raise ValueError(f"`weights` must have one value per edge ({m}), "
                             f"not {w.size}.")
        if not np.all(np.isfinite(w)) or np.any(w < 0):
            raise ValueError("`weights` must contain only finite, non-negative "
                             "values.")
    if invert_weights:
        with np.errstate(divide="ignore"):
            w = np.where(w == 0, 0.0, 1.0 / np.where(w == 0, 1.0, w))
    if normalize_weights and m > 0:
        lo, hi = w.min(), w.max()
        w = (w - lo) / (hi - lo) if lo < hi else np.ones(m)
    w = np.ones(m) if edge_weight_influence == 0 else w ** edge_weight_influence
    if not np.all(np.isfinite(w)):
        raise ValueError("The edge weights are too large or too small to be "
                         "used.")
    return w


def _edges(n, edges):
    e = np.asarray(edges, dtype=np.int64).reshape(-1, 2) if edges is not None \
        else np.zeros((0, 2), np.int64)
    if e.size and (e.min() < 0 or e.max() >= n):
        bad = int(np.nonzero((e < 0).any(1) | (e >= n).any(1))[0][0])
        raise ValueError(f"edge {bad + 1} refers to a node that does not exist")
    if e.shape[0] > (1 << 30) - 1:
        raise ValueError("too many edges")
    return (np.ascontiguousarray(e[:, 0], dtype=np.int32),
            np.ascontiguousarray(e[:, 1], dtype=np.int32))


def _common(n, edges, iterations, scaling_ratio, gravity, strong_gravity,
            linlog, dissuade_hubs, prevent_overlap, node_size, weights,
            edge_weight_influence, normalize_weights, invert_weights, init,
            fixed, jitter_tolerance, theta, rng):
    """The checks and preprocessing of fa2_layout (R/forceatlas2.R:205-276)."""
    iterations = int(_check_number(iterations, "iterations", 0, whole=True))
    _check_number(gravity, "gravity", 0)
    _check_number(edge_weight_influence, "edge_weight_influence", 0)
    _check_number(jitter_tolerance, "jitter_tolerance", 0, strict=True)
    _check_number(theta, "theta", 0)
    for name, v in (("strong_gravity", strong_gravity), ("linlog", linlog),
                    ("dissuade_hubs", dissuade_hubs),
                    ("prevent_overlap", prevent_overlap),
                    ("normalize_weights", normalize_weights),
                    ("invert_weights", invert_weights)):
        _check_flag(v, name)
    if scaling_ratio is None:
        scaling_ratio = 2.0 if n >= 100 else 10.0
    else:
        _check_number(scaling_ratio, "scaling_ratio", 0, strict=True)
    frm, to = _edges(n, edges)
    m = frm.size
    w = prepare_weights(m, weights, edge_weight_influence, normalize_weights,
                        invert_weights)
    size = np.broadcast_to(np.asarray(node_size, dtype=np.float64), (n,)).copy() \
        if np.ndim(node_size) == 0 else np.asarray(node_size, dtype=np.float64)
    if size.shape != (n,) or not np.all(np.isfinite(size)):
        raise ValueError("`node_size` must be a single number or one finite "
                         f"number per node ({n}).")
    if np.any(size < 0):
        raise ValueError("`node_size` must be non-negative.")
    if fixed is None:
        fx = np.zeros(n, dtype=np.int32)
    else:
        fx = np.asarray(fixed)
        if fx.dtype != bool or fx.shape != (n,):
            raise ValueError("`fixed` must be a logical vector with one value "
                             f"per node ({n}).")
        fx = fx.astype(np.int32)
    if init is None:
        if fx.any():
            raise ValueError("`fixed` nodes need a position: please provide "
                             "`init`.")
        rng = np.random.default_rng(rng)
        r = 10.0 * np.sqrt(n)
        pos = rng.uniform(-r, r, size=(n, 2))
    else:
        pos = np.asarray(init, dtype=np.float64)
        if pos.shape != (n, 2):
            raise ValueError("`init` must provide two numeric coordinates for "
                             f"each of the {n} nodes.")
        if not np.all(np.isfinite(pos)):
            raise ValueError("`init` must not contain missing or infinite "
                             "values.")
    dpar = np.array([scaling_ratio, gravity, jitter_tolerance, theta], np.float64)
    return iterations, frm, to, w, size, fx, pos, dpar


def _resolve(device, precision):
    """gpu_resolve (R/gpu.R:151-182): 0-based device and the FP64 flag."""
    if precision not in ("single", "double"):
        raise ValueError("'arg' should be one of \"single\", \"double\"")
    devices = gpu_devices()
    if not devices:
        raise GPUError("No GPU device was found on this computer: "
                       f"{last_error()}")
    if device is None:
        device = 1
    elif not (isinstance(device, (int, np.integer)) and 1 <= device <= len(devices)):
        raise ValueError("`device` must be the row number of a device in "
                         f"`gpu_devices()` (between 1 and {len(devices)}).")
    return int(device) - 1, int(precision == "double")


def layout_forceatlas2_gpu(n, edges, iterations=1000, scaling_ratio=None,
                           gravity=1.0, strong_gravity=False, linlog=False,
                           dissuade_hubs=False, prevent_overlap=False,
                           node_size=1.0, weights=None,
                           edge_weight_influence=1.0, normalize_weights=False,
                           invert_weights=False, init=None, fixed=None,
                           jitter_tolerance=1.0, barnes_hut=None, theta=1.2,
                           threads=1, device=None, precision="single",
                           rng=None, timings=False):
    """Repulsion on the GPU, the rest on the host (R/gpu.R:115-139 ->
    C_forceatlas2, here fa2_layout_run)."""
    iterations, frm, to, w, size, fx, pos, dpar = _common(
        n, edges, iterations, scaling_ratio, gravity, strong_gravity, linlog,
        dissuade_hubs, prevent_overlap, node_size, weights,
        edge_weight_influence, normalize_weights, invert_weights, init, fixed,
        jitter_tolerance, theta, rng)
    if barnes_hut is None:
        barnes_hut = n >= 10000
    _check_flag(barnes_hut, "barnes_hut")
    dev, dbl = _resolve(device, precision)
    xy = np.ascontiguousarray(pos.T).reshape(-1).copy()
    flags = np.array([strong_gravity, linlog, dissuade_hubs, prevent_overlap,
                      barnes_hut], np.int32)
    gpu = np.array([dev, dbl], np.int32)
    times = np.zeros(7)
    if n > 0 and iterations > 0:
        st = lib.fa2_layout_run(n, frm.size, _ptr(frm), _ptr(to), _ptr(w),
                                _ptr(xy), _ptr(size), _ptr(fx), iterations,
                                _ptr(dpar), _ptr(flags), int(threads),
                                _ptr(gpu), _ptr(times))
        if st != 0:
            raise GPUError(last_error())
    out = xy.reshape(2, n).T.copy()
    if timings:
        return out, dict(zip(["total", "tree", "repulsion", "gpu_send",
                              "gpu_compute", "gpu_fetch", "other"], times))
    return out


def layout_forceatlas2_large(n, edges, iterations=1000, scaling_ratio=None,
                             gravity=1.0, strong_gravity=False, linlog=False,
                             dissuade_hubs=False, prevent_overlap=False,
                             node_size=1.0, weights=None,
                             edge_weight_influence=1.0,
                             normalize_weights=False, invert_weights=False,
                             init=None, fixed=None, jitter_tolerance=1.0,
                             theta=1.2, device=None, precision="single",
                             rng=None, timings=False):
    """Every stage on the GPU (R/large.R:107-131 -> C_forceatlas2_large, here
    fa2_layout_large_run). Fewer than two nodes go to the repulsion offload,
    as the reference sends them to the CPU (R/forceatlas2.R:247-248)."""
    if n < 2:
        return layout_forceatlas2_gpu(
            n, edges, iterations, scaling_ratio, gravity, strong_gravity,
            linlog, dissuade_hubs, prevent_overlap, node_size, weights,
            edge_weight_influence, normalize_weights, invert_weights, init,
            fixed, jitter_tolerance, False, theta, 1, device, precision, rng,
            timings)
    iterations, frm, to, w, size, fx, pos, dpar = _common(
        n, edges, iterations, scaling_ratio, gravity, strong_gravity, linlog,
        dissuade_hubs, prevent_overlap, node_size, weights,
        edge_weight_influence, normalize_weights, invert_weights, init, fixed,
        jitter_tolerance, theta, rng)
    dev, dbl = _resolve(device, precision)
    xy = np.ascontiguousarray(pos.T).reshape(-1).copy()
    flags = np.array([strong_gravity, linlog, dissuade_hubs, prevent_overlap],
                     np.int32)
    gpu = np.array([dev, dbl, int(bool(timings))], np.int32)
    times = np.zeros(FA2_FULL_NTIMES)
    if iterations > 0:
        st = lib.fa2_layout_large_run(n, frm.size, _ptr(frm), _ptr(to),
                                      _ptr(w), _ptr(xy), _ptr(size), _ptr(fx),
                                      iterations, _ptr(dpar), _ptr(flags),
                                      _ptr(gpu), _ptr(times))
        if st != 0:
            raise GPUError(last_error())
    out = xy.reshape(2, n).T.copy()
    if timings:
        return out, dict(zip(["total", "codes", "sort", "tree", "repulsion",
                              "attraction", "move"], times))
    return out


def prepare_large(n, frm, to, w, outbound):
    """The preprocessing of C_forceatlas2_large (src/forceatlas2.c:870-902):
    mass = 1 + degree (self-loops count twice), incidence lists in edge
    order without self-loops, inc_coef = coef * w (/ mass[from] with
    outbound). Returns mass, inc_off, inc_node, inc_coef."""
    frm = np.asarray(frm, np.int64)
    to = np.asarray(to, np.int64)
    w = np.asarray(w, np.float64)
    mass = 1.0 + np.bincount(frm, minlength=n) + np.bincount(to, minlength=n)
    mass = mass.astype(np.float64)
    coef = mass.sum() / n if outbound else 1.0  # integer sums: exact
    keep = frm != to
    e = np.nonzero(keep)[0]
    a, b = frm[keep], to[keep]
    c = coef * w[keep] / mass[a] if outbound else coef * w[keep]
    node = np.concatenate([a, b])
    other = np.concatenate([b, a])
    cc = np.concatenate([c, c])
    ee = np.concatenate([e, e])
    order = np.lexsort((ee, node))
    inc_node = np.ascontiguousarray(other[order], dtype=np.int32)
    inc_coef = np.ascontiguousarray(cc[order], dtype=np.float64)
    inc_off = np.zeros(n + 1, np.int32)
    np.cumsum(np.bincount(node, minlength=n), out=inc_off[1:])
    return mass, inc_off, inc_node, inc_coef


def _fill_params(p, gr, device, precision, profile, edges=None):
    """Fill the fa2_full_params p from the graph dict gr (see Batch); returns
    the arrays p points into (to be kept alive) and the positions (x then
    y). With an Edges structure `edges`, the edge list goes there for the
    preprocessing on the device, and p's mass and incidence lists stay
    empty."""
    n = int(gr["n"])
    if "frm" in gr:
        frm = np.ascontiguousarray(gr["frm"], np.int32)
        to = np.ascontiguousarray(gr["to"], np.int32)
    else:
        frm, to = _edges(n, gr.get("edges"))
        frm = np.ascontiguousarray(frm, np.int32)
        to = np.ascontiguousarray(to, np.int32)
    w = None if gr.get("weights") is None else np.ascontiguousarray(gr["weights"], np.float64)
    outbound = bool(gr.get("dissuade_hubs", False))
    if edges is not None:
        edges.m = int(frm.size)
        edges.frm = frm.ctypes.data_as(POINTER(c_int))
        edges.to = to.ctypes.data_as(POINTER(c_int))
        edges.w = w.ctypes.data_as(POINTER(c_double)) if w is not None else None
        edges.outbound = int(outbound)
        mass = off = node = coef = None
    elif "inc_off" in gr:
        mass = np.ascontiguousarray(gr["mass"], np.float64)
        off = np.ascontiguousarray(gr["inc_off"], np.int32)
        node = np.ascontiguousarray(gr["inc_node"], np.int32)
        coef = np.ascontiguousarray(gr["inc_coef"], np.float64)
        if (mass.shape != (n,) or off.shape != (n + 1,) or off[0] != 0
                or np.any(np.diff(off) < 0) or node.shape != (int(off[-1]),)
                or coef.shape != node.shape):
            raise ValueError("mass, inc_off, inc_node and inc_coef must hold n, n + 1, "
                             "inc_off[n] and inc_off[n] values, inc_off rising from 0")
        if node.size and (node.min() < 0 or node.max() >= n):
            raise ValueError("inc_node holds a node outside 0..n-1")
    else:
        mass, off, node, coef = prepare_large(n, frm, to, np.ones(frm.size) if w is None else w,
                                              outbound)
    adjust = bool(gr.get("adjust", False))
    size = np.ascontiguousarray(
        np.broadcast_to(np.asarray(gr.get("size", 1.0), np.float64), (n,)))
    fixed = gr.get("fixed")
    fixed = np.zeros(n, np.int32) if fixed is None else np.asarray(fixed, np.int32)
    init = np.asarray(gr["init"], np.float64)
    xy = np.ascontiguousarray(init.T).reshape(-1).copy()
    p.n = n
    if off is not None:
        p.inc_off = off.ctypes.data_as(POINTER(c_int))
        p.inc_node = node.ctypes.data_as(POINTER(c_int))
        p.inc_coef = coef.ctypes.data_as(POINTER(c_double))
        p.mass = mass.ctypes.data_as(POINTER(c_double))
    p.size = size.ctypes.data_as(POINTER(c_double))
    p.fixed = fixed.ctypes.data_as(POINTER(c_int))
    p.any_fixed = int(fixed.any())
    p.scaling = float(gr.get("scaling", 2.0 if n >= 100 else 10.0))
    p.gravity = float(gr.get("gravity", 1.0))
    p.theta = float(gr.get("theta", 1.2))
    p.jitter_tolerance = float(gr.get("jitter_tolerance", 1.0))
    p.strong = int(bool(gr.get("strong", False)))
    p.linlog = int(bool(gr.get("linlog", False)))
    p.adjust = int(adjust)
    p.iterations = int(gr.get("iterations", 1 << 30))
    p.device = int(device)
    p.use_double = int(precision == "double")
    p.profile = int(bool(profile))
    p.source = None
    p.interrupted = INTERRUPT_FN(0)
    return [mass, off, node, coef, size, fixed, xy, frm, to, w], xy


def _tuning_dict(t):
    return dict(attract_unroll=tuple(int(t.attract_unroll[c]) for c in range(3)),
                repulse_threads=int(t.repulse_threads), repulse_xcd_run=int(t.repulse_xcd_run),
                attract_fork=int(t.attract_fork))


class Batch:
    """Graphs laid out together on one device (fa2_batch_* of fa2_hip.h).

    ``graphs`` is a list of dicts with keys n, edges (m x 2) or frm/to,
    weights (already transformed, default 1), init ((n, 2)), iterations, and
    the settings scaling, gravity, theta, jitter_tolerance, strong, linlog,
    dissuade_hubs, adjust, size, fixed. The device (0-based), the precision
    and ``profile`` are common to the batch. ``ingest`` chooses where the
    mass and incidence lists are built: "host" (numpy, or the prepared
    mass/inc_off/inc_node/inc_coef of the dicts) or "device"
    (fa2_batch_open_edges). ``autotune`` times the candidate launch shapes
    (fa2_batch_tune) before the first iteration and keeps the fastest; by
    default on for batches of at least AUTOTUNE_NODES nodes. The shapes never
    change a result.
    """

    AUTOTUNE_NODES = 65536

    def __init__(self, graphs, device=0, precision="single", profile=False, ingest="host",
                 autotune=None):
        if ingest not in ("host", "device"):
            raise ValueError("ingest must be 'host' or 'device'")
        self._keep = []
        G = len(graphs)
        params = (FullParams * G)()
        edges = (Edges * G)() if ingest == "device" else None
        pos_ptrs = (c_void_p * G)()
        self.n = []
        for g, gr in enumerate(graphs):
            keep, xy = _fill_params(params[g], gr, device, precision, profile,
                                    edges[g] if edges is not None else None)
            pos_ptrs[g] = xy.ctypes.data
            self._keep += keep
            self.n.append(int(gr["n"]))
        if edges is not None:
            self.h = lib.fa2_batch_open_edges(G, params, edges, pos_ptrs)
        else:
            self.h = lib.fa2_batch_open(G, params, pos_ptrs)
        self._keep = None
        if not self.h:
            raise GPUError(last_error())
        self.G = G
        self._autotune = (sum(self.n) >= self.AUTOTUNE_NODES) if autotune is None else autotune

    def run(self, iterations):
        if self._autotune:
            self._autotune = False
            self.tune(force=False)
        if lib.fa2_batch_run(self.h, int(iterations)) != 0:
            raise GPUError(last_error())

    def tune(self, force=True):
        """Time the candidate launch shapes on the current positions and keep
        the fastest (fa2_batch_tune). Returns the chosen shapes and the
        timings in ms (traversal at 64, 128, 256, 512, 1024 threads; at XCD runs of 0,
        8, 32, 128, 512 blocks; unrolls 1, 2, 4, 8 of each attraction class; then
        phases 1-6 of an iteration with the attraction forked at 0, 1, 2, 3, then
        phases 1-6 with the chosen traversal workgroup size and with 256 threads,
        0 when 256 was chosen),
        or None for timings when a cached choice was reused."""
        t = Tuning()
        times = np.full(FA2_TUNE_NTIMES, -1.0)
        if lib.fa2_batch_tune(self.h, int(bool(force)), ctypes.byref(t), _ptr(times)) != 0:
            raise GPUError(last_error())
        return _tuning_dict(t), (None if times[0] < 0 else times)

    @property
    def tuning(self):
        t = Tuning()
        if lib.fa2_batch_get_tuning(self.h, ctypes.byref(t)) != 0:
            raise GPUError(last_error())
        return _tuning_dict(t)

    def set_tuning(self, attract_unroll=(4, 4, 4), repulse_threads=256, repulse_xcd_run=32,
                   attract_fork=2):
        t = Tuning()
        for c in range(3):
            t.attract_unroll[c] = int(attract_unroll[c])
        t.repulse_threads = int(repulse_threads)
        t.repulse_xcd_run = int(repulse_xcd_run)
        t.attract_fork = int(attract_fork)
        if lib.fa2_batch_set_tuning(self.h, ctypes.byref(t)) != 0:
            raise GPUError(last_error())

    def sync(self):
        if lib.fa2_batch_sync(self.h) != 0:
            raise GPUError(last_error())

    def positions(self, g=0):
        n = self.n[g]
        out = np.empty(2 * n)
        if lib.fa2_batch_positions(self.h, g, _ptr(out)) != 0:
            raise GPUError(last_error())
        return out.reshape(2, n).T.copy()

    def forces(self, g=0):
        n = self.n[g]
        out = np.empty(2 * n)
        if lib.fa2_batch_forces(self.h, g, _ptr(out)) != 0:
            raise GPUError(last_error())
        return out.reshape(2, n).T.copy()

    def speed(self, g=0):
        out = np.empty(4)
        if lib.fa2_batch_speed(self.h, g, _ptr(out)) != 0:
            raise GPUError(last_error())
        return out

    def progress(self):
        """(iterations applied, status) per graph: status 0, or 2 once the
        graph's swing or traction totals stopped being numbers."""
        done = np.zeros(self.G, np.int32)
        status = np.zeros(self.G, np.int32)
        if lib.fa2_batch_progress(self.h, _ptr(done), _ptr(status)) != 0:
            raise GPUError(last_error())
        return done, status

    def tree(self, g=0):
        n = self.n[g]
        ks = np.empty(n, np.uint32)
        order = np.empty(n, np.uint32)
        child = np.empty(2 * (n - 1), np.int32)
        rng_ = np.empty(2 * (n - 1), np.int32)
        parent = np.empty(2 * n - 1, np.int32)
        if lib.fa2_batch_tree(self.h, g, _ptr(ks), _ptr(order), _ptr(child),
                              _ptr(rng_), _ptr(parent)) != 0:
            raise GPUError(last_error())
        return dict(keys=ks, order=order, child=child, range=rng_,
                    parent=parent)

    def csr(self, g=0):
        """The mass and incidence lists of graph g as the device holds them:
        mass (n), inc_off (n + 1), inc_node, inc_coef (in the batch's
        precision, as float64)."""
        n = self.n[g]
        off = np.empty(n + 1, np.int32)
        if lib.fa2_batch_csr(self.h, g, None, _ptr(off), None, None) != 0:
            raise GPUError(last_error())
        mass = np.empty(n)
        node = np.empty(int(off[-1]), np.int32)
        coef = np.empty(int(off[-1]))
        if lib.fa2_batch_csr(self.h, g, _ptr(mass), _ptr(off), _ptr(node), _ptr(coef)) != 0:
            raise GPUError(last_error())
        return mass, off, node, coef

    def regions(self, g=0):
        """Traversal records of graph g's regions in the reference numbering:
        an (n-1, 4) array of (centre x, centre y, mass, size^2) and the skip
        link of each region (region 0..n-2, leaf n-1.., -1 at the end)."""
        n = self.n[g]
        rec = np.empty(4 * (n - 1))
        skip = np.empty(n - 1, np.int32)
        if lib.fa2_batch_regions(self.h, g, _ptr(rec), _ptr(skip)) != 0:
            raise GPUError(last_error())
        return rec.reshape(n - 1, 4), skip

    def stage_times(self):
        out = np.zeros(FA2_FULL_NTIMES)
        lib.fa2_batch_stage_times(self.h, _ptr(out))
        return dict(zip(["total", "codes", "sort", "tree", "repulsion",
                         "attraction", "move"], out))

    def set_state(self, g, pos, f_old=None, speed=1.0, efficiency=1.0):
        """Replace the state of graph g: positions (n, 2), previous force
        (n, 2) or None, speed and speed efficiency."""
        n = self.n[g]
        xy = np.ascontiguousarray(np.asarray(pos, np.float64).T).reshape(-1)
        fo = None if f_old is None else \
            np.ascontiguousarray(np.asarray(f_old, np.float64).T).reshape(-1)
        if lib.fa2_batch_set_state(self.h, g, _ptr(xy), _ptr(fo), float(speed),
                                   float(efficiency)) != 0:
            raise GPUError(last_error())

    def count_work(self):
        """(accepted regions, opened regions, node pairs, wave steps) of the
        repulsion on the last tree."""
        out = np.zeros(4, np.uint64)
        if lib.fa2_batch_count_work(self.h, _ptr(out)) != 0:
            raise GPUError(last_error())
        return tuple(int(v) for v in out)

    @property
    def stream(self):
        return lib.fa2_batch_stream(self.h)

    def close(self):
        if getattr(self, "h", None):
            lib.fa2_batch_close(self.h)
            self.h = None

    def __del__(self):
        self.close()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


class Server:
    """A layout server on one device (fa2_serve_* of fa2_hip.h): requests
    are admitted by device-memory footprint into batches that run at the
    same time on their own streams; each request's positions come back as
    soon as its own iterations are done.

    ``submit(graph, precision)`` takes a graph dict as for Batch (with
    ``iterations``) and returns a request id; ``result(id)`` waits for it
    and returns the positions (n, 2) and the times (queued, running, total)
    in seconds. Requests of both precisions can be mixed: each batch holds
    one precision.

    The streams share the hardware queues HIP gives the process
    (``GPU_MAX_HW_QUEUES``, 4 by default, read when the HIP runtime starts):
    streams that share a queue wait for each other's work, so batch rebuilds
    wait longer with 4 queues than with ``max_batches`` of them. More queues
    can also put a later batch's two streams (the attraction runs on a
    second one beside the traversal) on one queue, which serializes them;
    measured on the 200-request stream, 4 queues served as many layouts per
    second as 16.
    """

    def __init__(self, device=0, budget_bytes=0.0, max_batches=8, chunk=25):
        self.h = lib.fa2_serve_open(int(device), float(budget_bytes), int(max_batches),
                                    int(chunk))
        if not self.h:
            raise GPUError(last_error())
        self.device = int(device)
        self.n = {}

    def submit(self, graph, precision="single", ingest="host"):
        """Queue one layout; ``ingest`` as for Batch ("device": the edge
        list is sent and prepared on the device)."""
        if precision not in ("single", "double"):
            raise ValueError("precision must be 'single' or 'double'")
        if ingest not in ("host", "device"):
            raise ValueError("ingest must be 'host' or 'device'")
        p = FullParams()
        e = Edges() if ingest == "device" else None
        keep, xy = _fill_params(p, graph, self.device, precision, False, e)
        if e is not None:
            job = lib.fa2_serve_submit_edges(self.h, ctypes.byref(p), ctypes.byref(e),
                                             xy.ctypes.data)
        else:
            job = lib.fa2_serve_submit(self.h, ctypes.byref(p), xy.ctypes.data)
        del keep
        if job < 0:
            raise GPUError(last_error())
        self.n[job] = int(graph["n"])
        return job

    def status(self, job):
        return ("queued", "running", "done", "failed")[lib.fa2_serve_status(self.h, int(job))]

    def result(self, job):
        n = self.n[job]
        xy = np.empty(2 * n)
        t = np.zeros(3)
        if lib.fa2_serve_wait(self.h, int(job), _ptr(xy), _ptr(t)) != 0:
            raise GPUError(last_error())
        return xy.reshape(2, n).T.copy(), dict(queued=t[0], running=t[1], total=t[2])

    def stats(self):
        out = np.zeros(7)
        lib.fa2_serve_stats(self.h, _ptr(out))
        return dict(batches=int(out[0]), admitted=int(out[1]), peak_running=int(out[2]),
                    budget_bytes=out[3], refills=int(out[4]), refill_s=float(out[5]),
                    uncaptured=int(out[6]))

    def close(self):
        if self.h:
            lib.fa2_serve_close(self.h)
            self.h = None

    def __enter__(self):
        return self

    def __exit__(self, *exc):
        self.close()

    def __del__(self):
        self.close()


def serve_footprint(n, incidences, precision="single"):
    """Device bytes a request of n nodes and `incidences` incidence entries
    (twice the edges without self-loops) takes on the server."""
    return lib.fa2_serve_footprint(int(precision == "double"), int(n), int(incidences))

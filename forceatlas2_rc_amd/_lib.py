"""ctypes binding of libfa2hip.so (the C ABI of include/fa2.h and
include/fa2_hip.h).

The library is built in-tree (``python -c "import __graft_entry__ as g;
g.build()"`` or ``make -C forceatlas2_rc_amd/csrc``). Importing this module
without it raises: there is no other implementation of the GPU path.
"""
import ctypes
import os
from ctypes import (CFUNCTYPE, POINTER, Structure, c_char, c_char_p, c_double,
                    c_int, c_uint32, c_void_p)

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "libfa2hip.so")

FA2_FULL_NTIMES = 7
INTERRUPT_FN = CFUNCTYPE(c_int)


class Region(Structure):
    """fa2_region (src/fa2.h:21-27)."""
    _fields_ = [("mass", c_double), ("cx", c_double), ("cy", c_double),
                ("size", c_double), ("start", c_int), ("count", c_int),
                ("child", c_int), ("nchild", c_int), ("skip", c_int),
                ("buf", c_int)]


class Node(Structure):
    """fa2_node (src/fa2.h:30-33)."""
    _fields_ = [("x", c_double), ("y", c_double), ("m", c_double),
                ("idx", c_int)]


class Tree(Structure):
    """fa2_tree (src/fa2.h:35-44)."""
    _fields_ = [("reg", POINTER(Region)), ("nreg", c_int),
                ("perm", POINTER(c_int)), ("cnt", POINTER(c_int)),
                ("buf", POINTER(Node) * 2), ("px", POINTER(c_double)),
                ("py", POINTER(c_double)), ("pm", POINTER(c_double)),
                ("ps", POINTER(c_double))]


class DevInfo(Structure):
    """fa2_gpu_devinfo (src/fa2.h:50-58)."""
    _fields_ = [("name", c_char * 128), ("vendor", c_char * 128),
                ("platform", c_char * 128), ("version", c_char * 128),
                ("is_gpu", c_int), ("has_double", c_int),
                ("memory", c_double)]


class Edges(Structure):
    """fa2_edges (include/fa2_hip.h): an edge list for the preprocessing on
    the device."""
    _fields_ = [("m", c_int), ("frm", POINTER(c_int)), ("to", POINTER(c_int)),
                ("w", POINTER(c_double)), ("outbound", c_int)]


class Tuning(Structure):
    """fa2_tuning (include/fa2_hip.h): launch shapes of a batch."""
    _fields_ = [("attract_unroll", c_int * 3), ("repulse_threads", c_int),
                ("repulse_xcd_run", c_int), ("attract_fork", c_int)]


FA2_TUNE_NTIMES = 28


class FullParams(Structure):
    """fa2_full_params (src/fa2.h:97-115)."""
    _fields_ = [("n", c_int), ("inc_off", POINTER(c_int)),
                ("inc_node", POINTER(c_int)), ("inc_coef", POINTER(c_double)),
                ("mass", POINTER(c_double)), ("size", POINTER(c_double)),
                ("fixed", POINTER(c_int)), ("any_fixed", c_int),
                ("scaling", c_double), ("gravity", c_double),
                ("theta", c_double), ("jitter_tolerance", c_double),
                ("strong", c_int), ("linlog", c_int), ("adjust", c_int),
                ("iterations", c_int), ("device", c_int),
                ("use_double", c_int), ("profile", c_int),
                ("source", c_char_p), ("interrupted", INTERRUPT_FN)]


# every symbol the headers declare, with its prototype
PROTOTYPES = {
    # include/fa2.h
    "fa2_gpu_error": (c_char_p, []),
    "fa2_gpu_ndevices": (c_int, []),
    "fa2_gpu_device_info": (c_int, [c_int, POINTER(DevInfo)]),
    "fa2_gpu_open": (c_void_p, [c_int, c_int, c_int, c_int, c_int, c_int]),
    "fa2_gpu_times": (None, [c_void_p, c_void_p]),
    "fa2_gpu_close": (None, [c_void_p]),
    "fa2_gpu_repulse_exact": (c_int, [c_void_p, c_void_p, c_void_p, c_void_p,
                                      c_void_p, c_double, c_void_p, c_void_p]),
    "fa2_gpu_repulse_bh": (c_int, [c_void_p, POINTER(Tree), c_double, c_double,
                                   c_void_p, c_void_p]),
    "fa2_gpu_shutdown": (None, []),
    "fa2_now": (c_double, []),
    "fa2_full_run": (c_int, [POINTER(FullParams), c_void_p, c_void_p]),
    "fa2_update_speed": (None, [c_int, c_double, c_double, c_double,
                                POINTER(c_double), POINTER(c_double)]),
    # include/fa2_hip.h
    "fa2_layout_run": (c_int, [c_int, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p, c_void_p, c_int, c_void_p,
                               c_void_p, c_int, c_void_p, c_void_p]),
    "fa2_layout_large_run": (c_int, [c_int, c_int, c_void_p, c_void_p,
                                     c_void_p, c_void_p, c_void_p, c_void_p,
                                     c_int, c_void_p, c_void_p, c_void_p,
                                     c_void_p]),
    "fa2_tree_build": (None, [POINTER(Tree), c_int, c_void_p, c_void_p,
                              c_void_p, c_void_p, c_int]),
    "fa2_gpu_kernel_time": (c_int, [c_void_p, POINTER(c_double)]),
    "fa2_gpu_repulse_exact_noalloc": (c_int, [c_int, c_int, c_int, c_int, c_void_p, c_void_p,
                                              c_void_p, c_void_p, c_double, c_void_p,
                                              c_void_p, c_void_p]),
    "fa2_batch_open": (c_void_p, [c_int, POINTER(FullParams),
                                  POINTER(c_void_p)]),
    "fa2_batch_open_edges": (c_void_p, [c_int, POINTER(FullParams), POINTER(Edges),
                                        POINTER(c_void_p)]),
    "fa2_batch_csr": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p, c_void_p]),
    "fa2_batch_run": (c_int, [c_void_p, c_int]),
    "fa2_batch_sync": (c_int, [c_void_p]),
    "fa2_batch_positions": (c_int, [c_void_p, c_int, c_void_p]),
    "fa2_batch_forces": (c_int, [c_void_p, c_int, c_void_p]),
    "fa2_batch_speed": (c_int, [c_void_p, c_int, c_void_p]),
    "fa2_batch_progress": (c_int, [c_void_p, c_void_p, c_void_p]),
    "fa2_batch_regions": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "fa2_batch_tree": (c_int, [c_void_p, c_int, c_void_p, c_void_p, c_void_p,
                               c_void_p, c_void_p]),
    "fa2_batch_stage_times": (c_int, [c_void_p, c_void_p]),
    "fa2_batch_count_work": (c_int, [c_void_p, c_void_p]),
    "fa2_batch_set_state": (c_int, [c_void_p, c_int, c_void_p, c_void_p,
                                    c_double, c_double]),
    "fa2_batch_tune": (c_int, [c_void_p, c_int, POINTER(Tuning), c_void_p]),
    "fa2_batch_get_tuning": (c_int, [c_void_p, POINTER(Tuning)]),
    "fa2_batch_set_tuning": (c_int, [c_void_p, POINTER(Tuning)]),
    "fa2_batch_stream": (c_void_p, [c_void_p]),
    "fa2_batch_close": (None, [c_void_p]),
    "fa2_serve_open": (c_void_p, [c_int, c_double, c_int, c_int]),
    "fa2_serve_submit": (c_int, [c_void_p, POINTER(FullParams), c_void_p]),
    "fa2_serve_submit_edges": (c_int, [c_void_p, POINTER(FullParams), POINTER(Edges),
                                       c_void_p]),
    "fa2_serve_status": (c_int, [c_void_p, c_int]),
    "fa2_serve_wait": (c_int, [c_void_p, c_int, c_void_p, c_void_p]),
    "fa2_serve_stats": (c_int, [c_void_p, c_void_p]),
    "fa2_serve_close": (None, [c_void_p]),
    "fa2_serve_footprint": (c_double, [c_int, c_int, ctypes.c_longlong]),
}


def _load():
    if not os.path.exists(LIB_PATH):
        raise ImportError(
            f"{LIB_PATH} is missing: build it first (make -C "
            "forceatlas2_rc_amd/csrc, or __graft_entry__.build()); the GPU "
            "path has no other implementation")
    lib = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in PROTOTYPES.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    return lib


lib = _load()


def last_error():
    msg = lib.fa2_gpu_error()
    return msg.decode() if msg else ""


class GPUError(RuntimeError):
    """An error reported by libfa2hip.so."""

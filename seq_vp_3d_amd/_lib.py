"""ctypes binding of libseqvp_hip.so (the C-ABI declared in include/seqvp.h).

The library is built in-tree by `seq_vp_3d_amd.build.build()`. There is no fallback: if the library
is missing or a call fails, an exception is raised.
"""
import ctypes
from pathlib import Path

import torch

LIB_PATH = Path(__file__).resolve().with_name("libseqvp_hip.so")

c_int, c_i64, c_dbl, c_vp, c_sz = ctypes.c_int, ctypes.c_int64, ctypes.c_double, ctypes.c_void_p, ctypes.c_size_t

# name -> argument types (all return int status unless listed in _RESTYPES)
SIGNATURES = {
    "seqvp_version": [],
    "seqvp_workspace_bytes": [c_int, c_i64, c_i64, c_i64],
    "seqvp_q1_soft_energy": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_soft_energy_noalloc": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_q1_trial_energy": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_trial_energy_base_min": [c_vp, c_vp, c_vp, c_vp, c_int, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz,
                                       c_vp],
    "seqvp_q1_trial_energy_batch": [c_vp, c_int, c_int, c_i64, c_vp],
    "seqvp_q1_barrier_energy": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_soft_energy_f32": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_barrier_energy_f32": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_node_energy": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64,
                          c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_node_terms": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_vp,
                         c_vp, c_vp, c_vp],
    "seqvp_q1_gradient": [c_vp, c_vp, c_vp, c_int, c_int, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_gradient_free": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_i64, c_i64,
                               c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_joint_free": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_i64, c_i64, c_i64,
                            c_vp, c_vp, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_joint": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_sz, c_vp],
    "seqvp_q1_ggn_diagonal": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_node_min_jd": [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_node_min_jd_extended_workspace_bytes": [c_i64, c_i64, c_i64],
    "seqvp_node_min_jd_extended": [c_vp, c_vp, c_i64, c_i64, c_i64, c_dbl, c_vp, c_vp, c_vp, c_sz, c_vp],
    "seqvp_face_interior_minima": [c_vp, c_i64, c_vp, c_vp],
    "seqvp_closed27_tile_minima": [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_witness_coefficients": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_i64, c_vp, c_vp],
    "seqvp_point_jd_interval_host": [c_vp, c_vp, c_vp, c_vp],
    "seqvp_face_interior_minima_host": [c_vp, c_i64, c_vp],
    "seqvp_jd_root": [c_vp, c_vp, c_vp, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_jd_root_free": [c_vp, c_vp, c_vp, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_jd_root_free_dev": [c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_pg_step": [c_vp, c_int, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp],
    "seqvp_capped_step": [c_vp, c_vp, c_dbl, c_vp, c_vp],
    "seqvp_prescreen": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_prescreen_flags": [c_vp, c_vp, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp],
    "seqvp_spline_warp": [c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_dbl, c_dbl,
                          c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp],
    "seqvp_spline_points": [c_vp, c_i64, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_i64, c_vp, c_i64, c_vp, c_i64, c_dbl,
                            c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_vp, c_vp, c_vp, c_vp],
    "seqvp_dst_poisson": [c_vp, c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_dst_poisson_f32": [c_vp, c_vp, c_vp, c_int, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_poisson_denominator": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_int, c_dbl, c_dbl,
                                  c_dbl, c_vp, c_vp],
    "seqvp_dots": [c_vp, c_vp, c_int, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_inject": [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_restrict": [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_axpy_inject": [c_vp, c_vp, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp],
    "seqvp_axpy_inject_dev": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp],
    "seqvp_direction_scalars": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                c_i64, c_i64, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_graph_begin": [c_vp],
    "seqvp_graph_end": [c_vp, ctypes.POINTER(c_vp)],
    "seqvp_graph_launch": [c_vp, c_vp],
    "seqvp_graph_destroy": [c_vp],
    "seqvp_copy_async": [c_vp, c_vp, c_sz, c_vp],
    "seqvp_bb2_products": [c_vp, c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_step_bound": [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_sz,
                         c_vp],
    "seqvp_max_abs_diff": [c_vp, c_vp, c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_metric_scale": [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp],
    "seqvp_mfma_probe": [c_vp, c_vp, c_int, c_vp, c_vp],
    "seqvp_certificate_workspace_bytes": [c_i64, c_i64, c_i64, c_i64],
    "seqvp_edge_hull": [c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp],
    "seqvp_certificate_screen": [c_vp, c_vp, c_dbl, c_dbl, c_dbl, c_dbl, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp,
                                 c_i64, c_vp, c_vp, c_sz, c_vp],
    "seqvp_gather_vertices": [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_vp, c_vp],
    "seqvp_bernstein_arrays": [c_vp, c_i64, c_vp, c_dbl, c_vp, c_vp, c_vp, c_vp, c_vp],
    "seqvp_bernstein_cell": [c_vp, c_vp, c_dbl, c_int, c_i64, c_vp],
    "seqvp_certificate_finish": [c_vp, c_i64, c_vp, c_vp, c_i64, c_vp, c_dbl, c_i64, c_int, c_i64, c_int, c_vp],
    "seqvp_edge_hull_host": [c_vp, c_i64, c_i64, c_i64, c_vp, c_dbl, c_int, c_vp, c_vp],
    "seqvp_certify_map_host": [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_dbl, c_vp, c_i64, c_int, c_i64,
                               c_int, c_vp],
    "seqvp_preopt_workspace_bytes": [c_i64, c_i64, c_i64, c_int],
    "seqvp_preopt_energy_grad": [c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl, c_dbl,
                                 c_dbl, c_int, c_vp, c_vp, c_vp, c_sz, c_vp],
    "seqvp_preopt_step": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_int, c_vp, c_int, c_dbl, c_dbl,
                          c_dbl, c_dbl, c_dbl, c_int, c_vp, c_vp, c_vp],
    "seqvp_gauss_axis": [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_int, c_vp, c_int, c_vp],
    "seqvp_lerp_axis": [c_vp, c_vp, c_i64, c_i64, c_i64, c_i64, c_int, c_i64, c_vp, c_vp, c_vp, c_vp],
    "seqvp_taper": [c_vp, c_i64, c_i64, c_i64, c_vp, c_vp, c_vp, c_vp, c_vp],
    "seqvp_scale": [c_vp, c_dbl, c_vp, c_i64, c_vp],
    "seqvp_fp64_ceiling_probe": [c_int, c_i64, c_i64, c_vp, c_vp],
    "seqvp_field_to_interior": [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_int, c_vp, c_vp],
    "seqvp_interior_to_field": [c_vp, c_i64, c_i64, c_i64, c_dbl, c_dbl, c_dbl, c_int, c_vp, c_vp],
    "seqvp_axis_transform": [c_vp, c_vp, c_vp, c_vp, c_vp, c_i64, c_i64, c_i64, c_vp, c_sz, c_vp],
}
_RESTYPES = {"seqvp_version": ctypes.c_char_p, "seqvp_workspace_bytes": c_sz,
             "seqvp_certificate_workspace_bytes": c_sz, "seqvp_preopt_workspace_bytes": c_sz,
             "seqvp_node_min_jd_extended_workspace_bytes": c_sz}


class Certificate(ctypes.Structure):
    """SeqvpCertificate (include/seqvp.h)."""
    _fields_ = [("status", ctypes.c_int32), ("reason", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("arithmetic_failure", ctypes.c_int32), ("n_cells_checked", c_i64), ("n_uncertain", c_i64),
                ("n_infeasible", c_i64), ("n_boxes", c_i64), ("n_fast_cells", c_i64), ("n_prescreen_cells", c_i64),
                ("n_fallback_cells", c_i64), ("witness_cell", c_i64), ("witness_point", c_dbl * 3),
                ("min_margin", c_dbl), ("min_jd", c_dbl), ("sampled_min_jd", c_dbl)]


class CellCertificate(ctypes.Structure):
    """SeqvpCellCertificate (include/seqvp.h)."""
    _fields_ = [("status", ctypes.c_int32), ("reason", ctypes.c_int32), ("max_depth", ctypes.c_int32),
                ("has_witness", ctypes.c_int32), ("n_boxes", c_i64), ("witness_point", c_dbl * 3),
                ("min_margin", c_dbl), ("min_jd", c_dbl), ("sampled_min_jd", c_dbl)]

ERRORS = {-1: "invalid argument", -2: "kernel launch failed", -3: "workspace too small"}

_LIB = None


class SeqvpError(RuntimeError):
    """A C-ABI call returned a nonzero status."""


def load():
    """The loaded library with every declared signature bound (raises if it was not built)."""
    global _LIB
    if _LIB is None:
        if not LIB_PATH.exists():
            raise RuntimeError(f"{LIB_PATH.name} is not built; run seq_vp_3d_amd.build.build() first")
        lib = ctypes.CDLL(str(LIB_PATH))
        for name, args in SIGNATURES.items():
            fn = getattr(lib, name)
            fn.argtypes = args
            fn.restype = _RESTYPES.get(name, c_int)
        _LIB = lib
    return _LIB


class CallTimer:
    """Records a pair of HIP events (on the current stream) around every C-ABI call while active, or only around
    the calls of the entry points in `names` (each bracket costs the host a few microseconds)."""

    def __init__(self, names=None):
        self.events = {}
        self.names = None if names is None else frozenset(names)

    def summary(self):
        """name -> (calls, mean ms) over the recorded calls (synchronizes on the last event)."""
        out = {}
        for name, pairs in self.events.items():
            if pairs:
                pairs[-1][1].synchronize()
                ms = [a.elapsed_time(b) for a, b in pairs]
                out[name] = (len(ms), sum(ms) / len(ms))
        return out


TIMER = None


def call(name, *args):
    timer = TIMER
    if timer is not None and timer.names is not None and name not in timer.names:
        timer = None
    if timer is not None and torch.cuda.is_current_stream_capturing():
        timer = None                # a captured call runs later, inside its graph: nothing to bracket now
    if timer is not None:
        start, end = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
        start.record()
    code = getattr(load(), name)(*args)
    if timer is not None:
        end.record()
        timer.events.setdefault(name, []).append((start, end))
    if code != 0:
        raise SeqvpError(f"{name}: {ERRORS.get(code, code)}")


def raw(name, *args):
    """The C-ABI call's integer result, unchecked (for entry points whose result is a value)."""
    return getattr(load(), name)(*args)


def hptr(a):
    """Host address of a C-contiguous numpy array (None -> NULL)."""
    if a is None:
        return None
    if not a.flags.c_contiguous:
        raise ValueError("C-contiguous host array expected")
    # the array interface's address (a.ctypes builds a helper object per call, ~10 us)
    return ctypes.c_void_p(a.__array_interface__["data"][0])


def ptr(t):
    """Device address of a tensor (None -> NULL)."""
    if t is None:
        return None
    if not isinstance(t, torch.Tensor):
        raise TypeError("tensor expected")
    if not t.is_cuda:
        raise ValueError("device tensor expected: the MI355X path has no CPU fallback")
    if not t.is_contiguous():
        raise ValueError("contiguous tensor expected")
    return ctypes.c_void_p(t.data_ptr())


def stream_handle(device=None):
    """The current HIP stream of `device` (torch's current stream: a caller's `torch.cuda.stream` context applies),
    read as the raw handle without building a Stream object (a few microseconds per call on the hot path)."""
    if isinstance(device, torch.device) and device.index is not None:
        index = device.index
    elif isinstance(device, int):
        index = device
    else:
        index = torch.cuda.current_device() if device is None else torch.device(device).index
        if index is None:
            index = torch.cuda.current_device()
    return ctypes.c_void_p(torch._C._cuda_getCurrentRawStream(index))


_WS_BYTES = {}


def workspace_bytes(op, shape_zyx):
    key = (op, tuple(shape_zyx))
    n = _WS_BYTES.get(key)
    if n is None:
        nz, ny, nx = shape_zyx
        n = _WS_BYTES[key] = int(load().seqvp_workspace_bytes(op, nx, ny, nz))
    return n


class Workspace:
    """One growable device byte buffer shared by the calls of a model on one stream."""

    def __init__(self, device):
        self.device = device
        self.buf = None

    def get(self, nbytes):
        nbytes = max(int(nbytes), 256)
        if self.buf is None or self.buf.numel() < nbytes:
            self.buf = torch.empty(nbytes, dtype=torch.uint8, device=self.device)
        return self.buf, nbytes

"""Whole-cell JD certificate of a map (the certify_map interface with the prescreen configuration).

The reference proves J > delta on every closed cell with three stages (jd_curl/three_d/feasibility.py:
571-769, PrescreenCertificateConfig): the outward-rounded edge hull, the T75 floating-point Bernstein
prescreen, and the outward-rounded interval Bernstein fallback with subdivision, which stays authoritative
for every cell the first two leave undecided.

Here the first two run on the GPU over every cell (seqvp_certificate_screen, bitwise the CPU arithmetic);
only the undecided cells (usually none) are gathered to the host and classified by the C++ fallback on
worker threads (seqvp_certificate_finish). certify_map() runs all three stages on the CPU
(seqvp_certify_map_host): the independent re-certification of an output.
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .problem import InputValidationError, q1_coefficients

STATUS = ("certified_feasible", "uncertain", "infeasible")
REASONS = ("positive_enclosure_cover", "positive_bernstein_cover", "JACOBIAN_VIOLATION", "arithmetic_uncertain",
           "subdivision_budget", "NONFINITE_STATE", "NONZERO_BOUNDARY", "INVALID_DOMAIN")
CELL_REASONS = ("positive_bernstein_cover", "sample_witness", "subdivision_witness", "box_budget", "depth_budget",
                "arithmetic_uncertain")
CERTIFICATE_VERSION = "q1-3d-parallel-hoisted-edge-hull-float-bernstein-prescreen-compiled-bernstein-v1"


@dataclass(frozen=True)
class CertificateConfig:
    """The reference's PrescreenCertificateConfig defaults (feasibility.py:21-117)."""
    delta: float = .01
    max_depth: int = 48
    max_boxes: int = 255
    tile_size: int = 8192
    fallback_workers: int = 8
    prescreen: bool = True

    def __post_init__(self):
        d = float(self.delta)
        if not (np.isfinite(d) and 0 <= d < 1):
            raise InputValidationError("delta: require 0 <= delta < 1")
        object.__setattr__(self, "delta", 0.0 if d == 0 else d)
        if not (0 <= int(self.max_depth) <= 120) or int(self.max_boxes) < 1 or int(self.tile_size) < 1:
            raise InputValidationError("certificate budgets out of range")
        if int(self.fallback_workers) < 1:
            raise InputValidationError("fallback_workers >= 1 required")


@dataclass(frozen=True)
class FeasibilityResult:
    """PrescreenFeasibilityResult fields (feasibility.py:136-173); None where the reference has None."""
    status: str
    min_jd_lower_bound: float | None
    min_margin_lower_bound: float | None
    sampled_min_jd: float | None
    n_cells_checked: int
    n_uncertain: int
    n_infeasible: int
    n_boxes: int
    max_depth: int
    witness_cell: int | None
    witness_reference_point: tuple | None
    reason: str
    n_fast_cells: int = 0
    n_fallback_cells: int = 0
    n_prescreen_cells: int = 0
    certificate_version: str = CERTIFICATE_VERSION

    @property
    def certified(self):
        return self.status == "certified_feasible"


@dataclass(frozen=True)
class CellCertificate:
    status: str
    min_jd_lower_bound: float | None
    min_margin_lower_bound: float | None
    sampled_min_jd: float | None
    n_boxes: int
    max_depth: int
    witness_reference_point: tuple | None
    reason: str


def _opt(x):
    return None if np.isnan(x) else float(x)


def _result(c):
    witness = int(c.witness_cell) if c.witness_cell >= 0 else None
    return FeasibilityResult(STATUS[c.status], _opt(c.min_jd), _opt(c.min_margin), _opt(c.sampled_min_jd),
                             int(c.n_cells_checked), int(c.n_uncertain), int(c.n_infeasible), int(c.n_boxes),
                             int(c.max_depth), witness, tuple(float(v) for v in c.witness_point) if witness is not None
                             else None, REASONS[c.reason], int(c.n_fast_cells), int(c.n_fallback_cells),
                             int(c.n_prescreen_cells))


def _host_call(name, *args):
    code = _lib.raw(name, *args)
    if code != 0:
        raise _lib.SeqvpError(f"{name}: {_lib.ERRORS.get(code, code)}")


def _spacing(grid):
    return np.ascontiguousarray(grid.spacing, dtype=np.float64)


def certify_map(u, grid, config=None, *, workers=None):
    """CPU certificate of a host map u (nz, ny, nx, 3) (certify_map, feasibility.py:395-431 and
    _certify_map_parallel): boundary and domain checks, edge hull, prescreen, Bernstein fallback."""
    config = CertificateConfig() if config is None else config
    u = np.ascontiguousarray(u.detach().cpu().numpy() if isinstance(u, torch.Tensor) else u, dtype=np.float64)
    if u.shape != tuple(grid.shape) + (3,):
        raise InputValidationError("u: expected (nz, ny, nx, 3)")
    nz, ny, nx = grid.shape
    axes = [np.ascontiguousarray(a, dtype=np.float64) for a in grid.axes]
    D = None
    if config.prescreen:
        try:
            D = np.ascontiguousarray(q1_coefficients(grid.spacing, "closed27_uniform")[0], dtype=np.float64)
        except FloatingPointError:
            D = None       # extreme spacing: no prescreen (feasibility.py:667-670)
    out = _lib.Certificate()
    _host_call("seqvp_certify_map_host", _lib.hptr(u), nx, ny, nz, _lib.hptr(axes[0]), _lib.hptr(axes[1]),
               _lib.hptr(axes[2]), _lib.hptr(_spacing(grid)), config.delta, _lib.hptr(D), int(config.tile_size),
               int(config.max_depth), int(config.max_boxes), int(workers or config.fallback_workers),
               ctypes.byref(out))
    return _result(out)


def certify_cell(vertices, spacing, config=None):
    """Bernstein classification of one cell's (8, 3) host vertex displacements (certify_cell without the
    edge-hull pre-proof, feasibility.py:375-393)."""
    config = CertificateConfig() if config is None else config
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    if v.shape != (8, 3):
        raise InputValidationError("vertices: expected (8, 3)")
    out = _lib.CellCertificate()
    _host_call("seqvp_bernstein_cell", _lib.hptr(v), _lib.hptr(np.ascontiguousarray(spacing, dtype=np.float64)),
               config.delta, int(config.max_depth), int(config.max_boxes), ctypes.byref(out))
    witness = tuple(float(x) for x in out.witness_point) if out.has_witness else None
    return CellCertificate(STATUS[out.status], _opt(out.min_jd), _opt(out.min_margin), _opt(out.sampled_min_jd),
                           int(out.n_boxes), int(out.max_depth), witness, CELL_REASONS[out.reason])


def fallback_arrays(vertices, spacing, delta):
    """(p_lo, p_hi, c_lo, c_hi, samples, ok) of the fallback for (n, 8, 3) host vertices, each (n, 27)
    (compiled_fallback.fallback_arrays, bitwise)."""
    v = np.ascontiguousarray(vertices, dtype=np.float64)
    n = len(v)
    arrs = [np.empty((n, 27)) for _ in range(5)]
    code = _lib.raw("seqvp_bernstein_arrays", _lib.hptr(v), n, _lib.hptr(np.ascontiguousarray(spacing, dtype=np.float64)),
                    float(delta), *(_lib.hptr(a) for a in arrs))
    if code < 0:
        raise _lib.SeqvpError(f"seqvp_bernstein_arrays: {_lib.ERRORS.get(code, code)}")
    return (*arrs, bool(code))


def edge_hull(u, grid, delta, margin=None, flag=None):
    """Device edge-hull margins and invalid-step flags of every cell (parallel_margins, bitwise)."""
    nz, ny, nx = grid.shape
    hx, hy, hz = (float(h) for h in grid.spacing)
    ncell = grid.n_cells
    margin = torch.empty(ncell, dtype=torch.float64, device=u.device) if margin is None else margin
    flag = torch.empty(ncell, dtype=torch.uint8, device=u.device) if flag is None else flag
    uc = u.contiguous()
    _lib.call("seqvp_edge_hull", _lib.ptr(uc), hx, hy, hz, float(delta), nx, ny, nz, _lib.ptr(margin), _lib.ptr(flag),
              _lib.stream_handle(u.device))
    return margin, flag


class DeviceCertificate:
    """certify_map of device maps on one grid: GPU screen of every cell, CPU fallback of the undecided ones.

    screen(u) enqueues the screen on the current stream; finish() (after the stream's results are needed)
    returns the FeasibilityResult. `capacity` bounds the undecided list kept on the device; a larger count
    re-runs the screen once with a list of that size."""

    def __init__(self, grid, device="cuda", config=None, D=None, capacity=65536, keep_codes=False):
        self.grid = grid
        self.device = torch.device(device)
        self.config = CertificateConfig() if config is None else config
        nz, ny, nx = grid.shape
        self.D = None
        if self.config.prescreen:
            if D is None:
                try:
                    D = q1_coefficients(grid.spacing, "closed27_uniform")[0]
                except FloatingPointError:
                    D = None
            if D is not None:
                self.D = torch.as_tensor(np.ascontiguousarray(D, dtype=np.float64), device=self.device).contiguous()
        self.ws_bytes = int(_lib.load().seqvp_certificate_workspace_bytes(nx, ny, nz, int(self.config.tile_size)))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=self.device)
        self.capacity = int(max(1, min(capacity, grid.n_cells)))
        self.undecided = torch.empty(self.capacity, dtype=torch.int64, device=self.device)
        self.summary = torch.empty(5, dtype=torch.float64, device=self.device)
        self.codes = torch.empty(grid.n_cells, dtype=torch.uint8, device=self.device) if keep_codes else None
        self._u = None

    def screen(self, u, summary=None):
        """Enqueue the screen of u (nz, ny, nx, 3) float64 on the current stream. `summary`: 5 device doubles that
        receive the screen values instead of self.summary (the caller then passes their host copy to finish())."""
        if u.dtype != torch.float64 or tuple(u.shape) != tuple(self.grid.shape) + (3,) or not u.is_contiguous():
            raise InputValidationError("u must be a contiguous (nz, ny, nx, 3) float64 device tensor")
        nz, ny, nx = self.grid.shape
        hx, hy, hz = (float(h) for h in self.grid.spacing)
        _lib.call("seqvp_certificate_screen", _lib.ptr(u), _lib.ptr(self.D), hx, hy, hz, self.config.delta, nx, ny,
                  nz, int(self.config.tile_size), _lib.ptr(self.codes), _lib.ptr(self.undecided), self.capacity,
                  _lib.ptr(self.summary if summary is None else summary), _lib.ptr(self.ws), self.ws_bytes,
                  _lib.stream_handle(self.device))
        self._u = u

    def finish(self, summary=None):
        """The certificate of the last screened map. `summary` = the 5 screen values if the caller has
        already copied them to the host (saves one synchronisation)."""
        u = self._u
        if u is None:
            raise RuntimeError("screen() first")
        s = np.asarray(self.summary.tolist() if summary is None else summary, dtype=np.float64)
        n = int(s[4])
        if n > self.capacity:
            self.capacity = n
            self.undecided = torch.empty(n, dtype=torch.int64, device=self.device)
            self.screen(u)
            s = np.asarray(self.summary.tolist(), dtype=np.float64)
            n = int(s[4])
        cells = np.empty(0, dtype=np.int64)
        verts = np.empty((0, 8, 3))
        if n:
            # gather in list order on the device, then order by cell id on the host (tile order of the jobs)
            verts_d = torch.empty((n, 8, 3), dtype=torch.float64, device=self.device)
            nz, ny, nx = self.grid.shape
            _lib.call("seqvp_gather_vertices", _lib.ptr(u), _lib.ptr(self.undecided), n, nx, ny, nz,
                      _lib.ptr(verts_d), _lib.stream_handle(self.device))
            listed = self.undecided[:n].cpu().numpy()
            order = np.argsort(listed, kind="stable")
            cells = np.ascontiguousarray(listed[order])
            verts = np.ascontiguousarray(verts_d.cpu().numpy()[order])
        out = _lib.Certificate()
        _host_call("seqvp_certificate_finish", _lib.hptr(np.ascontiguousarray(s)), self.grid.n_cells,
                   _lib.hptr(verts), _lib.hptr(cells), n, _lib.hptr(_spacing(self.grid)), self.config.delta,
                   int(self.config.tile_size), int(self.config.max_depth), int(self.config.max_boxes),
                   int(self.config.fallback_workers), ctypes.byref(out))
        return _result(out)

    def __call__(self, u):
        self.screen(u)
        return self.finish()

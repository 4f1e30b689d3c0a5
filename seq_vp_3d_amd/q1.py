"""FusedQ1: the device Q1 cell sweeps behind the reference's fused-kernel interface.

Same methods, argument meaning and return layout as jd_curl/three_d/backends/triton_q1.py:244-748
(FusedQ1): soft_energy, barrier_energy, gradient_nodes, joint, diagonal_nodes; plus node_energy
(the image/proximity part of AcceleratedHybridBackend._energy, accelerated.py:1583-1600) and
node_min_jd (the T62 measure, torch_ops.py:687-773). All arrays are contiguous float64 device
tensors; node fields are (nz, ny, nx, 3). Every call goes to libseqvp_hip.so; there is no
host fallback.
"""
import numpy as np
import torch

from . import _lib
from .problem import InputValidationError, problem_constants

OP_SOFT, OP_BARRIER, OP_GRAD, OP_GGN, OP_JOINT, OP_ROOT, OP_DST, OP_DOTS, OP_NODEMIN, OP_NODE, OP_TRIAL = range(11)


class FusedQ1:
    """Constants and launches for one problem (grid, weights, barrier) on one device."""

    def __init__(self, problem, device="cuda"):
        grid, k, D, delta = problem_constants(problem)
        self.grid = grid
        self.shape = tuple(grid.shape)
        self.nz, self.ny, self.nx = self.shape
        self.device = torch.device(device)
        self.k = torch.tensor(k, dtype=torch.float64, device=self.device)
        self.D = torch.tensor(D, dtype=torch.float64, device=self.device)
        self.delta = delta
        self.barrier_weight = float(k[35])
        self.activation = float(k[37])
        self.ws = _lib.Workspace(self.device)
        self.ncell = grid.n_cells

    # ---- helpers -----------------------------------------------------------------------------
    def _check(self, *arrays):
        for a in arrays:
            if (not isinstance(a, torch.Tensor) or a.dtype != torch.float64 or not a.is_contiguous()
                    or tuple(a.shape) != self.shape + (3,) or a.device != self.device):
                raise InputValidationError("fused kernels need contiguous float64 device node arrays")

    def _dims(self):
        return self.nx, self.ny, self.nz

    def _workspace(self, op):
        buf, n = self.ws.get(_lib.workspace_bytes(op, self.shape))
        return _lib.ptr(buf), n

    # ---- energies ----------------------------------------------------------------------------
    def soft_energy(self, u, ref, base=None):
        """[JD penalty, curl penalty, stable soft difference vs base] (triton_q1.py:572-581)."""
        self._check(u, ref) if base is None else self._check(u, ref, base)
        out = torch.empty(3, dtype=torch.float64, device=self.device)
        w, n = self._workspace(OP_SOFT)
        _lib.call("seqvp_q1_soft_energy", _lib.ptr(u), _lib.ptr(ref), _lib.ptr(base), _lib.ptr(self.k),
                  *self._dims(), _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def barrier_energy(self, u, base=None):
        """[barrier value, min J, unweighted barrier change, min base J] (triton_q1.py:680-691)."""
        self._check(u) if base is None else self._check(u, base)
        out = torch.empty(4, dtype=torch.float64, device=self.device)
        w, n = self._workspace(OP_BARRIER)
        _lib.call("seqvp_q1_barrier_energy", _lib.ptr(u), _lib.ptr(base), _lib.ptr(self.k), *self._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def trial_energy(self, u, ref, base=None, out=None, base_min=None, quiet=False):
        """soft_energy(u, ref, base) and barrier_energy(u, base) from one sweep: [jd, curl, difference, barrier
        value, min J, unweighted barrier change, min base J], the same bits as the two calls. base_min: a device
        double holding the base's least closed27 J when known; quiet: it is at or above the activation (the sweep
        then skips the base's closed27 J everywhere; otherwise where the base cell's edges bound it above both)."""
        self._check(u, ref) if base is None else self._check(u, ref, base)
        out = torch.empty(7, dtype=torch.float64, device=self.device) if out is None else out
        w, n = self._workspace(OP_TRIAL)
        if base_min is None or base is None:
            _lib.call("seqvp_q1_trial_energy", _lib.ptr(u), _lib.ptr(ref), _lib.ptr(base), _lib.ptr(self.k),
                      *self._dims(), _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        else:
            _lib.call("seqvp_q1_trial_energy_base_min", _lib.ptr(u), _lib.ptr(ref), _lib.ptr(base),
                      _lib.ptr(base_min), int(bool(quiet)), _lib.ptr(self.k), *self._dims(), _lib.ptr(out), w, n,
                      _lib.stream_handle(self.device))
        return out

    def soft_energy_f32(self, u, ref, base=None):
        """FP32 screening [jd, curl, difference] (triton_q1.py:652-662), Kahan-compensated."""
        self._check(u, ref) if base is None else self._check(u, ref, base)
        out = torch.empty(3, dtype=torch.float64, device=self.device)
        w, n = self._workspace(OP_SOFT)
        _lib.call("seqvp_q1_soft_energy_f32", _lib.ptr(u), _lib.ptr(ref), _lib.ptr(base), _lib.ptr(self.k),
                  *self._dims(), _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def barrier_energy_f32(self, u, base=None):
        """FP32 screening [value, min J, unweighted change, min base J] (triton_q1.py:663-678)."""
        self._check(u) if base is None else self._check(u, base)
        out = torch.empty(4, dtype=torch.float64, device=self.device)
        w, n = self._workspace(OP_BARRIER)
        _lib.call("seqvp_q1_barrier_energy_f32", _lib.ptr(u), _lib.ptr(base), _lib.ptr(self.k), *self._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def node_energy(self, warped, fixed, u, ref, lam, data_weight=None, base_warped=None, base_u=None, out=None):
        """[image, proximity, image difference vs base, proximity difference vs base]."""
        self._check(u, ref)
        for a in (warped, fixed):
            if a.dtype != torch.float64 or tuple(a.shape) != self.shape or not a.is_contiguous():
                raise InputValidationError("node scalar fields must be contiguous float64 (nz, ny, nx)")
        out = torch.empty(4, dtype=torch.float64, device=self.device) if out is None else out
        w, n = self._workspace(OP_NODE)
        hx, hy, hz = self.grid.spacing
        _lib.call("seqvp_node_energy", _lib.ptr(warped), _lib.ptr(fixed), _lib.ptr(data_weight), _lib.ptr(u),
                  _lib.ptr(ref), _lib.ptr(base_warped), _lib.ptr(base_u), float(lam), hx, hy, hz, *self._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def node_terms(self, warped, fixed, image_grad, u, ref, lam, data_weight=None, image=None, total=None, ggn=None):
        """Node terms: image = qd (W-F) b, total = image + lam q (u - ref), ggn = qd b^2 + lam q (fills the
        given outputs)."""
        hx, hy, hz = self.grid.spacing
        _lib.call("seqvp_node_terms", _lib.ptr(warped), _lib.ptr(fixed), _lib.ptr(data_weight), _lib.ptr(image_grad),
                  _lib.ptr(u), _lib.ptr(ref), float(lam), hx, hy, hz, *self._dims(), _lib.ptr(image), _lib.ptr(total),
                  _lib.ptr(ggn), _lib.stream_handle(self.device))

    # ---- derivatives -------------------------------------------------------------------------
    def gradient_nodes(self, base, u, ref, *, soft=True, barrier=True, out=None):
        """base + Gauss soft VJP + closed27 barrier VJP on the nodes (triton_q1.py:504-518)."""
        self._check(u, base) if not soft else self._check(u, ref, base)
        out = torch.empty_like(u) if out is None else out
        w, n = self._workspace(OP_GRAD)
        _lib.call("seqvp_q1_gradient", _lib.ptr(base), _lib.ptr(u), _lib.ptr(ref if soft else None), int(soft),
                  int(barrier), _lib.ptr(self.k), *self._dims(), _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def gradient_free(self, warped, fixed, image_grad, u, ref, lam, data_weight=None, out=None):
        """restrict(node terms total + Gauss soft VJP + closed27 barrier VJP) in one call: the energy gradient on
        the free nodes, component-major (3 nz-2 ny-2 nx-2 values)."""
        self._check(u, ref, image_grad)
        for a in (warped, fixed) + (() if data_weight is None else (data_weight,)):
            if a.dtype != torch.float64 or tuple(a.shape) != self.shape or not a.is_contiguous():
                raise InputValidationError("node scalar fields must be contiguous float64 (nz, ny, nx)")
        nz, ny, nx = self.shape
        F = (nx - 2) * (ny - 2) * (nz - 2)
        out = torch.empty(3 * F, dtype=torch.float64, device=self.device) if out is None else out
        if out.dtype != torch.float64 or out.numel() != 3 * F or not out.is_contiguous():
            raise InputValidationError("out must be a contiguous float64 vector of 3 (nz-2)(ny-2)(nx-2) values")
        w, n = self._workspace(OP_GRAD)
        hx, hy, hz = self.grid.spacing
        _lib.call("seqvp_q1_gradient_free", _lib.ptr(warped), _lib.ptr(fixed), _lib.ptr(data_weight),
                  _lib.ptr(image_grad), _lib.ptr(u), _lib.ptr(ref), float(lam), hx, hy, hz, _lib.ptr(self.k),
                  *self._dims(), _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def joint_free(self, warped, fixed, image_grad, u, ref, lam, data_weight=None, out=None):
        """(gd, gf, ge) on the free nodes in one call: restrict(image term), restrict(image term + barrier VJP),
        restrict(total term + soft VJP + barrier VJP). `out` = three contiguous float64 vectors of 3F values."""
        self._check(u, ref, image_grad)
        for a in (warped, fixed) + (() if data_weight is None else (data_weight,)):
            if a.dtype != torch.float64 or tuple(a.shape) != self.shape or not a.is_contiguous():
                raise InputValidationError("node scalar fields must be contiguous float64 (nz, ny, nx)")
        nz, ny, nx = self.shape
        F = (nx - 2) * (ny - 2) * (nz - 2)
        if out is None:
            out = tuple(torch.empty(3 * F, dtype=torch.float64, device=self.device) for _ in range(3))
        for o in out:
            if o.dtype != torch.float64 or o.numel() != 3 * F or not o.is_contiguous():
                raise InputValidationError("outputs must be contiguous float64 vectors of 3 (nz-2)(ny-2)(nx-2) values")
        w, n = self._workspace(OP_JOINT)
        hx, hy, hz = self.grid.spacing
        _lib.call("seqvp_q1_joint_free", _lib.ptr(warped), _lib.ptr(fixed), _lib.ptr(data_weight),
                  _lib.ptr(image_grad), _lib.ptr(u), _lib.ptr(ref), float(lam), hx, hy, hz, _lib.ptr(self.k),
                  *self._dims(), *(_lib.ptr(o) for o in out), w, n, _lib.stream_handle(self.device))
        return out

    def joint(self, image, total_base, u, ref):
        """(outer = image + barrier, total = total_base + soft + barrier) sharing the cell work
        (triton_q1.py:563-570)."""
        self._check(image, total_base, u, ref)
        outer, total = torch.empty_like(u), torch.empty_like(u)
        w, n = self._workspace(OP_JOINT)
        _lib.call("seqvp_q1_joint", _lib.ptr(image), _lib.ptr(total_base), _lib.ptr(u), _lib.ptr(ref),
                  _lib.ptr(self.k), *self._dims(), _lib.ptr(outer), _lib.ptr(total), w, n,
                  _lib.stream_handle(self.device))
        return outer, total

    def diagonal_nodes(self, base, u):
        """base + exact GGN diagonal of the geometry and barrier blocks (triton_q1.py:442-458)."""
        self._check(base, u)
        out = torch.empty_like(u)
        w, n = self._workspace(OP_GGN)
        _lib.call("seqvp_q1_ggn_diagonal", _lib.ptr(base), _lib.ptr(u), _lib.ptr(self.k), *self._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(self.device))
        return out

    def node_min_jd(self, u, extension_activation=None, stats=None):
        """T62 node JD measure (closed27 + exact edge minima), full grid shape, +inf off-grid.

        With extension_activation (in (0, 1]) the T69 face/interior extension acts on the candidate cells
        (jd_affine.py:154-225). stats: None, or a device tensor of 4 doubles that receives the number of
        candidates, of cells lowered by more than 0.1 %, of cells lowered at all, and the least estimate."""
        self._check(u)
        out = torch.empty(self.shape, dtype=torch.float64, device=self.device)
        if extension_activation is None:
            w, n = self._workspace(OP_NODEMIN)
            _lib.call("seqvp_node_min_jd", _lib.ptr(u), _lib.ptr(self.k), *self._dims(), _lib.ptr(out), w, n,
                      _lib.stream_handle(self.device))
            return out
        if stats is not None and (stats.dtype != torch.float64 or stats.numel() < 4 or stats.device != self.device):
            raise ValueError("stats must be a float64 device tensor of 4 entries")
        buf, n = self.ws.get(_lib.raw("seqvp_node_min_jd_extended_workspace_bytes", *self._dims()))
        _lib.call("seqvp_node_min_jd_extended", _lib.ptr(u), _lib.ptr(self.k), *self._dims(),
                  float(extension_activation), _lib.ptr(out), _lib.ptr(stats), _lib.ptr(buf), n,
                  _lib.stream_handle(self.device))
        return out

    @staticmethod
    def face_interior_minima(J27):
        """T69 estimate per row of a (n, 27) float64 device tensor (jd_affine.py:130-151)."""
        J27 = J27.contiguous()
        if J27.dtype != torch.float64 or J27.dim() != 2 or J27.shape[1] != 27:
            raise ValueError("closed27 JD values must be a float64 (cells, 27) tensor")
        mu = torch.empty(J27.shape[0], dtype=torch.float64, device=J27.device)
        _lib.call("seqvp_face_interior_minima", _lib.ptr(J27), J27.shape[0], _lib.ptr(mu),
                  _lib.stream_handle(J27.device))
        return mu


class FusedJDRoot:
    """T66 global closed27 first JD root (triton_jd_root.py:26-52): (root in s-units or inf,
    number of points with c0 <= 0, number of points)."""

    def __init__(self, problem, device="cuda", q1=None):
        self.q1 = q1 if q1 is not None else FusedQ1(problem, device)

    def __call__(self, u, v, limit):
        q1 = self.q1
        q1._check(u, v)
        if not (np.isfinite(limit) and limit > 0):
            raise InputValidationError("limit must be positive and finite")
        out = torch.empty(3, dtype=torch.float64, device=q1.device)
        w, n = q1._workspace(OP_ROOT)
        _lib.call("seqvp_jd_root", _lib.ptr(u), _lib.ptr(v), _lib.ptr(q1.k), float(limit), *q1._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(q1.device))
        root, invalid, bad = out.tolist()
        if bad:
            raise FloatingPointError("T66 nonfinite cubic coefficients (fused)")
        return float(root), int(invalid), 27 * q1.ncell

    def from_free(self, u, p, limit):
        """The same bound with the direction as its free-node vector p (no node field)."""
        return self.read(self.from_free_launch(u, p, limit))

    def from_free_launch(self, u, p, limit):
        """Enqueues from_free's kernels on the device's stream and returns the pending result for read()."""
        q1 = self.q1
        q1._check(u)
        nz, ny, nx = q1.shape
        if (p.dtype != torch.float64 or not p.is_contiguous() or p.device != q1.device
                or p.numel() != 3 * (nx - 2) * (ny - 2) * (nz - 2)):
            raise InputValidationError("p must be the contiguous float64 free-node vector")
        if not (np.isfinite(limit) and limit > 0):
            raise InputValidationError("limit must be positive and finite")
        out = torch.empty(3, dtype=torch.float64, device=q1.device)
        w, n = q1._workspace(OP_ROOT)
        _lib.call("seqvp_jd_root_free", _lib.ptr(u), _lib.ptr(p), _lib.ptr(q1.k), float(limit), *q1._dims(),
                  _lib.ptr(out), w, n, _lib.stream_handle(q1.device))
        return out

    def read(self, out):
        """(root, invalid, points) of a from_free_launch result (waits for it)."""
        root, invalid, bad = out.tolist()
        if bad:
            raise FloatingPointError("T66 nonfinite cubic coefficients (fused)")
        return float(root), int(invalid), 27 * self.q1.ncell

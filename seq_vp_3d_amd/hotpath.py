"""Device-resident state of one registration pair and the hot-path operations on it.

Every array of the pair lives in HBM for the whole solve (fields, images, spline coefficients,
metric, direction, BB history); each operation below is one or a few libseqvp_hip.so calls, and
only scalars (energies, roots, dot products, counts) come back to the host -- the reference moves
160-165 MB per gradient, direction and promotion across PCIe (accelerated.py:1525, 2801-2874;
hybrid.py:262-286).

Operations and the reference code they follow:
  metric()     -- _pg_metric_inverse with the image_shift_ggn_equilibrated_jd_affine policies
                  (poisson_gradient.py:400-576): the T62 node measure, with the T69 face/interior extension
                  on candidate cells when asked (jd_affine.py:154-225)
  gradient()   -- TorchLinearization.gradient (torch_ops.py:559-589) restricted to theta
  direction()  -- p = -S M^-1 S g (poisson_gradient.py:1241-1262, hybrid.py:262-286)
  jd_bound()   -- T66 global root on the direction (torch_ops.py:636-673)
  witness_coefficients() -- the cubics of the historical-witness predictor (jd_step.py:177-239)
  trial()      -- evaluate_trial: candidate map, whole-cell JD certificate (GPU edge hull + prescreen of every
                  cell, CPU interval Bernstein fallback of the undecided ones), sampling, energies vs base
                  (accelerated.py:1666-1730, 1574-1664; feasibility.py:571-769)
  promote()    -- promote_trial: the gradient sampling of an accepted trial (accelerated.py:2850-2874)
"""
import ctypes
from dataclasses import dataclass

import numpy as np
import torch

from . import _lib
from .poisson import DevicePoisson, SpectralMetric
from .certificate import CertificateConfig, DeviceCertificate
from .graphs import SegmentGraphs, key_of
from .jd_step import nominate_witness
from .q1 import OP_ROOT, OP_TRIAL, FusedJDRoot, FusedQ1
from .spline import DeviceSpline

JD_AFFINE_FLOOR = 1e-12
JD_AFFINE_ACTIVATION = .25
# FP32 screening band per term, relative to that term's values of the two states compared
# (accelerated.py:67-76): 64 ulps of binary32 for the JD/curl penalties, 4096 for the barrier.
FP32_EPS = 2. ** -24
FP32_BAND_SOFT = 64.
FP32_BAND_BARRIER = 4096.


@dataclass
class Energies:
    image: float
    proximity: float
    jd: float
    curl: float
    barrier: float

    @property
    def total(self):
        return self.image + self.proximity + self.jd + self.curl + self.barrier

    @property
    def outer(self):
        return self.image + self.barrier


@dataclass
class TrialResult:
    ok: bool
    reason: str
    energy: Energies | None
    total_difference: float = float("nan")     # E(trial) - E(base), stable
    outer_difference: float = float("nan")     # F(trial) - F(base), stable
    min_gap: float = float("nan")
    unproven_cells: int = 0                    # n_uncertain + n_infeasible of a failed certificate
    band: float | None = None                  # FP32 screening band of the differences (None: FP64 values)
    certificate: object = None                 # FeasibilityResult of the candidate (None: not reached)
    witness: tuple | None = None               # confirmed (cell, reference point) with J <= delta


class PairState:
    """Device buffers of one pair on one grid (node fields (nz, ny, nx, 3), free vectors (3F,)).

    The trial and the direction of each inner iteration are segments: library calls on fixed buffers that end with
    one copy of their scalars to page-locked host memory and one wait. With `graphs` each runs as a captured HIP graph
    (seq_vp_3d_amd.graphs), else its calls are issued one by one (the default: on MI355X the issued calls already keep
    the GPU busy from 96^3 up, and the graph launches measured 0-1% slower), with the same bits."""

    def __init__(self, problem, fixed, moving, device="cuda", data_weight=None, sampler=None, certificate=None,
                 graphs=False, batch=None):
        self.problem = problem
        self.grid = problem.grid
        self.device = torch.device(device)
        self.q1 = FusedQ1(problem, self.device)
        self.root = FusedJDRoot(problem, q1=self.q1)
        g = self.grid
        self.shape = tuple(g.shape)
        self.nfree = 3 * g.n_free_scalar
        t = lambda a: torch.as_tensor(a, dtype=torch.float64, device=self.device).contiguous()
        self.fixed = t(fixed)
        self.data_weight = None if data_weight is None else t(data_weight)
        if isinstance(sampler, DeviceSpline):           # shared device spline of the same moving image
            self.spline = sampler
        elif sampler is not None:                       # host coefficients (build_image_sampler)
            self.spline = DeviceSpline(sampler, g, self.device)
        else:                                           # coefficients computed on the device
            self.spline = DeviceSpline.from_image(t(moving), g, self.device)
        z = lambda *s: torch.zeros(s, dtype=torch.float64, device=self.device)
        self.ref = z(*self.shape, 3)                 # u_ref (proximal centre)
        self.u = z(*self.shape, 3)                   # accepted state
        self.theta = z(self.nfree)
        self.warped = z(*self.shape)
        self.bgrad = z(*self.shape, 3)
        self.u_t = z(*self.shape, 3)                 # trial
        self.theta_t = z(self.nfree)
        self.warped_t = z(*self.shape)
        self.bgrad_t = z(*self.shape, 3)             # image gradient at the trial (promotion swaps it in)
        self.node_a = z(*self.shape, 3)
        self.node_b = z(*self.shape, 3)
        self.g = z(self.nfree)
        self.p = z(self.nfree)
        self.v = z(*self.shape, 3)
        self.scale = z(self.nfree)
        self.diag = z(self.nfree)
        self.bad = torch.zeros(1, dtype=torch.int64, device=self.device)
        self.dot_out = z(4)
        # slope | gradient norm^2 | - | - | BB2 products (4) | step bound | step ratio (2) | -, read with one
        # synchronisation
        self.scalars = z(12)
        self._prev = None           # (theta, g, p) of the last accepted state, kept for the next BB2 products
        self._ratio_pair = None     # (new theta, old theta) of the last accepted step, its ratio not computed yet
        # The accepted state's version (bumped whenever it changes), the version at which the reference was set to
        # it, and the version whose outer gradient (image + barrier) the joint pass left in diag.
        self._version, self._ref_at, self._gf_at = 0, None, None
        self.dot_ws = torch.empty(max(_lib.workspace_bytes(7, self.shape), 256), dtype=torch.uint8, device=self.device)
        config = certificate if certificate is not None else CertificateConfig(delta=problem.delta)
        if config.delta != problem.delta:
            raise ValueError("certificate delta must equal the problem's delta")
        self.cert = DeviceCertificate(g, self.device, config)
        self._certificate = None    # certificate of the current candidate (kept across its energy evaluations)
        self.certificate = None     # certificate of the accepted state (None: not computed yet)
        self._witness_layout = None   # (key, cells, points) on the device, the last witness layout
        self.poisson = None
        self._den = None
        self.energy = None          # Energies of the accepted state
        self.metric_info = {}
        self._ext_stats = None
        # Segment buffers. Trial: node energies (4) | soft + barrier (7) | invalid-sample count (int64 bits) |
        # certificate screen (5); the step comes from a page-locked word through the segment's first copy.
        self._tau_host = torch.zeros(1, dtype=torch.float64, pin_memory=True)
        self._tau_dev = z(1)
        self._trial_dev = z(17)
        # the chained first attempt: [step, root sweep out (3), capped step] on the device and its host copy
        self._chain_dev = z(5)
        self._chain_host = torch.zeros(5, dtype=torch.float64, pin_memory=True)
        self._chain_np = self._chain_host.numpy()
        self.chain = True           # chain_launch enqueues (False: it declines, and the driver takes the host path)
        self.chained_trials = 0     # chained first attempts whose trial the driver used
        # least closed27 J of the accepted state when known (host float and device copy, the copy written before
        # its next use): at or above the activation, the trial sweep needs no closed27 J of the accepted state
        self._base_min, self._base_min_dev, self._base_min_stale, self._trial_min = None, z(1), False, None
        self._ref_min, self._ref_min_dev = None, z(1)   # the same for the reference, when it was set to a state
        self._trial_host = torch.zeros(17, dtype=torch.float64, pin_memory=True)
        self._scalars_host = torch.zeros(12, dtype=torch.float64, pin_memory=True)
        self._tau_np, self._trial_np, self._scalars_np = (t.numpy() for t in (self._tau_host, self._trial_host,
                                                                                 self._scalars_host))
        # the shared workspace at its largest size from the start, so that no call reallocates it under a captured
        # graph's addresses
        sizes = [_lib.workspace_bytes(op, self.shape) for op in range(OP_TRIAL + 1)]
        nz, ny, nx = self.shape
        sizes.append(_lib.raw("seqvp_node_min_jd_extended_workspace_bytes", nx, ny, nz))
        self.q1.ws.get(max(sizes))
        self.graphs = SegmentGraphs(self.device) if graphs else None
        # batching form: the trial-energy sweep goes through this BatchGroup's multi-pair launches
        self.batch = batch

    def _segment(self, key, body):
        """Runs a device-only segment: its captured graph for `key`, or its calls one by one without graphs."""
        if self.graphs is None:
            body()
        else:
            self.graphs.run(key, body)

    def _wait(self):
        torch.cuda.current_stream(self.device).synchronize()

    # ---- small helpers -------------------------------------------------------------------------
    def _stream(self):
        return _lib.stream_handle(self.device)

    def dots(self, *pairs):
        """Compensated dot products of (x, y) free-vector pairs (at most 4), as host floats."""
        xs = (ctypes.c_void_p * len(pairs))(*(a.data_ptr() for a, _ in pairs))
        ys = (ctypes.c_void_p * len(pairs))(*(b.data_ptr() for _, b in pairs))
        n = pairs[0][0].numel()
        _lib.call("seqvp_dots", xs, ys, len(pairs), n, _lib.ptr(self.dot_out), _lib.ptr(self.dot_ws),
                  self.dot_ws.numel(), self._stream())
        return self.dot_out[:len(pairs)].tolist()

    def inject(self, theta, out):
        nz, ny, nx = self.shape
        _lib.call("seqvp_inject", _lib.ptr(theta), nx, ny, nz, _lib.ptr(out), self._stream())
        return out

    def restrict(self, field, out):
        nz, ny, nx = self.shape
        _lib.call("seqvp_restrict", _lib.ptr(field), nx, ny, nz, _lib.ptr(out), self._stream())
        return out

    def set_state(self, theta=None, ref=None, certificate=None):
        """Set the accepted map from theta (and optionally the reference) and evaluate it. `certificate` is the
        map's certificate when the caller has one (else state_certificate() computes it on demand)."""
        self.certificate = certificate
        self._version += 1
        if ref is not None:
            self.ref.copy_(torch.as_tensor(ref, dtype=torch.float64))
            self._ref_at = None
            self._ref_min = None
        if theta is not None:
            self.theta.copy_(torch.as_tensor(theta, dtype=torch.float64))
        self.inject(self.theta, self.u)
        _, _, bad = self.spline.warp(self.u, value=self.warped, gradient=self.bgrad)
        if int(bad.item()):
            raise FloatingPointError("state outside the sampling support")
        self.energy = self.evaluate_energy(self.u, self.warped)
        self._set_base_min(self._evaluated_min)

    def _set_base_min(self, value):
        self._base_min = value
        self._base_min_stale = value is not None

    def _base_min_ptr(self):
        """The device double holding the accepted state's least closed27 J, written now if it changed (None when
        unknown). Called outside captured segments: a graph replays the value it captured."""
        if self._base_min is None:
            return None
        if self._base_min_stale:
            self._base_min_dev.fill_(self._base_min)
            self._base_min_stale = False
        return self._base_min_dev

    def _quiet(self):
        """Whether the accepted state has no active closed27 point (known from its last evaluation)."""
        return self._base_min is not None and self._base_min >= self.q1.activation

    def state_certificate(self):
        """The certificate of the accepted map: the one its trial or transfer passed, else screened now."""
        if self.certificate is None:
            self.cert.screen(self.u)
            self.certificate = self.cert.finish()
        return self.certificate

    def evaluate_energy(self, u, warped):
        q = self.q1
        node = q.node_energy(warped, self.fixed, u, self.ref, self.problem.weights.lam, data_weight=self.data_weight)
        sb = q.trial_energy(u, self.ref)
        n, v = node.tolist(), sb.tolist()
        s, b = v[0:3], v[3:7]
        self._evaluated_min = b[1]
        if not b[1] > q.delta:
            raise FloatingPointError("device barrier points contradict a feasible state")
        return Energies(n[0], n[1], s[0], s[1], b[0])

    # ---- metric --------------------------------------------------------------------------------
    def metric(self, *, shift_factor=1., jd_affine=True, activation=JD_AFFINE_ACTIVATION, extension=True):
        """P^-1 = S M^-1 S with M the shifted Q1 Poisson metric (poisson_gradient.py:269-293, 400-576); S carries
        the JD-affine factors of the node measure, with the T69 face/interior extension when `extension`."""
        q, w = self.q1, self.problem.weights
        q.node_terms(self.warped, self.fixed, self.bgrad, self.u, self.ref, w.lam, data_weight=self.data_weight,
                     ggn=self.node_a)
        diag_nodes = q.diagonal_nodes(self.node_a, self.u)
        self._gf_at = None
        self.restrict(diag_nodes, self.diag)
        # alpha = factor * mean over free nodes of (q/omega) |b|^2 / 3 (q/omega = 1 on free nodes)
        self.restrict(self.bgrad, self.p)          # p is free until direction(); g holds the gradient
        (bb,) = self.dots((self.p, self.p))
        alpha = shift_factor * (bb / (self.nfree // 3)) / 3
        a = w.lam + alpha
        bp = w.mu / 2 + w.nu / 2
        if not (np.isfinite(a) and a >= 0):
            raise FloatingPointError("shifted metric not representable")
        if self._den is None:
            mz, my, mx = self.grid.interior_shape
            self._den = torch.empty((mz, my, mx), dtype=torch.float64, device=self.device)
        self.poisson = DevicePoisson(SpectralMetric(self.grid, a, bp, "q1"), self.device, ws=q.ws, out=self._den)
        inv_h2 = sum(1 / h ** 2 for h in self.grid.spacing)
        m = self.grid.omega * (a + bp * (8 / 9) * inv_h2)
        stats = None
        if jd_affine and extension:
            if self._ext_stats is None:
                self._ext_stats = torch.empty(4, dtype=torch.float64, device=self.device)
            stats = self._ext_stats
            node = q.node_min_jd(self.u, extension_activation=activation, stats=stats)
        else:
            node = q.node_min_jd(self.u) if jd_affine else None
        nz, ny, nx = self.shape
        _lib.call("seqvp_metric_scale", _lib.ptr(self.diag), _lib.ptr(node), float(m), float(activation),
                  JD_AFFINE_FLOOR, nx, ny, nz, _lib.ptr(self.scale), _lib.ptr(self.bad), self._stream())
        # one synchronisation for the validity flag and the extension record
        flags = (self.bad.to(torch.float64) if stats is None else torch.cat((self.bad.to(torch.float64), stats))).tolist()
        if int(flags[0]):
            raise FloatingPointError("GGN diagonal or node JD invalid")
        self.metric_info = dict(shift_alpha=alpha, a=a, b=bp, mass_diagonal=m)
        if stats is not None:
            cand, lowered, lowered_any, mu_min = flags[1:5]
            self.metric_info["jd_affine"] = dict(activation=activation, extension=dict(
                triggered="always", candidates=int(cand), lowered_cells=int(lowered),
                lowered_cells_any=int(lowered_any), mu_min=mu_min))
        return self.metric_info

    # ---- gradient / direction / bound ------------------------------------------------------------
    def gradient(self):
        """Full energy gradient on theta (device), g = restrict(node terms + soft + barrier VJPs).

        Right after a self-reference refresh (ref = u) the proximity term lam q (u - ref) and every JD/curl change
        residual J - J_ref, c - c_ref are exactly zero, so the total gradient is the outer gradient (image +
        barrier) that outer_norms computed for the same state: that vector is taken instead of a second sweep (the
        first gradient of each inner solve of the outer loop). J and J_ref come out of the same fused
        multiply-adds, so the sweep's soft contributions are exact zeros there and both passes add the same barrier
        contributions in the same order to the same image term: the vectors are the same bits."""
        if not self._take_outer_gradient():
            self._gradient_sweep()
        return self.g

    def _take_outer_gradient(self):
        """Takes the joint pass's outer gradient as g when it is the total gradient of the current state."""
        if self._gf_at == self._version and self._ref_at == self._version:
            self.g, self.diag = self.diag, self.g
            self._gf_at = None
            return True
        return False

    def _gradient_sweep(self):
        q, w = self.q1, self.problem.weights
        q.gradient_free(self.warped, self.fixed, self.bgrad, self.u, self.ref, w.lam, data_weight=self.data_weight,
                        out=self.g)

    def direction(self, fp32=False):
        """p = -S M^-1 (S g)."""
        return self.poisson.equilibrated_direction(self.scale, self.g, out=self.p, fp32=fp32)

    def jd_bound(self, limit):
        return self.root.from_free(self.u, self.p, limit)

    def jd_bound_launch(self, limit):
        """jd_bound's kernels enqueued; jd_bound_read(pending) returns its result (the host can work meanwhile)."""
        return self.root.from_free_launch(self.u, self.p, limit)

    def jd_bound_read(self, pending):
        return self.root.read(pending)

    def witness_coefficients(self, cells, points):
        """(n, 4) host array: det(A + s V) - delta at reference points of cells, A = I + grad u of the accepted
        state, V = grad of the direction p (jd_step.py:13-17, 204-222)."""
        self.inject(self.p, self.v)
        cells = np.ascontiguousarray(cells, dtype=np.int64)
        points = np.ascontiguousarray(points, dtype=np.float64)
        key = (cells.tobytes(), points.tobytes())
        if self._witness_layout is None or self._witness_layout[0] != key:
            # the witness layout changes only when a witness is added: uploaded once per layout
            self._witness_layout = (key, torch.as_tensor(cells).to(self.device), torch.as_tensor(points).to(self.device))
        _, c, pts = self._witness_layout
        out = torch.empty((len(cells), 4), dtype=torch.float64, device=self.device)
        nz, ny, nx = self.shape
        _lib.call("seqvp_witness_coefficients", _lib.ptr(self.u), _lib.ptr(self.v), _lib.ptr(self.q1.k), nx, ny, nz,
                  _lib.ptr(c), _lib.ptr(pts), len(cells), _lib.ptr(out), self._stream())
        return out.cpu().numpy()

    # ---- trial and promotion -----------------------------------------------------------------------
    def trial(self, tau, screen=False):
        """Candidate theta + tau p: certificate screen, sampling and energies against the accepted state.
        screen=True evaluates the soft and barrier terms with the FP32 screening kernels and returns the
        screening band of the differences (accelerated.py:1603-1612, 67-76)."""
        self._certificate = None
        if screen:
            nz, ny, nx = self.shape
            _lib.call("seqvp_axpy_inject", _lib.ptr(self.theta), _lib.ptr(self.p), float(tau), nx, ny, nz,
                      _lib.ptr(self.theta_t), _lib.ptr(self.u_t), self._stream())
            self.cert.screen(self.u_t)
            _, _, self._trial_bad = self.spline.warp(self.u_t, value=self.warped_t, gradient=self.bgrad_t)
            return self.trial_energies(screen)
        # FP64 trial as one segment: the step through the page-locked word, the candidate, the certificate screen,
        # the warp (value and gradient at once: the 64-tap gather dominates the warp, and an accepted trial then
        # needs no second warp at promotion), the energies, and one copy of the 17 scalars to the host
        self._tau_np[0] = tau
        batched = self.batch is not None
        known = self._base_min_ptr() is not None     # writes the device copy now, outside any captured segment
        quiet = self._quiet()
        key = None if self.graphs is None else key_of(
            "trial", batched, known, quiet, self.theta, self.p, self.u, self.warped, self.ref, self.u_t, self.theta_t,
            self.warped_t, self.bgrad_t, self.fixed, self.data_weight, self.q1.ws.buf, self.cert.undecided,
            self.cert.capacity)
        self._segment(key, lambda: self._trial_segment(not batched, known, quiet))
        if batched:
            # the soft and barrier sweep in the group's multi-pair launch, then the scalars' copy
            self.batch.trial_energy(self, self._trial_dev[4:11])
            _lib.call("seqvp_copy_async", ctypes.c_void_p(self._trial_host.data_ptr()), _lib.ptr(self._trial_dev),
                      17 * 8, self._stream())
        self.cert._u = self.u_t
        return self._trial_read()

    def _trial_read(self):
        self._wait()
        vals = self._trial_np.tolist()
        nbad = int(self._trial_np[11:12].view(np.int64)[0])
        self._trial_min = vals[8]
        return self._trial_result(vals[0:4], vals[4:7], vals[7:11], nbad, vals[12:17], False)

    def _trial_segment(self, energies=True, known=False, quiet=False, tau_dev=None):
        """The trial's device work; the step from the page-locked word, or from tau_dev (one device double) when
        given."""
        nz, ny, nx = self.shape
        q, w, s = self.q1, self.problem.weights, self._stream()
        d = self._trial_dev
        if tau_dev is None:
            tau_dev = self._tau_dev
            _lib.call("seqvp_copy_async", _lib.ptr(tau_dev), ctypes.c_void_p(self._tau_host.data_ptr()), 8, s)
        _lib.call("seqvp_axpy_inject_dev", _lib.ptr(self.theta), _lib.ptr(self.p), _lib.ptr(tau_dev), nx, ny, nz,
                  _lib.ptr(self.theta_t), _lib.ptr(self.u_t), s)
        self.cert.screen(self.u_t, summary=d[12:17])
        self.spline.warp(self.u_t, value=self.warped_t, gradient=self.bgrad_t, status=d[11:12].view(torch.int64))
        q.node_energy(self.warped_t, self.fixed, self.u_t, self.ref, w.lam, data_weight=self.data_weight,
                      base_warped=self.warped, base_u=self.u, out=d[0:4])
        if energies:
            q.trial_energy(self.u_t, self.ref, self.u, out=d[4:11],      # soft and barrier terms in one sweep
                           base_min=self._base_min_dev if known else None, quiet=quiet)
            _lib.call("seqvp_copy_async", ctypes.c_void_p(self._trial_host.data_ptr()), _lib.ptr(d), 17 * 8, s)

    def trial_energies(self, screen=False):
        """Energies and stable differences of the current candidate (u_t, warped_t) against the accepted state,
        after trial(tau, screen=True) (its sampling count and certificate screen)."""
        q, w = self.q1, self.problem.weights
        node = q.node_energy(self.warped_t, self.fixed, self.u_t, self.ref, w.lam, data_weight=self.data_weight,
                             base_warped=self.warped, base_u=self.u)
        if screen:
            sb = torch.cat((q.soft_energy_f32(self.u_t, self.ref, self.u), q.barrier_energy_f32(self.u_t, self.u)))
        else:
            sb = q.trial_energy(self.u_t, self.ref, self.u,       # soft and barrier terms in one sweep
                                base_min=self._base_min_ptr(), quiet=self._quiet())
        vals = torch.cat((node, sb, self._trial_bad.to(torch.float64), self.cert.summary)).tolist()
        self._trial_min = None if screen else vals[8]      # the FP32 screen's min J is not the closed27 J
        return self._trial_result(vals[0:4], vals[4:7], vals[7:11], vals[11], vals[12:17], screen)

    # ---- the chained first attempt ---------------------------------------------------------------------
    def chain_launch(self, prev_tau, step_min, step_max, initial_step, safety, has_prev=True, root=True):
        """Enqueues, behind the pending direction segment, the next line search's first attempt as the host driver
        would take it when its witness memory is empty: the step from the direction scalars (seqvp_pg_step), with
        `root` the T66 root on that limit and the step capped by the root, and the trial segment on that step.
        Returns the pending chain (None when this state does not chain: captured segments or a batched group).
        The driver compares its
        own step with the chain's (chain_matches, chain_step) before it uses chain_root and chain_trial; a chain it
        does not use is simply not read (it writes only the trial's buffers)."""
        if not self.chain or self.graphs is not None or self.batch is not None:
            return None
        nz, ny, nx = self.shape
        s = self._stream()
        c = self._chain_dev
        _lib.call("seqvp_pg_step", _lib.ptr(self.scalars), int(has_prev), float(prev_tau), float(step_min),
                  float(step_max), float(initial_step), _lib.ptr(c[0:1]), s)
        if root:
            w, n = self.q1._workspace(OP_ROOT)
            _lib.call("seqvp_jd_root_free_dev", _lib.ptr(self.u), _lib.ptr(self.p), _lib.ptr(self.q1.k),
                      _lib.ptr(c[0:1]), nx, ny, nz, _lib.ptr(c[1:4]), w, n, s)
            _lib.call("seqvp_capped_step", _lib.ptr(c[0:1]), _lib.ptr(c[1:4]), float(safety), _lib.ptr(c[4:5]), s)
        _lib.call("seqvp_copy_async", ctypes.c_void_p(self._chain_host.data_ptr()), _lib.ptr(c), 5 * 8, s)
        root_done = torch.cuda.Event()
        root_done.record()
        known = self._base_min_ptr() is not None
        self._trial_segment(True, known, self._quiet(), tau_dev=c[4:5] if root else c[0:1])
        return dict(root_done=root_done, root=root, u_t=self.u_t, theta=self.theta, p=self.p, version=self._version)

    def chain_matches(self, chain, tau):
        """Whether the chain's step (waiting for its root sweep) is the host's `tau`, bit for bit, and the chain
        still belongs to the current accepted state and direction."""
        chain["root_done"].synchronize()
        return (chain["version"] == self._version and chain["theta"] is self.theta and chain["p"] is self.p
                and chain["u_t"] is self.u_t and float(self._chain_np[0]) == tau)

    def chain_root(self, chain):
        """(root, invalid, points) of the chain's T66 sweep (a chain launched with root=True), as jd_bound_read gives
        them."""
        root, invalid, bad = (float(v) for v in self._chain_np[1:4])
        if bad:
            raise FloatingPointError("T66 nonfinite cubic coefficients (fused)")
        return root, int(invalid), 27 * self.q1.ncell

    def chain_step(self, chain):
        """The step the chained trial used."""
        return float(self._chain_np[4 if chain["root"] else 0])

    def chain_trial(self, chain):
        """The chained trial's result, as trial(chain_step(chain)) gives it."""
        self._certificate = None
        self.cert._u = self.u_t
        self.chained_trials += 1
        return self._trial_read()

    def _trial_result(self, n, s, b, nbad, summary, screen):
        """The TrialResult of a candidate from its node energies, soft and barrier terms, invalid-sample count and
        certificate screen values."""
        q = self.q1
        if nbad:
            return TrialResult(False, "INVALID_DOMAIN", None)
        folded = not b[1] > q.delta or not np.isfinite(b[1])
        if folded:
            # a closed27 J at or below delta: the first tile's nominated point, confirmed by its enclosure, is a
            # witness and rejects the candidate; unconfirmed, the whole-cell certificate decides
            # (witness_screen.py:55-88; accelerated.py:1692-1729)
            witness = nominate_witness(self.u_t, q.k, self.grid, q.delta, self.device)
            if witness is not None:
                return TrialResult(False, "JACOBIAN_VIOLATION", None, witness=witness)
        cert = self._certificate if self._certificate is not None else self.cert.finish(summary=summary)
        self._certificate = cert
        if not cert.certified:
            reason = "FEASIBILITY_UNCERTAIN" if cert.status == "uncertain" else cert.reason
            return TrialResult(False, reason, None, unproven_cells=cert.n_uncertain + cert.n_infeasible,
                               certificate=cert)
        if folded:                      # certified, yet a float closed27 J sits at delta: no finite barrier
            return TrialResult(False, "JACOBIAN_VIOLATION", None, certificate=cert)
        energy = Energies(n[0], n[1], s[0], s[1], b[0])
        change = q.barrier_weight * b[2]
        total = n[2] + n[3] + s[2] + change
        outer = n[2] + change
        gap = min(b[1], b[3]) - q.delta
        band = None
        if screen:
            base = self.energy
            soft_sum = max(energy.jd + energy.curl + base.jd + base.curl, 0.)
            bar_sum = max(energy.barrier + base.barrier, 0.)
            band = FP32_EPS * (FP32_BAND_SOFT * soft_sum + FP32_BAND_BARRIER * bar_sum)
        return TrialResult(True, "", energy, total, outer, gap, 0, band, cert)

    def promote(self, energy, keep_history=False):
        """Make the trial the accepted state (swap buffers; the trial's image gradient came with its warp). With
        `keep_history` the replaced state's (theta, g, p) are returned for the next BB2 products, kept by moving
        buffers rather than copying them: g and p change places with history slots (the next gradient and
        direction write the freed ones), and the replaced theta leaves theta_t for a slot before the next trial."""
        if keep_history:
            if self._prev is None:
                self._prev = dict(theta=torch.empty_like(self.theta), g=torch.empty_like(self.g),
                                  p=torch.empty_like(self.p))
            self.g, self._prev["g"] = self._prev["g"], self.g
            self.p, self._prev["p"] = self._prev["p"], self.p
        self._version += 1
        self.u, self.u_t = self.u_t, self.u
        self.theta, self.theta_t = self.theta_t, self.theta
        self.warped, self.warped_t = self.warped_t, self.warped
        self.bgrad, self.bgrad_t = self.bgrad_t, self.bgrad
        self.energy = energy
        self.certificate = self._certificate
        self._set_base_min(self._trial_min)
        if not keep_history:
            return None
        self.theta_t, self._prev["theta"] = self._prev["theta"], self.theta_t
        return dict(self._prev)

    # ---- outer-loop support ----------------------------------------------------------------------
    def snapshot(self):
        """Device copies of the accepted state (the outer start)."""
        return dict(theta=self.theta.clone(), u=self.u.clone(), warped=self.warped.clone(), bgrad=self.bgrad.clone(),
                    energy=self.energy, certificate=self.certificate, version=self._version,
                    base_min=self._base_min)

    def restore(self, snap):
        self._version += 1
        self.theta.copy_(snap["theta"])
        self.u.copy_(snap["u"])
        self.warped.copy_(snap["warped"])
        self.bgrad.copy_(snap["bgrad"])
        self.energy = snap["energy"]
        self.certificate = snap.get("certificate")
        self._set_base_min(snap.get("base_min"))

    def compare_to(self, snap):
        """(E(current) - E(snap), F(current) - F(snap), min gap): stable differences of the current state
        against a snapshot with the same reference (sequential.py:549-555; accelerated.py:2965-3060)."""
        q, w = self.q1, self.problem.weights
        node = q.node_energy(self.warped, self.fixed, self.u, self.ref, w.lam, data_weight=self.data_weight,
                             base_warped=snap["warped"], base_u=snap["u"])
        # a snapshot of the state the reference was set to (the outer start after a refresh) is the reference
        # itself: passed as such, the sweep takes the base's point terms from the reference's
        same = snap.get("version") is not None and snap["version"] == self._ref_at
        sb = q.trial_energy(self.u, self.ref, self.ref if same else snap["u"],
                            base_min=self._ref_min_dev if same and self._ref_min is not None else None).tolist()
        n, s_, b = node.tolist(), sb[0:3], sb[3:7]
        change = q.barrier_weight * b[2]
        return n[2] + n[3] + s_[2] + change, n[2] + change, min(b[1], b[3]) - q.delta

    def outer_norms(self):
        """(GF, GD, GE) = mass norms of the outer (image + barrier), image-only and total gradients
        (sequential.py:288-338; linearization.py:217-238)."""
        q, w = self.q1, self.problem.weights
        # p, diag and g are free here (recomputed before use)
        gd, gf, ge = q.joint_free(self.warped, self.fixed, self.bgrad, self.u, self.ref, w.lam,
                                  data_weight=self.data_weight, out=(self.p, self.diag, self.g))
        self._gf_at = self._version
        ff, dd, ee = self.dots((gf, gf), (gd, gd), (ge, ge))
        r = float(np.sqrt(self.grid.omega))
        return float(np.sqrt(ff)) / r, float(np.sqrt(dd)) / r, float(np.sqrt(ee)) / r

    def refresh_reference(self):
        """Self reference: u_ref <- u (problem.py:397-403). The proximity, JD and curl change terms of the
        state are then exactly zero; image and barrier do not depend on the reference."""
        self.ref.copy_(self.u)
        self._ref_at = self._version
        self._ref_min = self._base_min
        if self._ref_min is not None:
            self._ref_min_dev.copy_(self._base_min_ptr())
        if self.energy is None:
            self.energy = self.evaluate_energy(self.u, self.warped)
        self.energy = Energies(self.energy.image, 0., 0., 0., self.energy.barrier)

    def gradient_norm(self):
        (gg,) = self.dots((self.g, self.g))
        return float(np.sqrt(gg)) / float(np.sqrt(self.grid.omega))

    def bb2_products(self, prev):
        """(s.y, |s|, |y|, y.iy) for the previous accepted (theta, g, p)."""
        out = self.dot_out
        _lib.call("seqvp_bb2_products", _lib.ptr(self.theta), _lib.ptr(prev["theta"]), _lib.ptr(self.g),
                  _lib.ptr(prev["g"]), _lib.ptr(prev["p"]), _lib.ptr(self.p), self.nfree, _lib.ptr(out),
                  _lib.ptr(self.dot_ws), self.dot_ws.numel(), self._stream())
        sy, ss, yy2, yiy = out.tolist()
        return sy, float(np.sqrt(ss)), float(np.sqrt(yy2)), yiy

    def direction_scalars(self, prev, norm=False, ratio=False, fp32=None, gradient=False):
        """(g.p, BB2 products for `prev` or None, domain step bound, gradient norm or None, ratio of the last
        accepted step or None) with one host synchronisation: the reductions are enqueued back to back into
        separate slots of one small buffer. `ratio` reads the slot step_ratio_launch filled. With `fp32` not None
        the direction p = -S M^-1 (S g) (FP32 transforms when true) is computed first, and with `gradient` the
        gradient before it (gradient()); all of it runs as one segment."""
        return self.direction_scalars_read(self.direction_scalars_launch(prev, norm, ratio, fp32, gradient))

    def direction_scalars_launch(self, prev, norm=False, ratio=False, fp32=None, gradient=False):
        """Enqueues direction_scalars' segment and returns at once; direction_scalars_read(pending) synchronises
        and returns its values. The host may do other work in between, as long as it changes no state."""
        sweep = gradient and not self._take_outer_gradient()
        prev_key = None if prev is None else (prev["theta"], prev["g"], prev["p"])
        # the pending step ratio comes with the one-pass scalars when its pair is (theta, the history's theta), the
        # pair promote leaves; any other pair gets its own pass inside the segment
        pair = self._ratio_pair
        extra = None if pair is None or (prev is not None and pair[0] is self.theta and pair[1] is prev["theta"]) \
            else pair
        self._ratio_pair = None
        key = None if self.graphs is None else key_of(
            "direction", sweep, fp32, *(prev_key or (None,)), *(extra or (None,)), self.g, self.p, self.u, self.theta,
            self.warped, self.bgrad, self.ref, self.fixed, self.data_weight, self.scale, self._den, self.q1.ws.buf,
            self.dot_ws, self.scalars)
        self._segment(key, lambda: self._direction_segment(prev, fp32, sweep, extra))
        done = torch.cuda.Event()
        done.record()
        return prev is not None, norm, ratio, done

    def direction_scalars_read(self, pending):
        """The values of a direction_scalars_launch (waits for its segment only, not for work enqueued after it)."""
        has_prev, norm, ratio, done = pending
        done.synchronize()
        v = self._scalars_np.tolist()
        bb2 = (v[4], float(np.sqrt(v[5])), float(np.sqrt(v[6])), v[7]) if has_prev else None
        G = float(np.sqrt(v[1])) / float(np.sqrt(self.grid.omega)) if norm else None
        return v[0], bb2, v[8], G, (self._ratio(v[9], v[10]) if ratio else None)

    def _direction_segment(self, prev, fp32, sweep, extra):
        if sweep:
            self._gradient_sweep()
        if fp32 is not None:
            self.direction(fp32=fp32)
        # g.p, g.g, the BB2 products with the previous accepted state, the step bound and the last step's maxima
        # in one pass (each value the bits of its separate entry point); the step ratio's maxima are those of
        # (theta, prev theta), the pair step_ratio_launch recorded
        out = self.scalars
        nz, ny, nx = self.shape
        hx, hy, hz = self.grid.spacing
        ox, oy, oz = self.grid.origin
        pv = (None, None, None) if prev is None else (prev["theta"], prev["g"], prev["p"])
        _lib.call("seqvp_direction_scalars", _lib.ptr(self.theta), _lib.ptr(pv[0]), _lib.ptr(self.g), _lib.ptr(pv[1]),
                  _lib.ptr(pv[2]), _lib.ptr(self.p), _lib.ptr(self.u), ox, oy, oz, hx, hy, hz, nx, ny, nz,
                  _lib.ptr(out), _lib.ptr(self.dot_ws), self.dot_ws.numel(), self._stream())
        if extra is not None:
            self._launch_step_ratio(*extra)
        _lib.call("seqvp_copy_async", ctypes.c_void_p(self._scalars_host.data_ptr()), _lib.ptr(out), 12 * 8,
                  self._stream())

    def step_bound(self):
        nz, ny, nx = self.shape
        hx, hy, hz = self.grid.spacing
        ox, oy, oz = self.grid.origin
        _lib.call("seqvp_step_bound", _lib.ptr(self.p), _lib.ptr(self.u), ox, oy, oz, hx, hy, hz, nx, ny, nz,
                  _lib.ptr(self.dot_out), _lib.ptr(self.dot_ws), self.dot_ws.numel(), self._stream())
        return float(self.dot_out[0].item())

    def step_ratio(self, a, b):
        """max|a - b| / max(L, max|b|), L = |axis lengths| (solvers/common.py:63-73)."""
        _lib.call("seqvp_max_abs_diff", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(self.dot_out),
                  _lib.ptr(self.dot_ws), self.dot_ws.numel(), self._stream())
        top, bmax = self.dot_out[:2].tolist()
        self.step_max_diff = top
        return self._ratio(top, bmax)

    def _ratio(self, top, bmax):
        length = float(np.linalg.norm([ax[-1] - ax[0] for ax in self.grid.axes]))
        return float(np.float64(top) / max(length, bmax))

    def step_ratio_launch(self, a, b):
        """step_ratio(a, b), read by the next direction_scalars(ratio=True) or read_step_ratio. Called with the
        accepted trial and the state it replaces, right before promote: the pair is recorded, and the next
        direction's one-pass scalars compute its maxima (a and b are then theta and the history's theta); a
        read_step_ratio before that launches them on its own."""
        self._ratio_pair = (a, b)

    def _launch_step_ratio(self, a, b):
        _lib.call("seqvp_max_abs_diff", _lib.ptr(a), _lib.ptr(b), a.numel(), _lib.ptr(self.scalars[9:11]),
                  _lib.ptr(self.dot_ws), self.dot_ws.numel(), self._stream())

    def read_step_ratio(self):
        if self._ratio_pair is not None:
            self._launch_step_ratio(*self._ratio_pair)
            self._ratio_pair = None
        top, bmax = self.scalars[9:11].tolist()
        return self._ratio(top, bmax)


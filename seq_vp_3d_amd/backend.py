"""The reference's duck-typed `backend=` object over the device path.

AmdBackend has the method set that solve_sequential(..., backend=) and solve_poisson_gradient call on the
reference's AcceleratedHybridBackend (jd_curl/three_d/sequential.py:165-689, solvers/poisson_gradient.py:387-1646,
solvers/common.py:50-61; backends/accelerated.py:363, 917-964, 1451, 1666-1730, 2730, 2850-2938, 3113-3173;
backends/hybrid.py:51-286), with the same arguments, return shapes and error conventions:

  evaluate_trial, promote_trial, linearize, gradient, final_gradient, poisson (apply / prepare_equilibrated /
  apply_equilibrated), diagonal, diagonal_launch(...).result(), node_min_jd, global_jd_root,
  stable_energy_reduction(s), outer_gradient_vectors, outer_norms, exact_energy, cpu_state, cpu_outer_recheck,
  refresh_reference, device (upload / download / scalar / synchronize), evaluation_backend_id, backend_id,
  last_comparison.

Every optional feature flag of the reference backend (fp32_screen, lean_checks, boundary_reuse, metric_prestart,
...) is absent, so the reference's driver takes its plain branches. The arithmetic runs in libseqvp_hip.so: a trial
is the device domain check, closed27 witness screen, whole-cell certificate (GPU edge hull + prescreen, CPU
interval fallback), spline value and gradient and energies; comparisons are device double-double differences with
the resolution 16 eps max(E) (the reference's CPU rule, energy_comparison.py:117-124).

With a reference Problem3D the results are the reference's own classes (TrialState3D, AcceptedState3D,
InvalidTrial, EnergyTerms3D, EnergyReduction, PrescreenFeasibilityResult), so its validate_state checks pass. With
this package's `make_problem` they are the mirror classes below, which carry the same fields.
"""
import hashlib
from dataclasses import dataclass, field

import numpy as np
import torch

from . import _lib
from .certificate import CertificateConfig, DeviceCertificate
from .jd_step import nominate_witness
from .poisson import DevicePoisson
from .problem import InputValidationError, Q1Problem, Weights3D
from .q1 import FusedJDRoot, FusedQ1
from .spline import DeviceSpline, SplineCoefficients, build_image_sampler

EPS = float(np.finfo(float).eps)
BACKEND_ID = "seq_vp_3d_amd-hip-gfx950-float64-certified"


# ---- mirror classes (the reference's field names, residuals.py:23-55, types.py:48-53, energy_comparison.py:10-16)
@dataclass(frozen=True)
class EnergyTerms3D:
    image: float
    proximity: float
    jd: float
    curl: float
    barrier: float
    curvature: float = 0.

    @property
    def total(self):
        return self.image + self.proximity + self.jd + self.curl + self.barrier + self.curvature

    @property
    def outer(self):
        return self.image + self.barrier + self.curvature


@dataclass(frozen=True)
class TrialState3D:
    theta: np.ndarray
    u: np.ndarray
    warped: np.ndarray
    energy: EnergyTerms3D
    certificate: object
    problem_id: str
    eval_id: str


@dataclass(frozen=True)
class AcceptedState3D(TrialState3D):
    image_grad: np.ndarray


@dataclass(frozen=True)
class InvalidTrial:
    reason: str
    message: str
    invalid_count: int


@dataclass(frozen=True)
class EnergyReduction:
    actual: float
    resolution: float

    @property
    def certified(self):
        return self.actual - self.resolution


@dataclass(frozen=True)
class Linearization3D:
    problem: object
    state: object


def build_linearization(problem, state):
    return Linearization3D(problem, state)


@dataclass(frozen=True)
class Controls:
    u_ref: np.ndarray
    controls_id: str


@dataclass(frozen=True)
class Problem:
    """The fields of the reference's Problem3D (problem.py:283-311) that the hot path reads."""
    grid: object
    fixed_image: np.ndarray
    sampler: SplineCoefficients
    controls: Controls
    weights: Weights3D
    certificate_config: CertificateConfig
    problem_id: str
    barrier_activation: float = .05
    barrier_location: str = "closed27_uniform"
    data_weight: np.ndarray | None = None
    curvature_weight: float = 0.


def _digest(*parts):
    h = hashlib.sha256()
    for p in parts:
        h.update(p if isinstance(p, bytes) else str(p).encode())
    return h.hexdigest()


def make_problem(*, grid, fixed_image, moving_image, u_ref, weights, certificate_config=None, barrier_activation=.05):
    """make_problem (problem.py:333-378) for the closed27 localized-barrier route: the moving image's not-a-knot
    spline coefficients (SciPy, as build_image_sampler), the frozen reference displacement and an identity."""
    fixed = np.array(fixed_image, dtype=np.float64, copy=True)
    u_ref = np.array(u_ref, dtype=np.float64, copy=True)
    if fixed.shape != tuple(grid.shape) or u_ref.shape != tuple(grid.shape) + (3,):
        raise InputValidationError("fixed image / u_ref do not match the grid")
    sampler = build_image_sampler(moving_image, grid)
    config = CertificateConfig(delta=0.) if certificate_config is None else certificate_config
    controls = Controls(u_ref, _digest(grid.shape, grid.spacing, u_ref.tobytes()))
    key = _digest(grid.shape, grid.spacing, fixed.tobytes(), sampler.coefficients.tobytes(), controls.controls_id,
                  weights, config, barrier_activation)
    return Problem(grid, fixed, sampler, controls, weights, config, key, float(barrier_activation))


def with_reference(problem, u_ref):
    """The same problem with another frozen reference displacement (make_self_reference + with_controls,
    problem.py:381-403): a new controls id and problem id."""
    u_ref = np.array(u_ref, dtype=np.float64, copy=True)
    controls = Controls(u_ref, _digest(problem.grid.shape, problem.grid.spacing, u_ref.tobytes()))
    key = _digest(problem.problem_id, controls.controls_id)
    return Problem(problem.grid, problem.fixed_image, problem.sampler, controls, problem.weights,
                   problem.certificate_config, key, problem.barrier_activation, problem.barrier_location,
                   problem.data_weight, problem.curvature_weight)


def _reference_types(problem):
    """The result classes for `problem`: the reference's own when it is a jd_curl problem (imported lazily, inside the
    reference's process), else this module's mirrors."""
    if type(problem).__module__.startswith("jd_curl"):
        from jd_curl.energy_comparison import EnergyReduction as ER
        from jd_curl.three_d.residuals import AcceptedState3D as AS, EnergyTerms3D as ET, TrialState3D as TS
        from jd_curl.types import InvalidTrial as IT
        return dict(energy=ET, trial=TS, accepted=AS, invalid=IT, reduction=ER, reference=True)
    return dict(energy=EnergyTerms3D, trial=TrialState3D, accepted=AcceptedState3D, invalid=InvalidTrial,
                reduction=EnergyReduction, reference=False)


class DeviceIO:
    """The reference's Device (torch_ops.py:42-95): upload / download / scalar / synchronize on one GPU."""

    def __init__(self, device):
        self.device = torch.device(device)

    def upload(self, a):
        # a copy: the reference freezes its host arrays, and a read-only array cannot back a tensor
        return torch.tensor(np.asarray(a, dtype=np.float64), dtype=torch.float64, device=self.device)

    def download(self, t):
        out = t.detach().cpu().numpy().copy()
        out.flags.writeable = False
        return out

    def scalar(self, t):
        return float(t.item())

    def synchronize(self):
        torch.cuda.synchronize(self.device)


class HostPoisson:
    """backend.poisson(inverse): the reference _HostPoisson protocol on host arrays (hybrid.py:224-286)."""

    def __init__(self, device, inverse):
        self.inverse = DevicePoisson(inverse, device.device)
        self.io = device

    def apply(self, rhs):
        return self.io.download(self.inverse.apply(self.io.upload(rhs).reshape(-1)))

    def prepare_equilibrated(self, scale, gradient):
        return self.inverse.prepare_equilibrated(scale, gradient)

    def apply_equilibrated(self, prepared, between=None):
        return self.inverse.apply_equilibrated(prepared, between=between)


class _PendingDiagonal:
    def __init__(self, io, tensor):
        self.io, self.tensor = io, tensor

    def result(self):
        t, self.tensor = self.tensor, None
        host = self.io.download(t)
        if not np.all(np.isfinite(host)) or np.any(host <= 0):
            raise FloatingPointError("invalid device GGN diagonal")
        return host


@dataclass
class _Model:
    """The device side of one problem family (grid, images, weights, delta, activation): kernels, spline, certificate."""
    q1: FusedQ1
    root: FusedJDRoot
    spline: DeviceSpline
    fixed: torch.Tensor
    data_weight: torch.Tensor | None
    cert: DeviceCertificate
    lam: float
    refs: dict = field(default_factory=dict)


class _DeviceState:
    def __init__(self, theta, u, warped, grad):
        self.theta, self.u, self.warped, self.grad = theta, u, warped, grad


class AmdBackend:
    """backend= for solve_sequential / solve_poisson_gradient on one MI355X."""

    evaluation_backend_id = BACKEND_ID
    backend_id = BACKEND_ID

    def __init__(self, device="cuda", *, certificate_workers=8):
        self.device = DeviceIO(device)
        self._models = {}
        self._states = {}           # eval_id -> _DeviceState (the most recent states only)
        self._workers = int(certificate_workers)
        self.last_comparison = dict(policy="device-double-double-16eps-v1")

    # ---- problem and state bookkeeping ---------------------------------------------------------------
    def _model(self, problem):
        grid = problem.grid
        sampler = problem.sampler
        key = (tuple(grid.shape), tuple(float(h) for h in grid.spacing), id(sampler), id(problem.fixed_image),
               float(problem.weights.lam), float(problem.weights.mu), float(problem.weights.nu),
               float(problem.weights.kappa), float(problem.certificate_config.delta), problem.barrier_activation)
        model = self._models.get(key)
        if model is None:
            if problem.barrier_location != "closed27_uniform" or problem.barrier_activation is None:
                raise InputValidationError("the device path needs the closed27 localized barrier")
            if getattr(problem, "curvature_weight", 0.):
                raise InputValidationError("the device path has no curvature term (alpha = 0 on the native level)")
            dev = self.device.device
            q1problem = Q1Problem(grid, problem.weights, float(problem.certificate_config.delta),
                                  float(problem.barrier_activation))
            if hasattr(sampler, "_spline"):          # the reference's ImageSampler3D: SciPy's own coefficients
                coefficients = SplineCoefficients(np.ascontiguousarray(sampler._spline.c, dtype=np.float64),
                                                  tuple(np.asarray(t, dtype=np.float64) for t in sampler._spline.t))
            else:
                coefficients = sampler
            q1 = FusedQ1(problem if hasattr(problem, "_gauss") else q1problem, dev)
            cfg = problem.certificate_config
            config = CertificateConfig(delta=float(cfg.delta), max_depth=int(getattr(cfg, "max_depth", 48)),
                                       max_boxes=int(getattr(cfg, "max_boxes", 255)),
                                       tile_size=int(getattr(cfg, "tile_size", 8192)), fallback_workers=self._workers)
            weight = getattr(problem, "data_weight", None)
            model = _Model(q1, FusedJDRoot(q1problem, q1=q1), DeviceSpline(coefficients, grid, dev),
                           self.device.upload(problem.fixed_image),
                           None if weight is None else self.device.upload(weight),
                           DeviceCertificate(grid, dev, config), float(problem.weights.lam))
            if len(self._models) > 8:
                self._models.clear()
            self._models[key] = model
        return model

    def _ref(self, model, problem):
        key = problem.controls.controls_id
        ref = model.refs.get(key)
        if ref is None:
            if len(model.refs) > 4:
                model.refs.clear()
            ref = model.refs[key] = self.device.upload(problem.controls.u_ref)
        return ref

    def _keep(self, eval_id, dstate):
        if len(self._states) > 6:
            self._states.pop(next(iter(self._states)))
        self._states[eval_id] = dstate

    def _device_state(self, problem, state):
        """The device arrays of an evaluated state (re-uploaded and re-sampled if not kept)."""
        d = self._states.get(state.eval_id)
        if d is None:
            model = self._model(problem)
            u = self.device.upload(state.u)
            warped, grad, bad = model.spline.warp(u)
            if int(bad.item()):
                raise FloatingPointError("state outside the sampling support")
            d = _DeviceState(self.device.upload(state.theta), u, warped, grad)
            self._keep(state.eval_id, d)
        return d

    def _certificate(self, types, problem, cert, u_host):
        if not types["reference"]:
            return cert
        from jd_curl.three_d.feasibility import PrescreenFeasibilityResult, certificate_version
        config = problem.certificate_config
        return PrescreenFeasibilityResult(
            cert.status, cert.min_jd_lower_bound, cert.min_margin_lower_bound, cert.sampled_min_jd,
            cert.n_cells_checked, cert.n_uncertain, cert.n_infeasible, cert.n_boxes, cert.max_depth, cert.witness_cell,
            cert.witness_reference_point, cert.reason, hashlib.sha256(memoryview(u_host).cast("B")).hexdigest(),
            problem.grid.grid_id, config.delta.hex(), config, certificate_version(config), cert.n_fast_cells,
            cert.n_fallback_cells, cert.n_prescreen_cells)

    # ---- evaluation -------------------------------------------------------------------------------------
    def evaluate_trial(self, problem, theta, *, certificate_observer=None, witness_observer=None):
        """TrialState3D or InvalidTrial for the free displacement theta (accelerated.py:1666-1730): finite theta, the
        sampling domain, the closed27 witness screen, the whole-cell certificate, then the energies."""
        types = _reference_types(problem)
        invalid = types["invalid"]
        theta = np.asarray(theta, dtype=np.float64)
        if theta.shape != (3 * problem.grid.n_free_scalar,):
            raise InputValidationError("theta must be the (3F,) free displacement")
        if not np.all(np.isfinite(theta)):
            return invalid("NONFINITE_STATE", "nonfinite free displacement", int(np.sum(~np.isfinite(theta))))
        model = self._model(problem)
        q = model.q1
        nz, ny, nx = problem.grid.shape
        T = self.device.upload(theta)
        u = torch.empty(tuple(problem.grid.shape) + (3,), dtype=torch.float64, device=self.device.device)
        _lib.call("seqvp_inject", _lib.ptr(T), nx, ny, nz, _lib.ptr(u), _lib.stream_handle(self.device.device))
        warped, grad, bad = model.spline.warp(u)
        if int(bad.item()):
            return invalid("INVALID_DOMAIN", "id+u violates closed boundary / open free-node support", int(bad.item()))
        bar = q.barrier_energy(u).tolist()
        if not bar[1] > q.delta:
            witness = nominate_witness(u, q.k, problem.grid, q.delta, self.device.device)
            if witness is not None:
                if witness_observer is not None:
                    witness_observer(*witness)
                return invalid("JACOBIAN_VIOLATION", "closed27 witness with J <= delta", 1)
        model.cert.screen(u)
        cert = model.cert.finish()
        u_host = self.device.download(u)
        record = self._certificate(types, problem, cert, u_host)
        if certificate_observer is not None:
            certificate_observer(record)
        if not cert.certified:
            reason = "FEASIBILITY_UNCERTAIN" if cert.status == "uncertain" else cert.reason
            return invalid(reason, "trial not certified: " + cert.reason, cert.n_uncertain + cert.n_infeasible)
        if not bar[1] > q.delta:
            return invalid("JACOBIAN_VIOLATION", "certified, yet a float closed27 J is at delta", 1)
        ref = self._ref(model, problem)
        node = q.node_energy(warped, model.fixed, u, ref, model.lam, data_weight=model.data_weight).tolist()
        soft = q.soft_energy(u, ref).tolist()
        energy = types["energy"](node[0], node[1], soft[0], soft[1], bar[0], 0.)
        eval_id = _digest(problem.problem_id, hashlib.sha256(memoryview(u_host).cast("B")).hexdigest())
        theta = theta.copy()
        theta.flags.writeable = False
        self._keep(eval_id, _DeviceState(T, u, warped, grad))
        return types["trial"](theta, u_host, self.device.download(warped), energy, record, problem.problem_id,
                              eval_id)

    def promote_trial(self, problem, trial):
        """AcceptedState3D with the image gradient of the trial's own warp (accelerated.py:2850-2874)."""
        types = _reference_types(problem)
        d = self._device_state(problem, trial)
        return types["accepted"](trial.theta, trial.u, trial.warped, trial.energy, trial.certificate,
                                 trial.problem_id, trial.eval_id, self.device.download(d.grad))

    def exact_energy(self, problem, state):
        """The float64 energy terms of a state (accelerated.py:2730): the device evaluation is float64 already."""
        return state.energy

    def refresh_reference(self, problem, state):
        """No reuse of a refreshed reference (the reference certifies the new controls itself)."""
        return None

    def cpu_state(self, problem, state):
        """The reference's own CPU evaluation and promotion of the state's theta (accelerated.py:2911-2925): the
        sequential driver's near-threshold recheck of the outer norms."""
        from jd_curl.three_d.residuals import evaluate_trial, promote_trial
        trial = evaluate_trial(problem, state.theta)
        return trial if isinstance(trial, _reference_types(problem)["invalid"]) else promote_trial(problem, trial)

    def cpu_outer_recheck(self, problem, state):
        """The three outer gradient vectors of the CPU re-evaluation (accelerated.py:2927-2938)."""
        from jd_curl.three_d.linearization import build_linearization as build, outer_gradient_vectors
        cpu = self.cpu_state(problem, state)
        if not hasattr(cpu, "image_grad"):
            raise FloatingPointError("CPU outer gradient recheck state invalid: " + cpu.reason)
        return outer_gradient_vectors(build(problem, cpu))

    # ---- derivatives -----------------------------------------------------------------------------------
    def linearize(self, lin):
        """The device arrays of the linearization's state."""
        return self._device_state(lin.problem, lin.state)

    def gradient(self, lin, target):
        """The energy gradient on the free nodes as a host array (hybrid.py:70-77)."""
        model, d = self._model(lin.problem), self.linearize(lin)
        ref = self._ref(model, lin.problem)
        g = model.q1.gradient_free(d.warped, model.fixed, d.grad, d.u, ref, model.lam, data_weight=model.data_weight)
        out = self.device.download(g)
        if not np.all(np.isfinite(out)):
            raise FloatingPointError("device gradient")
        return out

    final_gradient = gradient

    def outer_gradient_vectors(self, lin):
        """(outer F, image D, total E) gradients on the free nodes (accelerated.py:3125-3160)."""
        model, d = self._model(lin.problem), self.linearize(lin)
        ref = self._ref(model, lin.problem)
        image, outer, total = model.q1.joint_free(d.warped, model.fixed, d.grad, d.u, ref, model.lam,
                                                  data_weight=model.data_weight)
        result = tuple(self.device.download(v) for v in (outer, image, total))
        for v in result:
            if not np.all(np.isfinite(v)):
                raise FloatingPointError("device outer gradient")
        return result

    def outer_norms(self, lin, omega):
        return self.outer_vector_norms(self.outer_gradient_vectors(lin), omega)

    @staticmethod
    def outer_vector_norms(vectors, omega):
        """gradient_norm_mass of each vector (solvers/common.py:182-191)."""
        return tuple(float(np.linalg.norm(v)) / float(np.sqrt(omega)) for v in vectors)

    def poisson(self, inverse):
        return HostPoisson(self.device, inverse)

    def diagonal(self, lin):
        """The GGN diagonal on the free nodes as a device vector (hybrid.py:80-86)."""
        model, d = self._model(lin.problem), self.linearize(lin)
        ref = self._ref(model, lin.problem)
        q = model.q1
        base = torch.empty_like(d.u)
        q.node_terms(d.warped, model.fixed, d.grad, d.u, ref, model.lam, data_weight=model.data_weight, ggn=base)
        nodes = q.diagonal_nodes(base, d.u)
        out = torch.empty(3 * lin.problem.grid.n_free_scalar, dtype=torch.float64, device=self.device.device)
        nz, ny, nx = lin.problem.grid.shape
        _lib.call("seqvp_restrict", _lib.ptr(nodes), nx, ny, nz, _lib.ptr(out), _lib.stream_handle(self.device.device))
        return out

    def diagonal_launch(self, lin):
        return _PendingDiagonal(self.device, self.diagonal(lin))

    def node_min_jd(self, lin, *, extension_activation=None, stats=None, compiled=False):
        """The T62 node JD measure (host array of the grid shape), with the T69 extension when
        extension_activation is given; stats accumulates candidates / lowered_cells / lowered_cells_any / mu_min as
        jd_affine._count does (jd_affine.py:205-225, hybrid.py:97-101)."""
        model, d = self._model(lin.problem), self.linearize(lin)
        if extension_activation is None:
            return self.device.download(model.q1.node_min_jd(d.u))
        st = torch.zeros(4, dtype=torch.float64, device=self.device.device)
        node = model.q1.node_min_jd(d.u, extension_activation=float(extension_activation), stats=st)
        cand, lowered, lowered_any, mu_min = st.tolist()
        if stats is not None and cand:
            stats["candidates"] = stats.get("candidates", 0) + int(cand)
            stats["lowered_cells"] = stats.get("lowered_cells", 0) + int(lowered)
            stats["lowered_cells_any"] = stats.get("lowered_cells_any", 0) + int(lowered_any)
            stats["mu_min"] = min(stats.get("mu_min", float("inf")), float(mu_min))
        return self.device.download(node)

    def global_jd_root(self, lin, direction, limit):
        """(first root in s-units or inf, base points with c0 <= 0, points) of the T66 bound (hybrid.py:92-95)."""
        model, d = self._model(lin.problem), self.linearize(lin)
        return model.root.from_free(d.u, self.device.upload(direction), float(limit))

    # ---- comparisons ------------------------------------------------------------------------------------
    def _differences(self, problem, base, trial):
        model = self._model(problem)
        a, b = self._device_state(problem, base), self._device_state(problem, trial)
        ref = self._ref(model, problem)
        q = model.q1
        node = q.node_energy(b.warped, model.fixed, b.u, ref, model.lam, data_weight=model.data_weight,
                             base_warped=a.warped, base_u=a.u)
        vals = torch.cat((node, q.soft_energy(b.u, ref, a.u), q.barrier_energy(b.u, a.u))).tolist()
        n, s, bar = vals[0:4], vals[4:7], vals[7:11]
        if not min(bar[1], bar[3]) > q.delta:
            raise FloatingPointError("device comparison barrier point is not positive")
        change = q.barrier_weight * bar[2]
        return n[2] + n[3] + s[2] + change, n[2] + change

    def _reduction(self, problem, base, trial, difference, outer_only):
        f = (lambda e: e.outer) if outer_only else (lambda e: e.total)
        actual = -difference
        resolution = 16 * EPS * max(f(base.energy), f(trial.energy))
        if not (np.isfinite(actual) and np.isfinite(resolution)):
            raise FloatingPointError("stable energy reduction")
        return _reference_types(problem)["reduction"](float(actual), float(resolution))

    def stable_energy_reduction(self, problem, base, trial, *, outer_only=False, required_reduction=0.):
        dE, dF = self._differences(problem, base, trial)
        return self._reduction(problem, base, trial, dF if outer_only else dE, outer_only)

    def stable_energy_reductions(self, problem, base, trial, *, required_reduction=0.):
        dE, dF = self._differences(problem, base, trial)
        return (self._reduction(problem, base, trial, dE, False), self._reduction(problem, base, trial, dF, True))

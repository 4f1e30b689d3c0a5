"""FP32 8/4/2/1 pyramid pre-optimiser on the device (the reference's T88 pre-optimizer, preoptimize).

An uncertified initializer for the native FP64 solve (jd_curl/three_d/preopt.py:1-20, 518-614): on each of the
four levels of the pyramid (multiresolution.four_level_grids) the float32 pre-objective P(u) -- trilinear data
term, cell-centre JD/curl penalty, relu floor on J, seven-point curvature -- is minimised by Adam on the
Gaussian-smoothed gradient with a per-iteration move cap, then the field is carried to the next level by the
Q1 prolongation. The native handoff is tapered at the faces and certified (whole-cell J > 0 with the
alpha = .9^k back-off, multiresolution.prepare_transfer_problem) before the certified FP64 solver starts.

Everything runs in libseqvp_hip.so: the pyramid images (seqvp_gauss_axis + seqvp_lerp_axis, float64), the
energy/gradient (seqvp_preopt_energy_grad, float32 with Kahan/double-double sums) and the update
(seqvp_preopt_step: smoothing, Adam, cap and the iteration counter on the device). One level iteration needs no
host synchronisation, so the iterations after the first are replayed from a captured graph. Per level the record
states the FP32 path's energy and gradient error against the same kernels in float64 at the level's first and
last state.
"""
from dataclasses import asdict, dataclass, field
from time import perf_counter

import numpy as np
import torch

from . import _lib
from .problem import Grid3D, InputValidationError, make_grid_3d

FACTORS = (8, 4, 2, 1)


@dataclass(frozen=True)
class PreoptConfig:
    """PreoptConfig (preopt.py:27-60); native() = the native run's flags (docs/usage.md:91)."""
    iterations: tuple = (80, 80, 60, 40)
    lr: float = .25
    cap: float = .5
    sigma_grad: float = 1.
    sigma_field: float = 0.
    mu: tuple = (1., 1., .1, .03)
    alpha: tuple = (1e-3, 1e-3, 1e-3, 0.)
    kappa_pre: float = 10.
    j0: float = .2
    beta1: float = .9
    beta2: float = .99
    eps: float = 1e-8
    taper: int = 4

    def __post_init__(self):
        if len(self.iterations) != 4 or any(int(n) != n or n < 0 for n in self.iterations):
            raise InputValidationError("preopt iterations: four nonnegative integers")
        if len(self.mu) != 4 or len(self.alpha) != 4 or min(self.mu) <= 0 or min(self.alpha) < 0:
            raise InputValidationError("preopt mu (positive) and alpha (nonnegative): four values each")
        if not (self.lr > 0 and self.cap > 0 and self.sigma_grad >= 0 and self.sigma_field >= 0 and self.j0 < 1
                and 0 < self.beta1 < 1 and 0 < self.beta2 < 1 and self.eps > 0 and self.taper >= 1):
            raise InputValidationError("preopt: invalid step, smoothing or Adam parameters")

    @classmethod
    def native(cls):
        return cls(iterations=(50, 50, 100, 60), lr=.5, cap=.5, sigma_grad=1., sigma_field=1.,
                   mu=(1., 1., .01, .003), alpha=(1e-3, 1e-3, 1e-3, 0.), kappa_pre=10., j0=.2, taper=4)


def coarse_grid(grid, shape_zyx):
    """The same physical node-centre FOV on a coarser node shape (nifti.py:112-120)."""
    if any(n < 5 or n > old for n, old in zip(shape_zyx, grid.shape)):
        raise InputValidationError("target axes must be >= 5 and no larger than the source")
    spacing = tuple((axis[-1] - axis[0]) / (n - 1) for axis, n in zip(grid.axes, shape_zyx[::-1]))
    return make_grid_3d(nx=shape_zyx[2], ny=shape_zyx[1], nz=shape_zyx[0], hx=spacing[0], hy=spacing[1],
                        hz=spacing[2], x0=grid.axes[0][0], y0=grid.axes[1][0], z0=grid.axes[2][0])


def four_level_grids(native):
    """Shapes (n - 2 + s)//s + 1 for s = 8, 4, 2 and the native grid (multiresolution.py:45-52)."""
    shapes = [tuple((n - 2 + s) // s + 1 for n in native.shape) for s in FACTORS]
    if any(n < 5 for shape in shapes for n in shape):
        raise InputValidationError("all four levels need >= 5 nodes on every axis")
    return tuple(coarse_grid(native, shape) for shape in shapes[:-1]) + (native,)


def _stream(device):
    return _lib.stream_handle(device)


def _dev(a, device, dtype=torch.float64):
    return torch.as_tensor(np.ascontiguousarray(a), dtype=dtype, device=device)


def _lerp_tables(t, n_old, right_clamp):
    left = np.floor(t).astype(np.int64)
    if right_clamp:                      # map_coordinates order 1, mode "nearest": clamp to the last node
        left = np.clip(left, 0, n_old - 1)
        right = np.minimum(left + 1, n_old - 1)
    else:                                # prolongation: the last target node uses the last source cell
        left[-1] = n_old - 2
        right = left + 1
    return left, right, t - left


def downsample(values, source, shape_zyx, device=None):
    """Gaussian anti-aliasing (sigma = sqrt(r^2 - 1)/2 source voxels, truncate 4, nearest) then linear
    resampling at the coarse node positions (nifti.py:128-149), float64 on the device. Returns (values, grid)."""
    device = values.device if device is None else torch.device(device)
    grid = coarse_grid(source, tuple(shape_zyx))
    ratios = (np.asarray(source.shape) - 1) / (np.asarray(shape_zyx) - 1)
    sigma = .5 * np.sqrt(ratios ** 2 - 1)
    a = values.contiguous()
    dims = list(source.shape)
    for axis in range(3):
        s = float(sigma[axis])
        if s <= 1e-15:
            continue
        r = int(4. * s + .5)
        x = np.arange(-r, r + 1, dtype=np.float64)
        w = np.exp(-.5 / (s * s) * x ** 2)
        w = w / w.sum()
        wd = _dev(w, device)
        out = torch.empty_like(a)
        _lib.call("seqvp_gauss_axis", _lib.ptr(a), _lib.ptr(out), dims[0], dims[1], dims[2], 1, axis, _lib.ptr(wd), r,
                  _stream(device))
        a = out
    for axis in range(3):
        old, new = source.shape[axis], shape_zyx[axis]
        left, right, w = _lerp_tables(np.linspace(0, old - 1, new), old, True)
        tabs = [_dev(left, device, torch.int64), _dev(right, device, torch.int64), _dev(w, device)]
        od = list(dims)
        od[axis] = new
        out = torch.empty(od, dtype=torch.float64, device=device)
        _lib.call("seqvp_lerp_axis", _lib.ptr(a), _lib.ptr(out), dims[0], dims[1], dims[2], 1, axis, new,
                  *(_lib.ptr(t) for t in tabs), _stream(device))
        a, dims = out, od
    return a, grid


def prolong(u, source, target):
    """Physical Q1 interpolation of a (nz, ny, nx, 3) displacement to a finer grid of the same FOV
    (prolong_displacement, multiresolution.py:55-84), float64 on the device, bitwise the host arithmetic."""
    device = u.device
    a = u.contiguous()
    dims = list(source.shape)
    src_axes, tgt_axes = source.axes[::-1], target.axes[::-1]       # z, y, x
    for axis, (old, new, h) in enumerate(zip(src_axes, tgt_axes, source.spacing[::-1])):
        if len(new) < len(old) or abs(old[0] - new[0]) > 1e-9 * max(1., abs(old[0])) or \
                abs(old[-1] - new[-1]) > 1e-9 * max(1., abs(old[-1])):
            raise InputValidationError("prolongation must refine the same physical FOV")
        t = (new - old[0]) / h
        t[0], t[-1] = 0., len(old) - 1
        left, right, w = _lerp_tables(t, len(old), False)
        tabs = [_dev(left, device, torch.int64), _dev(right, device, torch.int64), _dev(w, device)]
        od = list(dims)
        od[axis] = len(new)
        out = torch.empty(od + [3], dtype=torch.float64, device=device)
        _lib.call("seqvp_lerp_axis", _lib.ptr(a), _lib.ptr(out), dims[0], dims[1], dims[2], 3, axis, len(new),
                  *(_lib.ptr(x) for x in tabs), _stream(device))
        a, dims = out, od
    return a


def taper(u, grid, L=4):
    """Nodes within L of a face scale linearly to zero on it (boundary_taper, preopt.py:617-624)."""
    rs = []
    for n in grid.shape:
        k = np.arange(n)
        rs.append(np.minimum(np.minimum(k, n - 1 - k) / float(L), 1.))
    rz, ry, rx = (_dev(r, u.device) for r in rs)
    out = torch.empty_like(u)
    nz, ny, nx = grid.shape
    _lib.call("seqvp_taper", _lib.ptr(u.contiguous()), nx, ny, nz, _lib.ptr(rx), _lib.ptr(ry), _lib.ptr(rz),
              _lib.ptr(out), _stream(u.device))
    return out


def _gauss32(sigma):
    """The smoothing kernel of preopt._gaussian_1d in float32: radius ceil(3 sigma), normalised."""
    r = int(np.ceil(3 * sigma))
    x = np.arange(-r, r + 1).astype(np.float32)
    k = np.exp(np.float32(-.5) * (x / np.float32(sigma)) ** 2).astype(np.float32)
    return (k / k.sum()).astype(np.float32), r


class LevelState:
    """Device buffers of one level: interior field, Adam moments, gradient, scratch, workspace, terms
    (light=True: the gradient, workspace and terms only, for an energy/gradient evaluation)."""

    def __init__(self, grid, device, fp32=True, light=False):
        self.grid = grid
        self.device = device
        self.fp32 = fp32
        self.dtype = torch.float32 if fp32 else torch.float64
        nz, ny, nx = grid.shape
        n = 3 * (nz - 2) * (ny - 2) * (nx - 2)
        z = lambda: torch.zeros(n, dtype=self.dtype, device=device)
        self.g = z()
        self.u = self.m = self.v = self.tmp = None
        if not light:
            self.u, self.m, self.v, self.tmp = z(), z(), z(), z()
        self.ws_bytes = int(_lib.load().seqvp_preopt_workspace_bytes(nx, ny, nz, int(fp32)))
        self.ws = torch.empty(self.ws_bytes, dtype=torch.uint8, device=device)
        self.terms = torch.zeros(6, dtype=torch.float64, device=device)
        self.state = torch.zeros(2, dtype=torch.int64, device=device)       # iteration t, cap hits
        self.pmax = torch.zeros(1024, dtype=self.dtype, device=device)


def energy_grad(st, u, F, M, mu, alpha, kappa, j0, grad=None, terms=None):
    """P(u) terms [total, data, geometry, floor, curvature, min J] (device) and dP/du (into grad)."""
    nz, ny, nx = st.grid.shape
    hx, hy, hz = st.grid.spacing
    grad = st.g if grad is None else grad
    terms = st.terms if terms is None else terms
    _lib.call("seqvp_preopt_energy_grad", _lib.ptr(u), _lib.ptr(F), _lib.ptr(M), nx, ny, nz, hx, hy, hz, float(mu),
              float(alpha), float(kappa), float(j0), int(st.fp32), _lib.ptr(grad), _lib.ptr(terms), _lib.ptr(st.ws),
              st.ws_bytes, _stream(st.device))
    return terms, grad


TERM_NAMES = ("total", "data", "geometry", "floor", "curvature", "min_det")


def capture(enqueue, device):
    """A graph of the launches `enqueue` makes (captured on a side stream; nothing runs during capture)."""
    timer, _lib.TIMER = _lib.TIMER, None          # no timing events inside a capture
    try:
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream(device=device)
        side.wait_stream(torch.cuda.current_stream(device))
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                enqueue()
        torch.cuda.current_stream(device).wait_stream(side)
    finally:
        _lib.TIMER = timer
    return g


class PreOptimizer:
    """preoptimize() on one device; `graph` replays a level's iterations from a captured graph."""

    def __init__(self, device="cuda", config=None, graph=True):
        self.device = torch.device(device)
        self.config = PreoptConfig() if config is None else config
        self.graph = graph
        c = self.config
        self._wg = None if c.sigma_grad <= 0 else _gauss32(c.sigma_grad)
        self._wf = None if c.sigma_field <= 0 else _gauss32(c.sigma_field)

    def _weights(self, fp32):
        dt = torch.float32 if fp32 else torch.float64
        out = []
        for w in (self._wg, self._wf):
            out.append((None, -1) if w is None else (torch.as_tensor(w[0], dtype=dt, device=self.device), w[1]))
        return out

    def _iteration(self, st, F, M, level, wts):
        c = self.config
        energy_grad(st, st.u, F, M, c.mu[level], c.alpha[level], c.kappa_pre, c.j0)
        (wg, rg), (wf, rf) = wts
        nz, ny, nx = st.grid.shape
        _lib.call("seqvp_preopt_step", _lib.ptr(st.u), _lib.ptr(st.m), _lib.ptr(st.v), _lib.ptr(st.g),
                  _lib.ptr(st.tmp), nx, ny, nz, _lib.ptr(wg), rg, _lib.ptr(wf), rf, c.lr, c.cap, c.beta1, c.beta2,
                  c.eps, int(st.fp32), _lib.ptr(st.state), _lib.ptr(st.pmax), _stream(st.device))

    def errors(self, st, F64, M64, F, M, level):
        """FP32 energy/gradient of the current interior field against the float64 kernels on the same field."""
        c = self.config
        u64 = st.u.to(torch.float64)
        d = LevelState(st.grid, self.device, fp32=False, light=True)
        t64, g64 = energy_grad(d, u64, F64, M64, c.mu[level], c.alpha[level], c.kappa_pre, c.j0)
        t32, g32 = energy_grad(st, st.u, F, M, c.mu[level], c.alpha[level], c.kappa_pre, c.j0, grad=st.tmp,
                               terms=torch.zeros(6, dtype=torch.float64, device=self.device))
        a, b = t32.tolist(), t64.tolist()
        g32d = g32.to(torch.float64)
        num = float(torch.linalg.vector_norm(g32d - g64))
        den = float(torch.linalg.vector_norm(g64))
        return dict(energy_fp32=a[0], energy_fp64=b[0], energy_rel_error=abs(a[0] - b[0]) / max(abs(b[0]), 1e-300),
                    grad_rel_l2_error=num / max(den, 1e-300), min_det_fp32=a[5], min_det_fp64=b[5])

    def run_level(self, level, grid, F64, M64, seed_field, report_errors=True):
        c = self.config
        tick = perf_counter()
        st = LevelState(grid, self.device, fp32=True)
        nz, ny, nx = grid.shape
        hx, hy, hz = grid.spacing
        _lib.call("seqvp_field_to_interior", _lib.ptr(seed_field.contiguous()), nx, ny, nz, hx, hy, hz, 1,
                  _lib.ptr(st.u), _stream(self.device))
        F, M = F64.to(torch.float32).contiguous(), M64.to(torch.float32).contiguous()
        wts = self._weights(True)
        rec = dict(level=level, shrink_factor=FACTORS[level], shape_zyx=list(grid.shape), spacing=list(grid.spacing))
        if report_errors:
            rec["start_errors"] = self.errors(st, F64, M64, F, M, level)
        n = c.iterations[level]
        first = last = None
        graphed = False
        if n:
            self._iteration(st, F, M, level, wts)
            first = dict(zip(TERM_NAMES, st.terms.tolist()))
            last = first
        if n > 1:
            if self.graph:
                g = capture(lambda: self._iteration(st, F, M, level, wts), self.device)
                for _ in range(n - 1):
                    g.replay()
                graphed = True
            else:
                for _ in range(n - 1):
                    self._iteration(st, F, M, level, wts)
            last = dict(zip(TERM_NAMES, st.terms.tolist()))
        t, hits = st.state.tolist()
        if t != n:
            raise RuntimeError(f"preopt level {level}: {t} iterations ran, {n} expected")
        if last is not None and not np.isfinite(last["total"]):
            raise FloatingPointError("preopt: nonfinite pre-objective")
        out = torch.empty(grid.shape + (3,), dtype=torch.float64, device=self.device)
        _lib.call("seqvp_interior_to_field", _lib.ptr(st.u), nx, ny, nz, hx, hy, hz, 1, _lib.ptr(out),
                  _stream(self.device))
        if not bool(torch.isfinite(out).all()):
            raise FloatingPointError("preopt: nonfinite field")
        if report_errors:
            rec["end_errors"] = self.errors(st, F64, M64, F, M, level)
        torch.cuda.synchronize(self.device)
        rec.update(iterations=n, cap_hits=int(hits), graph=graphed, first=first, last=last,
                   seconds=perf_counter() - tick)
        return out, rec

    def __call__(self, fixed, moving, grid, report_errors=True):
        """(u float64 (nz, ny, nx, 3) mm with zero faces on the device, record). fixed/moving: float64 device
        images on `grid`. The field is uncertified; taper() and the transfer certificate come next."""
        started = perf_counter()
        grids = four_level_grids(grid)
        levels = []
        u = None
        for level, lg in enumerate(grids):
            if level == 3:
                F64, M64 = fixed.contiguous(), moving.contiguous()
            else:
                F64, _ = downsample(fixed, grid, lg.shape)
                M64, _ = downsample(moving, grid, lg.shape)
            seed = (torch.zeros(lg.shape + (3,), dtype=torch.float64, device=self.device) if u is None
                    else prolong(u, grids[level - 1], lg))
            u, rec = self.run_level(level, lg, F64, M64, seed, report_errors)
            levels.append(rec)
        return u, dict(policy="fp32-preopt-device-v1", config=asdict(self.config), dtype="float32",
                       levels=levels, seconds=perf_counter() - started)


TRANSFER_ALPHAS = tuple(.9 ** k for k in range(7))


def certify_transfer(st, seed, alphas=TRANSFER_ALPHAS):
    """prepare_transfer_problem (multiresolution.py:111-195) on a PairState: the first alpha * seed that lies in the
    sampling domain and is certified J > 0 on every whole cell becomes the accepted state (each alpha relative to
    the original seed; the first attempt is the seed itself). Returns (alpha or None, attempts)."""
    from .certificate import CertificateConfig, DeviceCertificate
    cert = DeviceCertificate(st.grid, st.device, CertificateConfig(delta=0.))
    field = torch.empty_like(seed)
    attempts = []
    for index, a in enumerate(alphas):
        if index == 0:
            field.copy_(seed)
        else:
            _lib.call("seqvp_scale", _lib.ptr(seed), float(a), _lib.ptr(field), seed.numel(), _stream(st.device))
        st.restrict(field, st.theta_t)
        st.inject(st.theta_t, st.u_t)
        _, _, bad = st.spline.warp(st.u_t, with_gradient=False, value=st.warped_t)
        nbad = int(bad.item())
        c = None if nbad else cert(st.u_t)
        reason = "INVALID_DOMAIN" if nbad else c.reason
        ok = c is not None and c.certified
        attempts.append(dict(attempt_index=index, alpha=a, certificate_passed=ok, reason=reason,
                             min_jd_lower_bound=None if c is None else c.min_jd_lower_bound))
        if ok:
            st.set_state(theta=st.theta_t.clone(), certificate=c)
            return a, attempts
        if reason not in ("JACOBIAN_VIOLATION", "subdivision_budget", "INVALID_DOMAIN"):
            break
    return None, attempts


def preoptimized_solve(problem, fixed, moving, device="cuda", config=None, max_outer=None, observer=None,
                       sampler=None, report_errors=True, budget=None, on_state=None, inner_config=None):
    """The precision form: FP32 pyramid pre-optimisation, boundary taper, certified transfer, then the certified
    FP64 solve (solve_sequential) from the transferred state (the pre-optimizer route of
    scripts/run_stage_3d_e19_mri.py:930-960 and multiresolution.prepare_transfer_problem). The native solve keeps
    the zero-displacement baseline (multiresolution.py:496-502); `budget()` is its wall-budget predicate and
    `on_state(st)` is called once the PairState exists (before any outer row reaches `observer`); `inner_config` is
    the PGConfig of the certified solve (the native route's defaults when None).
    Returns (PairState, SequentialResult or None when no alpha certified, record)."""
    from .hotpath import PairState
    from .solver import PGConfig, SequentialConfig, solve_sequential
    device = torch.device(device)
    config = PreoptConfig.native() if config is None else config
    t0 = perf_counter()
    u, rec = PreOptimizer(device, config)(fixed, moving, problem.grid, report_errors)
    seed = taper(u, problem.grid, config.taper)
    torch.cuda.synchronize(device)
    t1 = perf_counter()
    st = PairState(problem, fixed, moving, device, sampler=sampler)
    if on_state is not None:
        on_state(st)
    alpha, attempts = certify_transfer(st, seed)
    torch.cuda.synchronize(device)
    t2 = perf_counter()
    record = dict(preopt=rec, preopt_seconds=t1 - t0, transfer=dict(alpha=alpha, attempts=attempts,
                                                                    seconds=t2 - t1))
    if alpha is None:
        return st, None, record
    record["handoff_energy"] = st.energy.total
    res = solve_sequential(st, SequentialConfig(), PGConfig() if inner_config is None else inner_config,
                           observer=observer, max_outer=max_outer, budget=budget, baseline_at_zero=True)
    record["solve_seconds"] = perf_counter() - t2
    return st, res, record

"""On-disk pair and run formats around the device solver.

In: a pair folder with `pair.npz` (`fixed`, `moving` float64 (nz, ny, nx); optional `initial_u` (nz, ny, nx, 3);
label volumes kept but unused) and `manifest.json` (`shape_zyx`, `spacing_xyz`, `origin_xyz`, `unit` mm,
`weights.lam`, `pair_sha256`) -- docs/usage.md:40-57, read as scripts/run_stage_3d_e.py:40-47 does, written with the
fields of scripts/prepare_stage_3d_l2r_val.py:45-62.

Out: a run folder with `final_native.npz` (`u`, `warped`), `final_native.json` (file record + certificate of the
final state), `summary.json`, `history.jsonl` (one outer row per line) and `commits.json` -- docs/usage.md:127-139,
written as scripts/run_stage_3d_e19_mri.py:677-680 (JSON lines), 1269-1299 (arrays), 1370-1395 (commits, history)
and 2037-2159 (final state, summary) do, with scripts/run_stage_3d_c.py:32-36's JSON conventions.

`recertify` is the independent re-certification of a stored displacement (docs/usage.md:150-165): the CPU
interval certificate with delta = 0 on the grid of the pair's manifest.
"""
import hashlib
import json
import os
import resource
from dataclasses import asdict
from pathlib import Path
from time import perf_counter

import numpy as np
import torch

from .certificate import CertificateConfig, FeasibilityResult, certify_map
from .problem import InputValidationError, Q1Problem, Weights3D, make_grid_3d

PAIR_FILE = "pair.npz"
MANIFEST_FILE = "manifest.json"
FINAL_FILE = "final_native.npz"
ROUTES = ("preopt", "native")


def file_sha256(path):
    """SHA-256 of a file read in 1 MiB chunks (jd_curl/three_d/nifti.py:22-27)."""
    digest = hashlib.sha256()
    with Path(path).open("rb") as stream:
        for chunk in iter(lambda: stream.read(1 << 20), b""):
            digest.update(chunk)
    return digest.hexdigest()


def _json_default(obj):
    return obj.item() if isinstance(obj, np.generic) else str(obj)


def dump(path, value):
    """JSON with indent 2, no NaN/Inf, NumPy scalars as Python numbers; written to a temporary file then renamed
    (scripts/run_stage_3d_c.py:32-36)."""
    path = Path(path)
    temp = path.with_suffix(path.suffix + ".tmp")
    temp.write_text(json.dumps(value, indent=2, allow_nan=False, default=_json_default) + "\n")
    temp.replace(path)


def append_jsonl(stream, row):
    """One JSON row per line, flushed (scripts/run_stage_3d_e19_mri.py:677-680)."""
    stream.write(json.dumps(row, allow_nan=False, default=_json_default) + "\n")
    stream.flush()


def save_npz(folder, name, **arrays):
    """np.savez into `<name>.partial.npz`, renamed to `name`; returns {path, sha256} of the stored file."""
    folder = Path(folder)
    temp = folder / (name + ".partial.npz")
    np.savez(temp, **arrays)
    temp.replace(folder / name)
    return dict(path=name, sha256=file_sha256(folder / name))


# ---- pairs -------------------------------------------------------------------------------------------------------

def _triple(manifest, key, positive):
    v = manifest.get(key)
    if not isinstance(v, (list, tuple)) or len(v) != 3:
        raise InputValidationError(f"manifest {key}: three values required")
    try:
        a = [float(x) for x in v]
    except (TypeError, ValueError):
        raise InputValidationError(f"manifest {key}: numbers required") from None
    if not all(np.isfinite(a)) or (positive and min(a) <= 0):
        raise InputValidationError(f"manifest {key}: finite{' positive' if positive else ''} values required")
    return a


def grid_from_manifest(manifest):
    """The fixed image's grid from a manifest (shape_zyx, spacing_xyz, origin_xyz; unit mm)."""
    shape = manifest.get("shape_zyx")
    if (not isinstance(shape, (list, tuple)) or len(shape) != 3
            or not all(isinstance(n, int) and not isinstance(n, bool) and n >= 2 for n in shape)):
        raise InputValidationError("manifest shape_zyx: three integers >= 2 required")
    if manifest.get("unit", "mm") != "mm":
        raise InputValidationError("manifest unit must be mm")
    hx, hy, hz = _triple(manifest, "spacing_xyz", True)
    x0, y0, z0 = _triple(manifest, "origin_xyz", False)
    nz, ny, nx = shape
    return make_grid_3d(nx=nx, ny=ny, nz=nz, hx=hx, hy=hy, hz=hz, x0=x0, y0=y0, z0=z0)


def read_manifest(folder):
    path = Path(folder) / MANIFEST_FILE
    try:
        manifest = json.loads(path.read_text())
    except FileNotFoundError:
        raise InputValidationError(f"no {MANIFEST_FILE} in {folder}") from None
    if not isinstance(manifest, dict):
        raise InputValidationError("manifest must be a JSON object")
    return manifest


def load_pair(folder, verify=True):
    """(arrays, grid, manifest) of a pair folder. With `verify`, pair.npz must hash to the manifest's pair_sha256.
    Images are checked for shape and finiteness; a missing initial_u is zeros. Nothing in the archive is
    unpickled (allow_pickle=False)."""
    folder = Path(folder)
    manifest = read_manifest(folder)
    grid = grid_from_manifest(manifest)
    path = folder / PAIR_FILE
    if not path.is_file():
        raise InputValidationError(f"no {PAIR_FILE} in {folder}")
    if verify:
        if "pair_sha256" not in manifest:
            raise InputValidationError("manifest has no pair_sha256")
        if file_sha256(path) != manifest["pair_sha256"]:
            raise InputValidationError(f"{PAIR_FILE} does not match the manifest's pair_sha256")
    with np.load(path, allow_pickle=False) as archive:
        arrays = {k: archive[k] for k in archive.files}
    shape = tuple(grid.shape)
    for key in ("fixed", "moving"):
        if key not in arrays:
            raise InputValidationError(f"{PAIR_FILE} has no {key!r} image")
        a = np.ascontiguousarray(arrays[key], dtype=np.float64)
        if a.shape != shape:
            raise InputValidationError(f"{key} has shape {a.shape}, the manifest says {shape}")
        if not np.isfinite(a).all():
            raise InputValidationError(f"{key} has non-finite values")
        arrays[key] = a
    if "initial_u" in arrays:
        u0 = np.ascontiguousarray(arrays["initial_u"], dtype=np.float64)
        if u0.shape != shape + (3,) or not np.isfinite(u0).all():
            raise InputValidationError("initial_u must be finite with shape (nz, ny, nx, 3)")
        arrays["initial_u"] = u0
    else:
        arrays["initial_u"] = np.zeros(shape + (3,))
    if "weights" in manifest:
        lam = manifest["weights"].get("lam") if isinstance(manifest["weights"], dict) else None
        if not isinstance(lam, (int, float)) or not np.isfinite(lam) or lam <= 0:
            raise InputValidationError("manifest weights.lam must be a positive number")
    return arrays, grid, manifest


def baseline_mse(fixed, moving, grid):
    """sum q (M - F)^2 / sum q in long double: the zero-displacement MSE (prepare_stage_3d_l2r_val.py:53-54)."""
    q = grid.node_weights()
    d = np.asarray(moving, dtype=np.float64) - np.asarray(fixed, dtype=np.float64)
    return float(np.sum(q * d ** 2, dtype=np.longdouble) / np.sum(q, dtype=np.longdouble))


def default_lam(grid):
    """lam = 0.1 / L^2, L the geometric mean of the axis extents (prepare_stage_3d_l2r_val.py:55, 62)."""
    L = float(np.prod([a[-1] - a[0] for a in grid.axes])) ** (1 / 3)
    return .1 / L ** 2


def write_pair(folder, fixed, moving, *, spacing_xyz=(1., 1., 1.), origin_xyz=(0., 0., 0.), pair=None, lam=None,
               kappa=.002, initial_u=None, extra=None, compressed=True):
    """Writes a pair folder (which must not exist) in the format load_pair reads; returns the manifest."""
    folder = Path(folder)
    fixed = np.ascontiguousarray(fixed, dtype=np.float64)
    moving = np.ascontiguousarray(moving, dtype=np.float64)
    if fixed.ndim != 3 or moving.shape != fixed.shape:
        raise InputValidationError("fixed and moving must be 3-D arrays of one shape")
    nz, ny, nx = fixed.shape
    manifest = dict(shape_zyx=[nz, ny, nx], spacing_xyz=[float(h) for h in spacing_xyz],
                    origin_xyz=[float(o) for o in origin_xyz], unit="mm")
    grid = grid_from_manifest(manifest)
    u0 = np.zeros(fixed.shape + (3,)) if initial_u is None else np.ascontiguousarray(initial_u, dtype=np.float64)
    folder.mkdir(parents=True, exist_ok=False)
    (np.savez_compressed if compressed else np.savez)(folder / PAIR_FILE, fixed=fixed, moving=moving, initial_u=u0)
    manifest.update(pair=folder.name if pair is None else pair, pair_sha256=file_sha256(folder / PAIR_FILE),
                    baseline_native_mse=baseline_mse(fixed, moving, grid),
                    weights=dict(lam=default_lam(grid) if lam is None else float(lam), kappa=float(kappa)),
                    **({} if extra is None else extra))
    dump(folder / MANIFEST_FILE, manifest)
    return manifest


# ---- runs --------------------------------------------------------------------------------------------------------

def _certificate_dict(cert):
    if cert is None:
        return None
    return asdict(cert) if isinstance(cert, FeasibilityResult) else dict(cert)


class RunWriter:
    """Writes a run folder as the solver goes: history.jsonl per outer row, commits.json per formal commit
    (with the committed u and the start's u in `committed_NNNN_L3.npz` when `save_commits`), then the final state
    and summary."""

    def __init__(self, out, save_commits=False, started=None):
        self.out = Path(out)
        self.out.mkdir(parents=True, exist_ok=False)
        self.started = perf_counter() if started is None else started
        self.save_commits = save_commits
        self.history_rows = 0
        self.commits = []
        self.io_seconds = 0.
        self._history = (self.out / "history.jsonl").open("x")

    def observer(self, st, level=3):
        """The solver's outer observer: called with each outer row while `st` holds the candidate."""
        def observe(row):
            append_jsonl(self._history, dict(level=level, **row))
            self.history_rows += 1
            if row["outer_accepted"]:
                tick = perf_counter()
                if self.save_commits:
                    record = save_npz(self.out, f"committed_{len(self.commits) + 1:04d}_L{level}.npz",
                                      u=st.u.cpu().numpy(), warped=st.warped.cpu().numpy())
                else:
                    record = dict(path=None, sha256=None, arrays_saved=False)
                record.update(level=level, native=level == 3, role="formal_outer_commit",
                              ratio=row["committed_mse_ratio"], GF=row["committed_GF"],
                              certificate=_certificate_dict(st.certificate),
                              solver_commit_seconds=row["elapsed_seconds"],
                              visible_seconds=perf_counter() - self.started)
                self.commits.append(record)
                dump(self.out / "commits.json", self.commits)
                self.io_seconds += perf_counter() - tick
        return observe

    def finish(self, st, result, *, task, record=None, solve_seconds=None, input_seconds=None):
        """final_native.npz/.json and summary.json for the final state of `st` (the last formal commit or the
        certified initializer); returns the summary."""
        self._history.close()
        commits = len(self.commits)
        role = ("no_native_result" if result is None else "formal_native_outer_commit" if commits
                else "certified_native_initializer_NOT_native_outer_commit")
        final = None
        e2e = export = None
        cert = None
        if result is not None:
            tick = perf_counter()
            cert = st.state_certificate()
            final = save_npz(self.out, FINAL_FILE, u=st.u.cpu().numpy(), warped=st.warped.cpu().numpy())
            export = perf_counter() - tick
            e2e = perf_counter() - self.started
            dump(self.out / "final_native.json", dict(final, role=role, certificate=_certificate_dict(cert)))
        energy = None
        if result is not None and result.energy is not None:
            energy = dict(asdict(result.energy), curvature=0.)
        device = getattr(st, "device", None)
        summary = dict(
            task=task, status=result.status if result is not None else "transfer_failed",
            success=bool(result is not None and result.success),
            reason=result.reason if result is not None else "no certified transfer alpha",
            metrics_role=role,
            ratio=None if result is None else result.mse_ratio,
            native_mse=None if result is None else result.native_mse,
            final_mse=None if result is None else result.final_mse,
            GF0=None if result is None else result.G0, GF=None if result is None else result.GF,
            GF_threshold=None if result is None else result.threshold,
            GD=None if result is None else result.GD, GE=None if result is None else result.GE,
            certificate=_certificate_dict(cert), energy=energy,
            native_commits=commits, outer_rows=self.history_rows,
            attempted_inner_candidates=None if result is None else sum(r["inner_trials"] for r in result.history),
            input_seconds=input_seconds, solve_call_wall_seconds=solve_seconds,
            stored_array_io_seconds=self.io_seconds,
            peak_rss_KiB=resource.getrusage(resource.RUSAGE_SELF).ru_maxrss,
            device_peak_allocated_bytes=(torch.cuda.max_memory_allocated(device)
                                         if device is not None and device.type == "cuda" else None),
            e2e_seconds=e2e, final_export_seconds=export,
            e2e_scope="run start to final_native.npz written",
            independent_audit_pending=True,
            **({} if record is None else dict(route_record=record)))
        dump(self.out / "summary.json", summary)
        return summary


def run_pair(pair_dir, out_dir, *, device="cuda", route="preopt", seconds=None, export_reserve=1., max_outer=None,
             mu=.03, kappa=.002, preopt=None, save_commits=False, verify=True):
    """One registration from a pair folder into a new run folder.

    route "preopt": FP32 pyramid pre-optimisation, certified transfer, certified FP64 native solve with the
    zero-displacement baseline (the documented default run, usage.md:70-85); route "native": the certified FP64
    solve from u = 0. Weights: lam from the manifest, mu = nu = `mu`, kappa (run_stage_3d_e19_mri.py:1470-1479).
    `seconds` is the wall budget (the last `export_reserve` of it kept for the export); `max_outer` bounds the
    formal commits. Returns the summary dict."""
    from .hotpath import PairState
    from .preopt import PreoptConfig, preoptimized_solve
    from .solver import PGConfig, SequentialConfig, solve_sequential
    if route not in ROUTES:
        raise InputValidationError(f"route must be one of {ROUTES}")
    started = perf_counter()
    arrays, grid, manifest = load_pair(pair_dir, verify=verify)
    input_seconds = perf_counter() - started
    if np.any(arrays["initial_u"]):
        raise InputValidationError("requires the original zero initial field, not a resumed run")
    lam = (manifest.get("weights") or {}).get("lam")
    weights = Weights3D(float(lam) if lam is not None else default_lam(grid), mu, mu, kappa)
    problem = Q1Problem(grid, weights, 0., .05)
    budget = None
    if seconds is not None:
        deadline = started + float(seconds) - float(export_reserve)
        budget = lambda: perf_counter() >= deadline
    writer = RunWriter(out_dir, save_commits=save_commits, started=started)
    device = torch.device(device)
    fixed = torch.from_numpy(arrays["fixed"]).to(device)
    moving = torch.from_numpy(arrays["moving"]).to(device)
    task = dict(pair=manifest.get("pair"), route=route, seconds=seconds, export_reserve=export_reserve,
                max_outer=max_outer, weights=asdict(weights), input_sha256=manifest.get("pair_sha256"))
    tick = perf_counter()
    st = None
    try:
        if route == "native":
            st = PairState(problem, fixed, moving, device)
            st.set_state(theta=torch.zeros(st.nfree, dtype=torch.float64, device=device))
            result = solve_sequential(st, SequentialConfig(), PGConfig(), observer=writer.observer(st),
                                      max_outer=max_outer, budget=budget)
            record = None
        else:
            config = PreoptConfig.native() if preopt is None else preopt
            late = _LateObserver(writer)
            st, result, record = preoptimized_solve(problem, fixed, moving, device, config, max_outer=max_outer,
                                                    observer=late, budget=budget, on_state=late.bind)
        solve_seconds = perf_counter() - tick
    except BaseException:
        writer._history.close()
        raise
    return writer.finish(st, result, task=task, record=_plain(record), solve_seconds=solve_seconds,
                         input_seconds=input_seconds)


class _LateObserver:
    """An outer observer whose PairState is bound after construction (preoptimized_solve builds the state)."""

    def __init__(self, writer):
        self.writer = writer
        self.st = None
        self._f = None

    def bind(self, st):
        self.st = st
        self._f = self.writer.observer(st)

    def __call__(self, row):
        self._f(row)


def _plain(record):
    """A route record with tensors and NumPy values made JSON-ready."""
    if record is None:
        return None
    if isinstance(record, dict):
        return {k: _plain(v) for k, v in record.items()}
    if isinstance(record, (list, tuple)):
        return [_plain(v) for v in record]
    if isinstance(record, torch.Tensor):
        return record.tolist()
    if isinstance(record, np.ndarray):
        return record.tolist()
    if isinstance(record, np.generic):
        return record.item()
    return record


def recertify(run_dir, pair_dir, *, workers=None, config=None):
    """Independent re-certification of a run's final displacement (docs/usage.md:150-165): final_native.npz's u on
    the manifest's grid, the CPU interval certificate with delta = 0. Checks the file against the hash recorded in
    final_native.json first. Returns the FeasibilityResult."""
    run_dir = Path(run_dir)
    grid = grid_from_manifest(read_manifest(pair_dir))
    record = json.loads((run_dir / "final_native.json").read_text())
    path = run_dir / record.get("path", FINAL_FILE)
    if file_sha256(path) != record.get("sha256"):
        raise InputValidationError("final_native.npz does not match the hash in final_native.json")
    with np.load(path, allow_pickle=False) as archive:
        u = np.ascontiguousarray(archive["u"], dtype=np.float64)
    if u.shape != tuple(grid.shape) + (3,):
        raise InputValidationError("the stored u does not match the manifest's grid")
    cfg = CertificateConfig(delta=0.) if config is None else config
    return certify_map(u, grid, cfg, workers=workers or min(8, os.cpu_count() or 1))

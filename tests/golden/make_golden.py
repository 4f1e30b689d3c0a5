"""Writes the golden vectors in tests/golden/ from the reference's own CPU functions.

Runs only where the reference is mounted (/root/reference, in the build container); the vectors it writes are
plain data (inputs and the reference's outputs) and are all the tests read. Usage:

    python tests/golden/make_golden.py

What it records, and why only an import pins it:
* T69 node JD measure, jd_curl/three_d/solvers/jd_affine.py:154-225 (node_min_jd with extension_activation and
  stats) and the face/interior estimator rows (jd_affine.py:96-151: eigh, ties of the stable argsort);
* T62 node JD measure (the same function without the extension);
* T66 global first root, solvers/global_jd_step.py:107-129 (global_root_cpu);
* the historical-witness prediction, solvers/jd_step.py:177-239 (JDStepMemory.predict: numpy.roots and its
  imaginary-part tolerance);
* the whole-cell certificate with EnclosureCertificateConfig, feasibility.py:393-506 (edge hull, then the
  adaptive Bernstein proof, serial route);
* the Gauss-point J and curl of geometry.py:60-89, and the DST-I Poisson inverse of poisson.py:99-131.

The reference's solver chain (residuals.py -> data_weight.py -> nifti.py) imports nibabel, which this image does
not have, so energies, gradients and the stepwise trajectory are not recorded here.
"""
import json
import sys
from pathlib import Path
from types import SimpleNamespace

import numpy as np

REFERENCE = Path("/root/reference")
OUT = Path(__file__).resolve().parent
sys.path.insert(0, str(REFERENCE))

from jd_curl.three_d.cell_q1 import CellQ1Operators3D  # noqa: E402
from jd_curl.three_d.feasibility import CertificateConfig, EnclosureCertificateConfig, certify_map  # noqa: E402
from jd_curl.three_d.geometry import _compute_cell_geometry  # noqa: E402
from jd_curl.three_d.grid import make_grid_3d  # noqa: E402
from jd_curl.three_d.poisson import build_poisson_inverse  # noqa: E402
from jd_curl.three_d.problem import Weights3D, make_problem  # noqa: E402
from jd_curl.three_d.solvers import global_jd_step, jd_affine, jd_step  # noqa: E402

SHAPE = (7, 8, 9)                 # (nz, ny, nx)
SPACING = (1., 1.2, .9)           # (hx, hy, hz)


def smooth(shape, amplitude, seed, modes=3):
    """Zero-boundary sum of sine modes, (nz, ny, nx, 3)."""
    rng = np.random.default_rng(seed)
    z, y, x = np.meshgrid(*(np.linspace(0., 1., n) for n in shape), indexing="ij")
    u = np.zeros(tuple(shape) + (3,))
    for c in range(3):
        for _ in range(modes):
            kx, ky, kz = rng.integers(1, 3, size=3)
            u[..., c] += rng.uniform(-1, 1) * np.sin(np.pi * kx * x) * np.sin(np.pi * ky * y) * np.sin(np.pi * kz * z)
    u[[0, -1]] = 0.
    u[:, [0, -1]] = 0.
    u[:, :, [0, -1]] = 0.
    return amplitude * u


def rough(shape, amplitude, seed):
    """Zero-boundary i.i.d. normal field on the free nodes."""
    u = np.random.default_rng(seed).normal(size=tuple(shape) + (3,)) * amplitude
    u[[0, -1]] = 0.
    u[:, [0, -1]] = 0.
    u[:, :, [0, -1]] = 0.
    return u


def problem_for(shape, spacing, delta=0.):
    nz, ny, nx = shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=spacing[0], hy=spacing[1], hz=spacing[2])
    rng = np.random.default_rng(99)
    fixed, moving = rng.uniform(size=shape), rng.uniform(size=shape)
    return make_problem(grid=grid, fixed_image=fixed, moving_image=moving, u_ref=np.zeros(shape + (3,)),
                        weights=Weights3D(lam=1e-3, mu=.03, nu=.03, kappa=.002),
                        certificate_config=CertificateConfig(delta=delta), barrier_location="closed27_uniform",
                        barrier_activation=.05)


def restrict(field):
    return np.moveaxis(field[1:-1, 1:-1, 1:-1], -1, 0).copy().ravel(order="C")


def certificate_fields(res):
    keys = ("status", "reason", "min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd", "n_cells_checked",
            "n_uncertain", "n_infeasible", "n_boxes", "max_depth", "witness_cell", "witness_reference_point",
            "n_fast_cells", "n_fallback_cells")
    out = {k: getattr(res, k) for k in keys}
    if out["witness_reference_point"] is not None:
        out["witness_reference_point"] = [float(x) for x in out["witness_reference_point"]]
    for k in ("min_jd_lower_bound", "min_margin_lower_bound", "sampled_min_jd"):
        if out[k] is not None:
            out[k] = float(out[k]).hex()
    return out


def main():
    arrays, meta = {}, dict(shape=SHAPE, spacing=SPACING)
    problem = problem_for(SHAPE, SPACING)

    # ---- T62 / T69 node JD measure ----------------------------------------------------------------
    fields = {"mild": smooth(SHAPE, .9, 1) + rough(SHAPE, .16, 2),
              "folded": smooth(SHAPE, .8, 3) + rough(SHAPE, .25, 4),
              "dense": smooth(SHAPE, .9, 7) + rough(SHAPE, .2, 8)}
    t69 = {}
    for name, u in fields.items():
        arrays[f"u_{name}"] = u
        state = SimpleNamespace(u=u)
        arrays[f"t62_{name}"] = jd_affine.node_min_jd(problem, state)
        stats = {}
        arrays[f"t69_{name}"] = jd_affine.node_min_jd(problem, state, extension_activation=.25, stats=stats)
        t69[name] = {k: (float(v) if isinstance(v, float) else v) for k, v in stats.items()}
        J = _compute_cell_geometry(u, problem.barrier_ops, tile=slice(0, problem.grid.n_cells)).J
        minima = jd_affine.closed27_cell_minima(J)
        cand = jd_affine.extension_candidates(J, minima, .25)
        arrays[f"t69_rows_{name}"] = np.asarray(J[cand])
        arrays[f"t69_mu_{name}"] = jd_affine.face_interior_minima(J[cand])
    meta["t69_stats"] = t69

    # ---- T66 global root ----------------------------------------------------------------------------
    base = fields["mild"]
    v = smooth(SHAPE, 1.5, 5) + rough(SHAPE, .3, 6)
    arrays["t66_base"], arrays["t66_direction"] = base, restrict(v)
    roots = []
    for limit in (.05, .5, 2., 10.):
        r = global_jd_step.global_root_cpu(problem, SimpleNamespace(u=base), restrict(v), limit)
        roots.append(dict(limit=limit, root=None if r["root"] is None else float(r["root"]).hex(),
                          points=r["points"], invalid_base_points=r["invalid_base_points"]))
    meta["t66"] = roots

    # ---- historical-witness prediction ---------------------------------------------------------------
    memory = jd_step.JDStepMemory()
    memory.bind(problem.grid)
    witnesses = [(17, (.3, .7, .5)), (17, (0., .5, 1.)), (200, (.25, .125, .875)), (311, (.6, .05, .45)),
                 (0, (1., 1., 1.))]
    for cell, point in witnesses:
        memory.witness(cell, point)
    predictions = []
    base_state = SimpleNamespace(u=base, eval_id="golden")
    for limit in (.05, .5, 2., 10.):
        tau, info = memory.predict(problem, base_state, restrict(v), limit, .9, direction_id="golden")
        predictions.append(dict(limit=limit, tau=float(tau).hex(), points=info["points"],
                                root=None if info["root"] is None else float(info["root"]).hex(),
                                fallback=info["fallback"]))
    meta["witnesses"] = [[c, list(p)] for c, p in witnesses]
    meta["predictions"] = predictions
    coefficients = np.array([[.9, -2.1, 1.3, .2], [1., -3., 3., -1.], [.5, .1, .1, .1], [2., -1., -4., 4.],
                             [1e-3, -1., 0., 0.], [1., 0., -1.5, 0.]])
    arrays["first_root_coefficients"] = coefficients
    meta["first_roots"] = [[None if (r := jd_step.first_root_before(c, lim)) is None else float(r).hex()
                            for lim in (.3, 1., 5.)] for c in coefficients]

    # ---- Gauss geometry -----------------------------------------------------------------------------
    geo = _compute_cell_geometry(fields["folded"], CellQ1Operators3D(problem.grid), tile=slice(100, 160))
    arrays["gauss_J_folded"], arrays["gauss_curl_folded"] = geo.J, geo.curl      # cells 100..159

    # ---- Poisson inverse ----------------------------------------------------------------------------
    rhs = np.random.default_rng(7).normal(size=3 * problem.grid.n_free_scalar)
    arrays["poisson_rhs"] = rhs
    for i, (a, b) in enumerate(((.1, .03), (2.8e-6, .03), (0., 1.))):
        arrays[f"poisson_out_{i}"] = build_poisson_inverse(problem.grid, a=a, b=b).apply(rhs)
    meta["poisson_ab"] = [[.1, .03], [2.8e-6, .03], [0., 1.]]

    # ---- whole-cell certificate (EnclosureCertificateConfig, serial route) ---------------------------
    cshape = (6, 7, 8)
    cgrid = problem_for(cshape, SPACING).grid
    cases = {"smooth": smooth(cshape, .3, 11),
             "edge": smooth(cshape, .52, 12) + rough(cshape, .03, 13),
             "near": smooth(cshape, .54, 12) + rough(cshape, .04, 13),
             "tight": smooth(cshape, .9, 12) + rough(cshape, .05, 13),
             "folded": smooth(cshape, 1.6, 14) + rough(cshape, .2, 15),
             "rough": rough(cshape, .3, 16)}
    certs = {}
    for name, u in cases.items():
        arrays[f"cert_u_{name}"] = u
        for delta in (0., .01):
            res = certify_map(u, cgrid, config=EnclosureCertificateConfig(delta=delta))
            certs[f"{name}/{delta}"] = certificate_fields(res)
    meta["cert_shape"] = cshape
    meta["certificates"] = certs

    np.savez_compressed(OUT / "reference_vectors.npz", **arrays)
    (OUT / "reference_vectors.json").write_text(json.dumps(meta, indent=1, sort_keys=True) + "\n")
    print(f"wrote {len(arrays)} arrays, {OUT / 'reference_vectors.npz'}")
    for k, v in certs.items():
        print(k, v["status"], v["reason"], v["n_fast_cells"], v["n_fallback_cells"], v["n_boxes"])
    print("t69", t69)


if __name__ == "__main__":
    main()

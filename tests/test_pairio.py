"""On-disk formats: pair folders (pair.npz + manifest.json), the JSON conventions of the run files, and the
independent re-certification of a stored final displacement (seq_vp_3d_amd/pairio.py)."""
import hashlib
import json

import numpy as np
import pytest

from conftest import smooth_field
from seq_vp_3d_amd import _lib
from seq_vp_3d_amd import pairio
from seq_vp_3d_amd.problem import InputValidationError


def _images(shape, seed=0):
    rng = np.random.default_rng(seed)
    return rng.normal(size=shape), rng.normal(size=shape)


def test_pair_round_trip_and_manifest_fields(tmp_path):
    shape = (6, 7, 9)
    fixed, moving = _images(shape)
    labels = np.arange(np.prod(shape), dtype=np.int16).reshape(shape) % 5
    m = pairio.write_pair(tmp_path / "p", fixed, moving, spacing_xyz=(1.5, 1., .5), origin_xyz=(-3., 2., 7.),
                          extra=dict(dataset="synthetic"))
    # a label volume as the reference's pair sets carry (evaluation only): kept, not used
    with np.load(tmp_path / "p" / "pair.npz") as z:
        arrays = {k: z[k] for k in z.files}
    np.savez_compressed(tmp_path / "p" / "pair.npz", fixed_seg4=labels, **arrays)
    m["pair_sha256"] = hashlib.sha256((tmp_path / "p" / "pair.npz").read_bytes()).hexdigest()
    pairio.dump(tmp_path / "p" / "manifest.json", m)
    arrays, grid, manifest = pairio.load_pair(tmp_path / "p")
    np.testing.assert_array_equal(arrays["fixed"], fixed)
    np.testing.assert_array_equal(arrays["moving"], moving)
    np.testing.assert_array_equal(arrays["fixed_seg4"], labels)
    assert not arrays["initial_u"].any() and arrays["initial_u"].shape == shape + (3,)
    assert tuple(grid.shape) == shape and tuple(grid.spacing) == (1.5, 1., .5) and tuple(grid.origin) == (-3., 2., 7.)
    x, y, z = grid.axes
    assert x[0] == -3. and x[-1] == -3. + 8 * 1.5 and z[-1] == 7. + 5 * .5
    assert manifest["shape_zyx"] == list(shape) and manifest["unit"] == "mm" and manifest["dataset"] == "synthetic"
    # lam = 0.1 / L^2 with L the geometric mean of the extents (12, 6, 2.5 mm)
    assert abs(manifest["weights"]["lam"] - .1 / (12 * 6 * 2.5) ** (2 / 3)) <= 1e-15
    assert manifest["weights"]["kappa"] == .002
    # baseline MSE: trapezoid node weights (1/2 on each end node per axis) by brute force
    q = np.ones(shape)
    for ax in range(3):
        idx = [slice(None)] * 3
        for end in (0, -1):
            idx[ax] = end
            q[tuple(idx)] *= .5
    want = float(np.sum(q * (moving - fixed) ** 2) / np.sum(q))
    assert abs(manifest["baseline_native_mse"] - want) <= 1e-13 * want


def test_pair_validation(tmp_path):
    shape = (5, 5, 6)
    fixed, moving = _images(shape, 1)
    pairio.write_pair(tmp_path / "ok", fixed, moving)
    # a changed archive no longer matches the manifest's hash
    with np.load(tmp_path / "ok" / "pair.npz") as z:
        arrays = {k: z[k] for k in z.files}
    arrays["moving"] = arrays["moving"] + 1e-12
    np.savez_compressed(tmp_path / "ok" / "pair.npz", **arrays)
    with pytest.raises(InputValidationError, match="pair_sha256"):
        pairio.load_pair(tmp_path / "ok")
    pairio.load_pair(tmp_path / "ok", verify=False)

    def bad(name, mutate_manifest=None, **arrays):
        folder = tmp_path / name
        pairio.write_pair(folder, fixed, moving)
        if arrays:
            with np.load(folder / "pair.npz") as z:
                cur = {k: z[k] for k in z.files}
            cur.update(arrays)
            np.savez(folder / "pair.npz", **{k: v for k, v in cur.items() if v is not None})
        m = json.loads((folder / "manifest.json").read_text())
        m["pair_sha256"] = pairio.file_sha256(folder / "pair.npz")
        if mutate_manifest:
            mutate_manifest(m)
        pairio.dump(folder / "manifest.json", m)
        return folder

    with pytest.raises(InputValidationError, match="shape"):
        pairio.load_pair(bad("shape", moving=moving[:, :, :-1]))
    with pytest.raises(InputValidationError, match="non-finite"):
        f = fixed.copy()
        f[2, 2, 2] = np.nan
        pairio.load_pair(bad("nan", fixed=f))
    with pytest.raises(InputValidationError, match="initial_u"):
        pairio.load_pair(bad("u0", initial_u=np.zeros(shape + (2,))))
    with pytest.raises(InputValidationError, match="unit"):
        pairio.load_pair(bad("unit", lambda m: m.update(unit="cm")))
    with pytest.raises(InputValidationError, match="spacing"):
        pairio.load_pair(bad("spacing", lambda m: m.update(spacing_xyz=[1., 0., 1.])))
    with pytest.raises(InputValidationError, match="shape_zyx"):
        pairio.load_pair(bad("dims", lambda m: m.update(shape_zyx=[5, 5])))
    with pytest.raises(InputValidationError, match="lam"):
        pairio.load_pair(bad("lam", lambda m: m["weights"].update(lam=-1.)))
    with pytest.raises(InputValidationError, match="manifest"):
        pairio.load_pair(tmp_path / "missing")
    # object arrays are never unpickled
    with pytest.raises(ValueError):
        pairio.load_pair(bad("pickle", extra=np.array([{"a": 1}], dtype=object)))
    with pytest.raises(FileExistsError):
        pairio.write_pair(tmp_path / "ok", fixed, moving)


def test_json_conventions(tmp_path):
    path = tmp_path / "x.json"
    pairio.dump(path, dict(a=np.float64(1.5), b=np.int64(3), c=(1, 2), d=None, e=np.bool_(True)))
    assert json.loads(path.read_text()) == dict(a=1.5, b=3, c=[1, 2], d=None, e=True)
    assert path.read_text().endswith("\n") and not (tmp_path / "x.json.tmp").exists()
    with pytest.raises(ValueError):
        pairio.dump(path, dict(a=float("nan")))
    assert json.loads(path.read_text())["b"] == 3           # the failed dump left the old file
    with (tmp_path / "h.jsonl").open("w") as s:
        pairio.append_jsonl(s, dict(i=0, v=np.float32(.5)))
        pairio.append_jsonl(s, dict(i=1, v=None))
    rows = [json.loads(line) for line in (tmp_path / "h.jsonl").read_text().splitlines()]
    assert rows == [dict(i=0, v=.5), dict(i=1, v=None)]
    rec = pairio.save_npz(tmp_path, "a.npz", u=np.arange(6.).reshape(2, 3))
    assert rec["sha256"] == hashlib.sha256((tmp_path / "a.npz").read_bytes()).hexdigest()
    assert not (tmp_path / "a.npz.partial.npz").exists()


@pytest.mark.skipif(not _lib.LIB_PATH.exists(), reason="library not built")
def test_recertify_a_stored_displacement(tmp_path):
    shape = (8, 9, 10)
    fixed, moving = _images(shape, 2)
    pairio.write_pair(tmp_path / "pair", fixed, moving)
    run = tmp_path / "run"
    run.mkdir()
    u = smooth_field(shape, .05, 4)
    u[[0, -1]] = u[:, [0, -1]] = u[:, :, [0, -1]] = 0.          # exactly zero on the boundary, as solver maps are
    rec = pairio.save_npz(run, "final_native.npz", u=u, warped=moving)
    pairio.dump(run / "final_native.json", dict(rec, role="formal_native_outer_commit", certificate=None))
    cert = pairio.recertify(run, tmp_path / "pair", workers=2)
    assert cert.status == "certified_feasible" and cert.min_jd_lower_bound > 0
    # a node pushed past its neighbour folds the map: not certified
    folded = u.copy()
    folded[4, 4, 5] += np.array([2.5, 0., 0.])
    rec = pairio.save_npz(run, "final_native.npz", u=folded, warped=moving)
    pairio.dump(run / "final_native.json", dict(rec, role="formal_native_outer_commit", certificate=None))
    assert pairio.recertify(run, tmp_path / "pair", workers=2).status != "certified_feasible"
    # a file that is not the recorded one is refused
    np.savez(run / "final_native.npz", u=u, warped=moving)
    with pytest.raises(InputValidationError, match="hash"):
        pairio.recertify(run, tmp_path / "pair")


def test_run_pair_refuses_a_resumed_field(tmp_path):
    shape = (5, 6, 7)
    fixed, moving = _images(shape, 3)
    u0 = np.zeros(shape + (3,))
    u0[2, 3, 3] = .1
    pairio.write_pair(tmp_path / "p", fixed, moving, initial_u=u0)
    with pytest.raises(InputValidationError, match="zero initial field"):
        pairio.run_pair(tmp_path / "p", tmp_path / "run", device="cpu")
    with pytest.raises(InputValidationError, match="route"):
        pairio.run_pair(tmp_path / "p", tmp_path / "run", device="cpu", route="lm")
    assert not (tmp_path / "run").exists()

"""A registration from a pair folder on the device: the run files, their agreement with a direct solve of the same
inputs, the zero-displacement baseline of the pre-optimised route, the wall budget and the re-certification."""
import json

import numpy as np
import pytest

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")

SHAPE = (28, 30, 32)          # the coarsest of the four pyramid levels needs >= 5 nodes per axis
TIMING = ("seconds", "elapsed_seconds")


@pytest.fixture(scope="module")
def pair_dir(tmp_path_factory, gpu_device):
    from seq_vp_3d_amd import pairio
    from seq_vp_3d_amd.problem import make_grid_3d
    from seq_vp_3d_amd.synthetic import device_min_jd, device_sampler, make_pair, scaled_spec
    nz, ny, nx = SHAPE
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1., hz=1.)
    pair = make_pair(scaled_spec(SHAPE, 4100), gpu_device, device_sampler(gpu_device), device_min_jd(grid, gpu_device))
    folder = tmp_path_factory.mktemp("pairs") / "synthetic_28x30x32"
    pairio.write_pair(folder, pair.fixed.cpu().numpy(), pair.moving.cpu().numpy())
    return folder


def _rows(path):
    return [json.loads(line) for line in path.read_text().splitlines()]


def test_native_route_files_match_a_direct_solve(tmp_path, pair_dir, gpu_device):
    from seq_vp_3d_amd import pairio
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D
    from seq_vp_3d_amd.solver import PGConfig, SequentialConfig, solve_sequential
    print("case run_pair native 28x30x32, 2 outer steps", flush=True)
    run = tmp_path / "run"
    summary = pairio.run_pair(pair_dir, run, device=gpu_device, route="native", max_outer=2)
    # the same inputs solved directly
    arrays, grid, manifest = pairio.load_pair(pair_dir)
    w = Weights3D(manifest["weights"]["lam"], .03, .03, .002)
    st = PairState(Q1Problem(grid, w, 0., .05), arrays["fixed"], arrays["moving"], gpu_device)
    st.set_state(theta=torch.zeros(st.nfree, dtype=torch.float64, device=gpu_device))
    rows = []
    res = solve_sequential(st, SequentialConfig(), PGConfig(), max_outer=2, observer=rows.append)
    assert res.outer_accepted == 2
    with np.load(run / "final_native.npz", allow_pickle=False) as z:
        assert np.array_equal(z["u"], st.u.cpu().numpy()) and np.array_equal(z["warped"], st.warped.cpu().numpy())
    final = json.loads((run / "final_native.json").read_text())
    assert final["path"] == "final_native.npz" and final["sha256"] == pairio.file_sha256(run / "final_native.npz")
    assert final["role"] == "formal_native_outer_commit"
    assert final["certificate"]["status"] == "certified_feasible" and final["certificate"]["min_jd_lower_bound"] > 0
    history = _rows(run / "history.jsonl")
    assert len(history) == len(rows) == summary["outer_rows"]
    for a, b in zip(history, rows):
        assert a["level"] == 3
        for k, v in b.items():
            if k not in TIMING:
                assert a[k] == v, k
    commits = json.loads((run / "commits.json").read_text())
    assert len(commits) == summary["native_commits"] == 2
    assert [c["GF"] for c in commits] == [r["committed_GF"] for r in rows if r["outer_accepted"]]
    assert all(c["certificate"]["status"] == "certified_feasible" for c in commits)
    assert commits[-1]["certificate"] == final["certificate"]
    assert summary["status"] == res.status and summary["success"] is False
    assert (summary["GF0"], summary["GF"], summary["GF_threshold"]) == (res.G0, res.GF, res.threshold)
    assert summary["ratio"] == res.mse_ratio and summary["final_mse"] == res.final_mse
    # at u = 0 the warped image is the moving image (spline interpolation), so the baseline is the manifest's
    assert abs(summary["native_mse"] - manifest["baseline_native_mse"]) <= 1e-10 * manifest["baseline_native_mse"]
    assert summary["energy"]["curvature"] == 0. and summary["certificate"] == final["certificate"]
    assert json.loads((run / "summary.json").read_text()) == summary
    # independent re-certification of the stored map (CPU interval certificate, delta = 0)
    cert = pairio.recertify(run, pair_dir)
    assert cert.status == "certified_feasible"
    assert cert.min_jd_lower_bound == final["certificate"]["min_jd_lower_bound"]


def test_preopt_route_keeps_the_zero_baseline(tmp_path, pair_dir, gpu_device):
    from seq_vp_3d_amd import pairio
    from seq_vp_3d_amd.hotpath import PairState
    from seq_vp_3d_amd.preopt import PreoptConfig
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D
    print("case run_pair preopt 28x30x32, 1 outer step", flush=True)
    cfg = PreoptConfig(iterations=(6, 6, 6, 6), lr=.5, sigma_field=1., mu=(1., 1., .01, .003))
    run = tmp_path / "run"
    summary = pairio.run_pair(pair_dir, run, device=gpu_device, route="preopt", preopt=cfg, max_outer=1)
    record = summary["route_record"]
    assert record["transfer"]["alpha"] is not None, record["transfer"]
    # G0 of the native solve is the zero map's, not the transferred start's (multiresolution.py:496-502)
    arrays, grid, manifest = pairio.load_pair(pair_dir)
    w = Weights3D(manifest["weights"]["lam"], .03, .03, .002)
    zero = PairState(Q1Problem(grid, w, 0., .05), arrays["fixed"], arrays["moving"], gpu_device)
    zero.set_state(theta=torch.zeros(zero.nfree, dtype=torch.float64, device=gpu_device))
    zero.refresh_reference()
    assert summary["GF0"] == zero.outer_norms()[0]
    assert summary["GF_threshold"] == 1e-8 + 1e-6 * summary["GF0"]
    assert abs(summary["native_mse"] - manifest["baseline_native_mse"]) <= 1e-10 * manifest["baseline_native_mse"]
    final = json.loads((run / "final_native.json").read_text())
    assert final["certificate"]["status"] == "certified_feasible"
    assert pairio.recertify(run, pair_dir).status == "certified_feasible"


def test_spent_budget_exports_the_initializer(tmp_path, pair_dir, gpu_device):
    from seq_vp_3d_amd import pairio
    print("case run_pair with a spent budget", flush=True)
    run = tmp_path / "run"
    summary = pairio.run_pair(pair_dir, run, device=gpu_device, route="native", seconds=0., export_reserve=0.)
    assert summary["status"] == "outer_time_budget" and summary["native_commits"] == 0
    assert summary["metrics_role"] == "certified_native_initializer_NOT_native_outer_commit"
    with np.load(run / "final_native.npz", allow_pickle=False) as z:
        assert not z["u"].any()
    assert json.loads((run / "final_native.json").read_text())["certificate"]["status"] == "certified_feasible"
    assert not (run / "commits.json").exists() and (run / "history.jsonl").read_text() == ""

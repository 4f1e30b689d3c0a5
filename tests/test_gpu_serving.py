"""Serving form: continuous batching on one GPU gives every request the result it gets alone."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _requests(gpu_device, shapes, max_outer=3):
    from seq_vp_3d_amd.serving import Request
    from seq_vp_3d_amd.synthetic import device_min_jd, device_sampler, make_pair, scaled_spec
    from seq_vp_3d_amd.problem import make_grid_3d
    out = []
    for i, shape in enumerate(shapes):
        spec = scaled_spec(shape, 3000 + i)
        grid = make_grid_3d(nx=shape[2], ny=shape[1], nz=shape[0], hx=1., hy=1., hz=1.)
        pair = make_pair(spec, gpu_device, device_sampler(gpu_device), device_min_jd(grid, gpu_device))
        out.append(Request(pair.fixed.cpu().numpy(), pair.moving.cpu().numpy(), max_outer=max_outer, request_id=i))
    return out


SHAPES = [(24, 26, 28), (20, 22, 24), (28, 24, 20), (22, 22, 22)]


def test_results_are_the_same_alone_and_in_the_mix(gpu_device, monkeypatch):
    from seq_vp_3d_amd import serving
    from seq_vp_3d_amd.serving import serial, serve
    reqs = _requests(gpu_device, SHAPES)
    print("case serving serial", flush=True)
    alone, _ = serial(reqs, gpu_device)
    print("case serving mix", flush=True)
    # the slots never capture graphs: a capture open on one thread fails the other threads' blocking copies
    seen = []
    original = serving.run_request

    def recording(*args, **kw):
        seen.append(kw.get("graphs", True))
        return original(*args, **kw)

    monkeypatch.setattr(serving, "run_request", recording)
    mixed, _, server = serve(reqs, None, gpu_device, slots=3)
    assert seen == [False] * len(reqs)
    assert max(a[4] for a in server.admissions) >= 2          # requests did overlap on the device
    for a, b in zip(alone, mixed):
        assert a.theta_digest == b.theta_digest and a.energy == b.energy and a.status == b.status
        assert a.outer_accepted == 3 and a.recertified == "certified_feasible"
        assert b.recertified == "certified_feasible"


def test_admission_respects_the_hbm_budget(gpu_device):
    from seq_vp_3d_amd.serving import footprint_bytes, serve
    reqs = _requests(gpu_device, SHAPES[:3], max_outer=1)
    budget = max(footprint_bytes(r.shape) for r in reqs) + 1
    print("case serving budget", flush=True)
    results, _, server = serve(reqs, [0., .01, .02], gpu_device, slots=3, hbm_budget=budget)
    assert all(a[4] == 1 for a in server.admissions)              # one at a time under this budget
    assert all(a[3] <= budget for a in server.admissions)
    assert [r.request_id for r in results] == [0, 1, 2]
    from seq_vp_3d_amd.serving import Server
    big = _requests(gpu_device, [(30, 30, 30)], max_outer=1)[0]
    s = Server(gpu_device, slots=1, hbm_budget=footprint_bytes((24, 24, 24)))
    with pytest.raises(MemoryError):
        s.submit(big).result()
    s.close()


def test_native_pair_in_the_mix_is_the_same_alone(gpu_device):
    """Uneven sizes in flight together: one 160x192x224 pair with 96^3 and 128^3 pairs on three slots. Every
    result is re-certified by the CPU interval certificate and is bit for bit its alone result."""
    from seq_vp_3d_amd.serving import serial, serve
    shapes = [(160, 192, 224), (96, 96, 96), (128, 128, 128), (96, 96, 96)]
    reqs = _requests(gpu_device, shapes, max_outer=2)
    print("case serving native mix serial", flush=True)
    alone, _ = serial(reqs, gpu_device)
    print("case serving native mix served", flush=True)
    mixed, _, server = serve(reqs, None, gpu_device, slots=3)
    assert max(a[4] for a in server.admissions) >= 2
    for a, b in zip(alone, mixed):
        assert a.theta_digest == b.theta_digest and a.energy == b.energy and a.status == b.status, a.request_id
        assert a.recertified == b.recertified == "certified_feasible"
        assert a.outer_accepted == b.outer_accepted == 2

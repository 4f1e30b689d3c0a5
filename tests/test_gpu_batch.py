"""Batching form: the multi-pair persistent trial-energy launch (seqvp_q1_trial_energy_batch) gives every pair the
bits of its own single-pair sweep for every launch shape, and pairs served through a BatchGroup get the results
they get alone."""
import numpy as np
import pytest

from conftest import rough_field, smooth_field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fields(shape, seed, device):
    u = smooth_field(shape, .08, seed) + rough_field(shape, .004, seed + 1)
    ref = smooth_field(shape, .05, seed + 2)
    base = u + rough_field(shape, .003, seed + 3)
    return [torch.tensor(a, device=device) for a in (u, ref, base)]


def test_batched_launch_is_each_pairs_own_sweep(gpu_device):
    """Six pairs of uneven sizes (one of 12 cells and one with a partial last block), every candidate launch shape
    and tiles up to 8: each pair's 7 results equal seqvp_q1_trial_energy's bit for bit."""
    from seq_vp_3d_amd import batch as B
    from seq_vp_3d_amd.problem import Q1Problem, make_grid_3d, native_weights
    from seq_vp_3d_amd.q1 import OP_TRIAL, FusedQ1
    from seq_vp_3d_amd import _lib
    shapes = [(20, 22, 24), (3, 4, 5), (33, 17, 29), (16, 16, 16), (40, 24, 18), (9, 31, 11)]
    print("case batched trial energy", shapes, flush=True)
    pairs = []
    for i, shape in enumerate(shapes):
        nz, ny, nx = shape
        grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1. + .1 * i, hy=.9, hz=1.2)
        q = FusedQ1(Q1Problem(grid, native_weights(shape), 0., .3), gpu_device)
        u, ref, base = _fields(shape, 10 * i, gpu_device)
        want = q.trial_energy(u, ref, base).cpu().numpy()
        ws = torch.empty(_lib.workspace_bytes(OP_TRIAL, shape), dtype=torch.uint8, device=gpu_device)
        pairs.append((q, u, ref, base, want, ws))
    handle = _lib.stream_handle(gpu_device)
    for tile, grid in tuple(B.CANDIDATES) + ((8, 0), (3, 8), (1, 16)):
        outs = [torch.full((7,), np.nan, dtype=torch.float64, device=gpu_device) for _ in pairs]
        entries = [B.entry(q, u, ref, base, o, ws) for (q, u, ref, base, _, ws), o in zip(pairs, outs)]
        B.trial_energy_batch(entries, tile, grid, handle)
        for (_, _, _, _, want, _), o, shape in zip(pairs, outs, shapes):
            got = o.cpu().numpy()
            assert np.array_equal(got, want), (tile, grid, shape, got, want)


def test_batched_launch_rejects_bad_tables(gpu_device):
    from seq_vp_3d_amd import batch as B, _lib
    from seq_vp_3d_amd.problem import Q1Problem, make_grid_3d, native_weights
    from seq_vp_3d_amd.q1 import OP_TRIAL, FusedQ1
    shape = (6, 7, 8)
    q = FusedQ1(Q1Problem(make_grid_3d(nx=8, ny=7, nz=6, hx=1., hy=1., hz=1.), native_weights(shape), 0., .3),
                gpu_device)
    u, ref, base = _fields(shape, 3, gpu_device)
    out = torch.empty(7, dtype=torch.float64, device=gpu_device)
    ws = torch.empty(_lib.workspace_bytes(OP_TRIAL, shape), dtype=torch.uint8, device=gpu_device)
    e = B.entry(q, u, ref, base, out, ws)
    handle = _lib.stream_handle(gpu_device)
    with pytest.raises(ValueError):
        B.trial_energy_batch([e] * (B.TRIAL_BATCH_MAX + 1), 1, 0, handle)
    for tile, grid in ((0, 0), (65, 0), (1, -1)):
        with pytest.raises(_lib.SeqvpError):
            B.trial_energy_batch([e], tile, grid, handle)
    small = B.TrialBatchEntry(*[getattr(e, f) for f, _ in B.TrialBatchEntry._fields_[:-1]], 8)
    with pytest.raises(_lib.SeqvpError):
        B.trial_energy_batch([small], 1, 0, handle)


def test_batch_group_results_are_the_same_alone(gpu_device):
    """Four uneven pairs served on three slots through one BatchGroup (its launches tuned online): every result is
    bit for bit its alone result and re-certified; the group's launches held more than one pair."""
    from test_gpu_serving import SHAPES, _requests
    from seq_vp_3d_amd.batch import BatchGroup
    from seq_vp_3d_amd.serving import serial, serve
    reqs = _requests(gpu_device, SHAPES)
    print("case batch group serial", flush=True)
    alone, _ = serial(reqs, gpu_device)
    print("case batch group served", flush=True)
    group = BatchGroup(gpu_device, max_wait=.02)
    mixed, _, server = serve(reqs, None, gpu_device, slots=3, batch=group)
    for a, b in zip(alone, mixed):
        assert a.theta_digest == b.theta_digest and a.energy == b.energy and a.status == b.status
        assert b.recertified == "certified_feasible"
    assert group.launches > 0 and group.pairs_launched > group.launches

"""The allocation-free soft energy (seqvp_q1_soft_energy_noalloc) through the C ABI against seqvp_q1_soft_energy
(the entry point with a caller workspace): the same bits, on two calls in a row whose second grid is smaller (the
fixed scratch keeps the first call's partials), with and without a base; a grid beyond the fixed scratch is refused
with SEQVP_EWORKSPACE before anything is launched."""
import numpy as np
import pytest

from conftest import rough_field, smooth_field

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _fields(shape, seed, device):
    ref = smooth_field(shape, .15, seed)
    base = ref + smooth_field(shape, .3, seed + 1) + rough_field(shape, .03, seed + 2)
    u = base + smooth_field(shape, .1, seed + 3) + rough_field(shape, .02, seed + 4)
    return [torch.tensor(a, dtype=torch.float64, device=device) for a in (u, ref, base)]


def _problem(shape):
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D, make_grid_3d
    nz, ny, nx = shape
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=1., hy=1.3, hz=.8)
    return Q1Problem(grid, Weights3D(lam=2e-3, mu=.03, nu=.05, kappa=.002), 0., .05)


def test_noalloc_soft_energy_equals_the_workspace_entry_point(gpu_device):
    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.q1 import FusedQ1
    stream = _lib.stream_handle(gpu_device)
    for shape, seed in (((30, 34, 38), 10), ((12, 13, 14), 20)):
        print("case noalloc soft energy", shape, flush=True)
        q = FusedQ1(_problem(shape), gpu_device)
        U, R, B = _fields(shape, seed, gpu_device)
        nz, ny, nx = shape
        for base in (B, None):
            want = q.soft_energy(U, R, base)
            got = torch.full((3,), float("nan"), dtype=torch.float64, device=gpu_device)
            _lib.call("seqvp_q1_soft_energy_noalloc", _lib.ptr(U), _lib.ptr(R), _lib.ptr(base), _lib.ptr(q.k),
                      nx, ny, nz, _lib.ptr(got), stream)
            torch.cuda.synchronize(gpu_device)
            assert torch.equal(got, want), (shape, base is None, got.tolist(), want.tolist())
            assert float(got[0]) > 0 and np.isfinite(got.cpu().numpy()).all()


def test_noalloc_refuses_grids_beyond_its_scratch(gpu_device):
    from seq_vp_3d_amd import _lib
    from seq_vp_3d_amd.q1 import FusedQ1
    shape = (12, 13, 14)
    q = FusedQ1(_problem(shape), gpu_device)
    U, R, _ = _fields(shape, 30, gpu_device)
    out = torch.zeros(3, dtype=torch.float64, device=gpu_device)
    # 299^3 = 26.7M cells need 104,524 blocks of partials, more than the 87,381 the scratch holds: refused on the
    # host before any launch (the small arrays are never read)
    rc = _lib.raw("seqvp_q1_soft_energy_noalloc", _lib.ptr(U), _lib.ptr(R), None, _lib.ptr(q.k), 300, 300, 300,
                  _lib.ptr(out), _lib.stream_handle(gpu_device))
    assert rc == -3
    assert torch.equal(out, torch.zeros(3, dtype=torch.float64, device=gpu_device))

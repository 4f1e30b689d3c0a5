import sys
from pathlib import Path

import numpy as np
import pytest

ROOT = Path(__file__).resolve().parents[1]
if str(ROOT) not in sys.path:
    sys.path.insert(0, str(ROOT))


def pytest_configure(config):
    config.addinivalue_line("markers", "gpu: needs an MI355X (gfx950) GPU and the built libseqvp_hip.so")


def smooth_field(shape_zyx, amplitude, seed, modes=3):
    """Zero-boundary smooth displacement (sum of sine modes), (nz, ny, nx, 3)."""
    rng = np.random.default_rng(seed)
    nz, ny, nx = shape_zyx
    z, y, x = np.meshgrid(*(np.linspace(0., 1., n) for n in shape_zyx), indexing="ij")
    u = np.zeros(shape_zyx + (3,))
    for c in range(3):
        for _ in range(modes):
            kx, ky, kz = rng.integers(1, 3, size=3)
            u[..., c] += rng.uniform(-1, 1) * np.sin(np.pi * kx * x) * np.sin(np.pi * ky * y) * np.sin(np.pi * kz * z)
    return amplitude * u


def rough_field(shape_zyx, amplitude, seed):
    """Zero-boundary random field (i.i.d. on the free nodes)."""
    rng = np.random.default_rng(seed)
    u = rng.normal(size=tuple(shape_zyx) + (3,)) * amplitude
    u[[0, -1]] = 0.
    u[:, [0, -1]] = 0.
    u[:, :, [0, -1]] = 0.
    return u


def counterexample_cell(d):
    """The T05 cell (edge_hull_gpu.py:214-218): positive corner JD, negative JD inside."""
    v = np.array([(x, y, z) for z in (0., 1.) for y in (0., 1.) for x in (0., 1.)])
    x, y, z = v.T
    return np.stack((x, y + 2 * (x - .5) * z, -2 * (x - .5) * y + d * z), axis=-1) - v


@pytest.fixture(scope="session")
def gpu_device():
    torch = pytest.importorskip("torch")
    if not torch.cuda.is_available():
        pytest.skip("no GPU")
    return torch.device("cuda:0")

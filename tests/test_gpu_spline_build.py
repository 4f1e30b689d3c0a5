"""Spline coefficients built on the device (seqvp_axis_transform) against the host build (SciPy)."""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("shape,spacing", [((9, 11, 13), (1., 1.2, .8)), ((40, 37, 45), (1., 1., 1.))])
def test_device_coefficients_match_make_interp_spline(gpu_device, shape, spacing):
    import torch
    from seq_vp_3d_amd.problem import make_grid_3d
    from seq_vp_3d_amd.spline import DeviceSpline, build_image_sampler
    grid = make_grid_3d(nx=shape[2], ny=shape[1], nz=shape[0], hx=spacing[0], hy=spacing[1], hz=spacing[2])
    img = np.random.default_rng(0).normal(size=shape).cumsum(axis=2)
    print("case spline build", shape, flush=True)
    host = build_image_sampler(img, grid)
    dev = DeviceSpline.from_image(torch.tensor(img, device=gpu_device), grid)
    c = dev.c.cpu().numpy()
    assert np.max(np.abs(c - host.coefficients)) <= 1e-12 * np.max(np.abs(host.coefficients))
    for a, b in zip(dev.t, host.knots):
        assert np.array_equal(a.cpu().numpy(), b)
    # values at the nodes reproduce the image (interpolation), gradients agree with the host spline's
    X = torch.tensor(grid.identity(), device=gpu_device)
    v, g = dev.sample(X.reshape(-1, 3))
    np.testing.assert_allclose(v.cpu().numpy().reshape(shape), img, rtol=0, atol=1e-11 * np.abs(img).max())
    vh, gh = DeviceSpline(host, grid, gpu_device).sample(X.reshape(-1, 3))
    np.testing.assert_allclose(g.cpu().numpy(), gh.cpu().numpy(), rtol=0, atol=1e-11 * np.abs(gh.cpu().numpy()).max())

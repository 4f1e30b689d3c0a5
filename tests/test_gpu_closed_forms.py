"""Device energies and gradient on affine maps against closed forms (no oracle in the loop).

u = (M - I)X gives F = M at every Gauss and closed point, so with C cells of volume omega:
E_jd = mu/2 C omega (det M - det M_r)^2, E_curl = nu/2 C omega |curl M - curl M_r|^2 (residuals.py:166-172),
B = kappa C omega psi(det M) (barrier.py:19-36, problem.py:309-311), the stable differences are the differences
of those (energy_comparison.py:196-240), the node terms are trapezoid-mass sums (grid.py:37-39, residuals.py:189-195)
and the soft + barrier gradient vanishes at every interior node.
"""
import numpy as np
import pytest

from test_oracle_pins import _curl, affine_case, psi_mp

pytestmark = pytest.mark.gpu
torch = pytest.importorskip("torch")


def _problem(og, m):
    from seq_vp_3d_amd.problem import Q1Problem, Weights3D, make_grid_3d
    nz, ny, nx = og.shape
    hx, hy, hz = og.spacing
    x0, y0, z0 = og.origin
    grid = make_grid_3d(nx=nx, ny=ny, nz=nz, hx=hx, hy=hy, hz=hz, x0=x0, y0=y0, z0=z0)
    return Q1Problem(grid, Weights3D(lam=m.lam, mu=m.mu, nu=m.nu, kappa=m.kappa), m.delta, m.activation)


@pytest.mark.parametrize("shape", [(5, 6, 7), (9, 17, 70)])
def test_affine_energies_closed_forms(gpu_device, shape):
    from seq_vp_3d_amd.q1 import FusedQ1
    print("case affine closed forms", shape, flush=True)
    og, m, (M, Mr, Mb), (u, ref, base) = affine_case(shape=shape)
    q = FusedQ1(_problem(og, m), gpu_device)
    U, R, B = (torch.tensor(a, device=gpu_device) for a in (u, ref, base))
    C, om = og.n_cells, og.omega
    dM, dR, dB = (np.linalg.det(x) for x in (M, Mr, Mb))
    cm, cr, cb = _curl(M), _curl(Mr), _curl(Mb)
    jd, curl, diff = q.soft_energy(U, R, B).cpu().numpy()
    want_jd = .5 * m.mu * C * om * (dM - dR) ** 2
    want_curl = .5 * m.nu * C * om * np.sum((cm - cr) ** 2)
    assert abs(jd - want_jd) <= 1e-12 * want_jd
    assert abs(curl - want_curl) <= 1e-12 * want_curl
    want = .5 * C * om * (m.mu * ((dM - dR) ** 2 - (dB - dR) ** 2) + m.nu * (np.sum((cm - cr) ** 2)
                                                                          - np.sum((cb - cr) ** 2)))
    assert abs(diff - want) <= 1e-12 * (want_jd + want_curl)
    val, jmin, chg, bmin = q.barrier_energy(U, B).cpu().numpy()
    psiM, psiB = (float(psi_mp(d, 0., m.activation)) for d in (dM, dB))
    assert abs(val - m.kappa * C * om * psiM) <= 1e-12 * val
    assert abs(jmin - dM) <= 1e-14 and abs(bmin - dB) <= 1e-14
    # the device returns the unweighted change; the weight is kappa omega / 27
    assert abs(chg * m.barrier_weight - m.kappa * C * om * (psiM - psiB)) <= 1e-11 * val
    # node terms: [image, proximity, ...] with constant residuals
    fixed = torch.zeros(og.shape, dtype=torch.float64, device=gpu_device)
    v = torch.tensor([.1, -.2, .05], dtype=torch.float64, device=gpu_device)
    image, prox = q.node_energy(fixed + .3, fixed, R + v, R, m.lam).cpu().numpy()[:2]
    assert abs(image - .5 * .09 * C * om) <= 1e-13 * image
    assert abs(prox - .5 * m.lam * C * om * float(v @ v)) <= 1e-13 * prox


def test_affine_gradient_vanishes_inside(gpu_device):
    from seq_vp_3d_amd.q1 import FusedQ1
    og, m, (M, Mr, _), (u, ref, _) = affine_case(shape=(9, 17, 70))
    q = FusedQ1(_problem(og, m), gpu_device)
    U, R = (torch.tensor(a, device=gpu_device) for a in (u, ref))
    out = q.gradient_nodes(torch.zeros_like(U), U, R).cpu().numpy()
    scale = np.abs(out).max()
    assert scale > 1e-6
    assert np.abs(out[1:-1, 1:-1, 1:-1]).max() <= 1e-13 * scale

"""The Bloch-wave Laplacian check of gpaw/test/parallel/ut_gucops.py, with
assertions (the reference script only prints the deviations).

Periodic cells (orthogonal, hexagonal and fcc: the last two take the
non-orthogonal NewGUCLaplace stencils) hold a Gaussian wave
f(r) = sum_n exp(-|r - r0 - n.L|^2 / 2 sigma^2) exp(i k.r) (its periodic
images n, gpaw/utilities/gauss.py:106-160), whose Laplacian is known in
closed form: sum over the images of
(sum_v (i k_v - d_v / sigma^2)^2 - 3 / sigma^2) g e^{ik.r}
(ut_gucops.py:261-271).  The finite-difference Laplacian with the Bloch
phases phase_cd = exp(2 pi i sdisp_cd k_c) must approach it at the order of
the stencil: the range-3 error is far below the range-1 error, and both are
bounded.  On the GPU the device operator must also equal the oracle bit for
bit.
"""
import numpy as np
import pytest

import oracle as O

CELLS = {
    'sc': np.diag([5.2, 5.6, 6.0]),
    'hexagonal': np.array([[5.6, 0.0, 0.0], [2.8, 5.6 * 3 ** 0.5 / 2, 0.0], [0.0, 0.0, 6.2]]),
    'fcc': 3.1 * np.array([[0.0, 1.0, 1.0], [1.0, 0.0, 1.0], [1.0, 1.0, 0.0]]),
}
N = (28, 28, 30)
K_C = np.array([0.25, -0.125, 0.375])


def bloch_gaussian(g, k_c):
    """The wave, its analytic Laplacian and the Bloch phases on grid g."""
    cell = g.cell_cv
    r_vG = g.grid_point_coordinates()
    k_v = 2 * np.pi * np.linalg.solve(cell, k_c)   # k . L_a = 2 pi k_a
    sigma = 0.15 * np.sqrt((cell ** 2).sum(1)).min()
    r0 = cell.sum(0) * 0.45
    f = np.zeros(r_vG.shape[1:], complex)
    lap = np.zeros_like(f)
    eikr = np.exp(1j * np.tensordot(k_v, r_vG, axes=1))
    for n in np.ndindex(5, 5, 5):
        shift = (np.array(n) - 2) @ cell
        d = r_vG - (r0 + shift)[:, None, None, None]
        gau = np.exp(-(d ** 2).sum(0) / (2 * sigma ** 2))
        fac = sum((1j * k_v[v] - d[v] / sigma ** 2) ** 2 for v in range(3)) - 3 / sigma ** 2
        f += gau * eikr
        lap += fac * gau * eikr
    return f, lap, g.phase_cd(k_c)


def _errors(apply):
    out = {}
    for name, cell in CELLS.items():
        g = O.Grid(N, cell, True)
        f, lap, ph = bloch_gaussian(g, K_C)
        scale = np.abs(lap).max()
        out[name] = {n: np.abs(apply(g, cell, n, f, ph) - lap).max() / scale for n in (1, 3)}
    return out


def _oracle_apply(g, cell, n, f, ph):
    b = np.zeros_like(f)
    O.laplace(g, 1.0, n, dtype=complex).apply(f, b, ph)
    return b


def _check(err):
    for name, e in err.items():
        assert e[1] < 2e-2, (name, e)
        assert e[3] < 2e-4, (name, e)
        assert e[3] < e[1] / 20, (name, e)


def test_bloch_gaussian_laplacian_oracle():
    _check(_errors(_oracle_apply))


@pytest.mark.gpu
def test_bloch_gaussian_laplacian_device():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    print('case ut_gucops Bloch Gaussian', flush=True)
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor

    def device_apply(g, cell, n, f, ph):
        gd = GridDescriptor(N, cell, True)
        a = torch.from_numpy(f).to('cuda')
        b = torch.empty_like(a)
        Laplace(gd, 1.0, n, complex).apply(a, b, ph)
        got = b.cpu().numpy()
        assert np.array_equal(got, _oracle_apply(g, cell, n, f, ph))
        return got

    _check(_errors(device_apply))


# --- the reference's other test scripts, through the HIP operators --------
def _torch():
    import torch
    if not torch.cuda.is_available():
        pytest.skip('no GPU')
    return torch


@pytest.mark.gpu
def test_laplace_gradient_polynomial_exactness_device():
    """gpaw/test/laplace.py through the device Laplace / Gradient: on the
    eleven cells and n = 1..5, every monomial of degree <= 2n has the exact Laplacian
    (2 for x^2-type terms, else 0) and gradient at the central point, to
    2e-12."""
    print('case laplace.py', flush=True)
    t = _torch()
    from test_oracle_math import CELLS as POLY_CELLS
    from gpaw_amd.fd_operators import Gradient, Laplace
    from gpaw_amd.grid_descriptor import GridDescriptor
    for name, D, cell in POLY_CELLS:
        for n in range(1, 6):
            Nn = 2 * n + 2
            g = O.Grid((Nn,) * 3, cell)
            gd = GridDescriptor((Nn,) * 3, cell)
            r_gv = g.grid_point_coordinates().transpose((1, 2, 3, 0)) - g.cell_cv.sum(0) / 2
            lap = Laplace(gd, 1.0, n)
            grads = [Gradient(gd, v, 1.0, n) for v in range(3)]
            b = t.empty((Nn,) * 3, dtype=t.float64, device='cuda')
            for m in range(0, 2 * n + 1):
                for ix in range(m + 1):
                    for iy in range(m - ix + 1):
                        iz = m - ix - iy
                        a = t.from_numpy(np.ascontiguousarray(
                            (r_gv ** (ix, iy, iz)).prod(3))).to('cuda')
                        r = 2.0 if (m == 2 and max(ix, iy, iz) == 2) else 0.0
                        lap.apply(a, b)
                        assert abs(float(b[n + 1, n + 1, n + 1]) - r) < 2e-12, (name, n, m)
                        for v in range(3):
                            grads[v].apply(a, b)
                            r = 1.0 if (m == 1 and (ix, iy, iz)[v] == 1) else 0.0
                            assert abs(float(b[n + 1, n + 1, n + 1]) - r) < 2e-12, (name, n, v)


@pytest.mark.gpu
@pytest.mark.parametrize('dtype', [float, complex])
def test_transformations_device(dtype):
    """gpaw/test/transformations.py through the device Transformer: K = 2,
    4, 6, 8 interpolation and restriction of random data on a periodic 20^3
    grid keep the integral to 1e-14."""
    print('case transformations.py', dtype, flush=True)
    t = _torch()
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.transformers import Transformer
    n = 20
    gd = GridDescriptor((n, n, n))
    rng = np.random.RandomState(8)
    a = rng.random_sample((n, n, n)).astype(dtype)
    if dtype == complex:
        a.imag = rng.random_sample((n, n, n))
    phase = np.ones((3, 2), complex)
    for g2 in (gd.refine(), gd.coarsen()):
        for k in (2, 4, 6, 8):
            b = t.empty(tuple(g2.n_c), dtype=t.float64 if dtype == float else t.complex128,
                        device='cuda')
            Transformer(gd, g2, k // 2, dtype).apply(t.from_numpy(a).to('cuda'), b, phase)
            ia = gd.dv * a.sum()
            ib = g2.dv * b.cpu().numpy().sum()
            assert abs(ia - ib) < 1e-14, (k, g2.N_c)


@pytest.mark.gpu
def test_non_periodic_device():
    """gpaw/test/non_periodic.py through the device Transformer: zero-boundary
    interpolation and restriction keep the integral to 1e-10."""
    print('case non_periodic.py', flush=True)
    t = _torch()
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.transformers import Transformer
    n = 20
    g1 = GridDescriptor((n, n, n), (8.0, 8.0, 8.0), pbc_c=False)
    a1 = np.random.RandomState(8).random_sample(tuple(g1.n_c))
    g2 = g1.refine()
    a2 = t.empty(tuple(g2.n_c), dtype=t.float64, device='cuda')
    Transformer(g1, g2).apply(t.from_numpy(a1).to('cuda'), a2)
    h2 = a2.cpu().numpy()
    assert abs(g1.dv * a1.sum() - g2.dv * h2.sum()) < 1e-10
    for sl in (0, -1):
        h2[sl] = 0.0
        h2[:, sl] = 0.0
        h2[:, :, sl] = 0.0
    back = t.empty(tuple(g1.n_c), dtype=t.float64, device='cuda')
    Transformer(g2, g1).apply(t.from_numpy(h2).to('cuda'), back)
    assert abs(g1.dv * back.cpu().numpy().sum() - g2.dv * h2.sum()) < 1e-10


@pytest.mark.gpu
def test_gauss_func_device():
    """gpaw/test/gauss_func.py (L = 0) through the device PoissonSolver: the
    potential of the generalized Gaussian on a 48^3 zero-boundary cell is
    within 0.6 (L2 norm) of the analytic erf potential."""
    print('case gauss_func.py', flush=True)
    t = _torch()
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import PoissonSolver
    g = O.Grid((48, 48, 48), (20.0, 20.0, 20.0), 0)
    gd = GridDescriptor((48, 48, 48), (20.0, 20.0, 20.0), False)
    ng, vg = O.gpaw_oracle.gaussian_monopole(g)
    ps = PoissonSolver(nn=3)
    ps.set_grid_descriptor(gd)
    ps.initialize()
    pot = t.zeros(tuple(gd.n_c), dtype=t.float64, device='cuda')
    ps.solve(pot, t.from_numpy(ng).to('cuda'), charge=None, zero_initial_phi=True)
    p = pot.cpu().numpy()
    assert (g.dv * ((p - vg) ** 2).sum()) ** 0.5 < 0.6

"""The oracle's restatement of the eigensolver preconditioner
(gpaw/preconditioner.py:13-70), pinned without the reference:

* closed forms derived by hand from the reference's sequence: a constant
  residual c on a periodic grid maps to -41 * step * c (every Laplacian
  annihilates constants, every transfer preserves them); a checkerboard
  residual restricts to exactly zero, so the result is exactly -step * r;
* a Bloch plane wave along axis 0, e^{i theta i0} times a period-4 pattern,
  stays in the span of such functions under every operator: each level is
  modelled by that pattern alone (Laplacian taps from the Chelikowsky table,
  transfer weights of the K=2 restriction / interpolation) and the model's
  prediction is compared to the oracle, which pins the phases handed to
  every call and the level bookkeeping.
"""
import numpy as np
import pytest

import oracle as O

TABLE = {1: [-2.0, 1.0], 3: [-49 / 18, 3 / 2, -3 / 20, 1 / 90]}


def _pc(N, h, dtype, pbc=True, nn0=3, kernels='port'):
    grid = O.Grid(N, np.array(N) * h, pbc)
    kin0 = O.laplace(grid, -0.5, nn0, dtype, kernels)
    return grid, O.Preconditioner(grid, kin0, dtype, kernels)


@pytest.mark.parametrize('dtype', [float, complex])
def test_constant_residual(dtype):
    grid, pc = _pc((16, 8, 8), 0.25, dtype)
    c = 1.7
    res = np.full((2,) + tuple(grid.n_c), c, dtype)
    ph = grid.phase_cd([0, 0, 0]) if dtype == complex else None
    out = pc(res, ph)
    expect = -41 * pc.step * c
    assert np.abs(out - expect).max() < 1e-12 * abs(expect)


def test_checkerboard_residual_is_exact():
    grid, pc = _pc((16, 8, 12), 0.3, float)
    i = np.indices(tuple(grid.n_c)).sum(0)
    res = np.where(i % 2 == 0, 1.25, -1.25)[None].copy()
    out = pc(res)
    assert np.array_equal(out, -pc.step * res)


# --------------------------------------------------------------- plane waves
def _lap(theta, a, taps):
    M = len(a)
    out = np.zeros(M, complex)
    for p in range(M):
        for k, c in enumerate(taps, 1):
            out[p] += c * (np.exp(1j * theta * k) * a[(p + k) % M] +
                           np.exp(-1j * theta * k) * a[(p - k) % M] - 2 * a[p])
    return out


def _restrict(theta, a):
    M = len(a)
    b = np.array([0.5 * (a[2 * p] + 0.5 * (np.exp(1j * theta) * a[(2 * p + 1) % M] +
                                          np.exp(-1j * theta) * a[(2 * p - 1) % M]))
                  for p in range(M // 2)])
    return 2 * theta, b


def _interp(theta_c, b):
    theta = theta_c / 2
    M = len(b)
    a = np.zeros(2 * M, complex)
    for p in range(M):
        a[2 * p] = b[p]
        a[2 * p + 1] = 0.5 * (np.exp(-1j * theta) * b[p] +
                              np.exp(1j * theta) * b[(p + 1) % M])
    return theta, a


def _model(theta0, a0, h, step, nn0):
    taps0 = [-0.5 * w / h ** 2 for w in TABLE[nn0][1:]]
    taps1 = [-0.5 / (2 * h) ** 2]
    taps2 = [-0.5 / (4 * h) ** 2]
    t1, r1 = _restrict(theta0, -a0)
    d1 = 4 * step * r1
    q1 = _lap(t1, d1, taps1) - r1
    t2, r2 = _restrict(t1, q1)
    d2 = 16 * step * r2
    q2 = _lap(t2, d2, taps2) - r2
    d2 = d2 - 16 * step * q2
    _, q1 = _interp(t2, d2)
    d1 = d1 - q1
    q1 = _lap(t1, d1, taps1) - r1
    d1 = d1 - 4 * step * q1
    _, d0 = _interp(t1, -d1)
    q0 = _lap(theta0, d0, taps0) - a0
    return -(d0 - step * q0)


@pytest.mark.parametrize('k0,nn0', [(0.3, 3), (-0.15, 1), (0.0, 3)])
def test_bloch_plane_wave(k0, nn0):
    N, h = (16, 8, 8), 0.25
    grid, pc = _pc(N, h, complex, nn0=nn0)
    theta = 2 * np.pi * k0 / N[0]
    a0 = np.array([1.0, -0.5 + 0.25j, 2.0, 0.75j])
    i0 = np.arange(N[0])
    line = np.exp(1j * theta * i0) * a0[i0 % 4]
    res = np.ascontiguousarray(np.broadcast_to(line[:, None, None], N))[None]
    out = pc(res, grid.phase_cd([k0, 0.0, 0.0]))
    pred = _model(theta, a0, h, pc.step, nn0)
    expect = np.exp(1j * theta * i0) * pred[i0 % 4]
    err = np.abs(out[0] - expect[:, None, None]).max()
    assert err < 1e-11 * np.abs(expect).max()


def test_model_catches_a_dropped_term():
    """The model is sensitive to the sequence: dropping the level-2 update
    changes the prediction well beyond the tolerance used above."""
    theta, a0, h = 0.2, np.array([1.0, 0.5, -1.0, 2.0]), 0.25
    step = 0.66666666 / (0.5 * 49 / 18 * 3 / h ** 2)
    full = _model(theta, a0, h, step, 3)
    t1, r1 = _restrict(theta, -a0)
    d1 = 4 * step * r1
    q1 = _lap(t1, d1, [-0.5 / (2 * h) ** 2]) - r1
    d1 = d1 - 4 * step * q1
    _, d0 = _interp(t1, -d1)
    q0 = _lap(theta, d0, [-0.5 * w / h ** 2 for w in TABLE[3][1:]]) - a0
    short = -(d0 - step * q0)
    assert np.abs(full - short).max() > 1e-6 * np.abs(full).max()


@pytest.mark.skipif(not O.Kernels.available('reference'),
                    reason='reference kernels not built')
@pytest.mark.parametrize('dtype,pbc', [(float, True), (complex, True),
                                       (float, False)])
def test_port_equals_reference_kernels(dtype, pbc):
    """The same sequence over the reference's own compiled kernels gives
    identical bits."""
    rng = np.random.default_rng(3)
    N = (16, 12, 8)
    grid, pc = _pc(N, 0.3, dtype, pbc)
    _, pr = _pc(N, 0.3, dtype, pbc, kernels='reference')
    shape = (3,) + tuple(grid.n_c)
    res = rng.standard_normal(shape)
    if dtype == complex:
        res = res + 1j * rng.standard_normal(shape)
    ph = grid.phase_cd([0.1, -0.2, 0.3]) if dtype == complex else None
    assert np.array_equal(pc(res, ph), pr(res, ph))

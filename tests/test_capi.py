"""The C-ABI library: loads, exports every declared symbol, and its host-side
plan validation behaves like the reference objects (no GPU needed)."""
import ctypes
import os
import re

import numpy as np
import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'gpaw_amd.h')


def _declared():
    text = open(HEADER).read()
    text = re.sub(r'/\*.*?\*/', '', text, flags=re.S)
    return sorted(set(re.findall(r'\b(gpaw_[a-z0-9_]+)\s*\(', text)))


def test_library_exports_header_symbols():
    from gpaw_amd import lib
    L = lib()
    names = _declared()
    assert len(names) >= 15
    for n in names:
        assert hasattr(L, n), n


def test_python_binding_covers_header():
    from gpaw_amd._lib import PROTOTYPES
    assert set(_declared()) == set(PROTOTYPES)


def test_operator_plan_validation():
    from gpaw_amd import GpawAmdError
    from gpaw_amd.ops import OperatorHandle
    # offset outside the declared range
    with pytest.raises(GpawAmdError):
        OperatorHandle([1.0, 2.0], [0, 10 ** 6], [8, 8, 8], 1,
                       np.zeros((3, 2), int), True)
    # MPI neighbour ranks are not a single domain
    with pytest.raises(GpawAmdError):
        OperatorHandle([1.0], [0], [8, 8, 8], 1, np.full((3, 2), 3), True)
    # periodic axis narrower than the stencil
    with pytest.raises(GpawAmdError):
        OperatorHandle([1.0, 1.0], [0, 1], [8, 8, 2], 3,
                       np.zeros((3, 2), int), True)


def test_operator_star_detection_and_diagonal():
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    gd = GridDescriptor((16, 16, 16), (4.0, 4.0, 4.0))
    lap = Laplace(gd, 1.0, 3)     # a star stencil needs no device memory
    assert abs(lap.get_diagonal_element() - lap.coef_p[0]) == 0.0
    assert lap.npoints == 19


def test_transformer_plan_validation():
    from gpaw_amd import GpawAmdError
    from gpaw_amd.ops import TransformerHandle
    nb = np.zeros((3, 2), int)
    pad = np.array([[5, 4]] * 3)
    npad = np.array([[4, 5]] * 3)
    skip = np.zeros((3, 2), int)
    TransformerHandle([16] * 3, [8] * 3, 6, pad, npad, skip, nb, True)
    with pytest.raises(GpawAmdError):   # wrong output size
        TransformerHandle([16] * 3, [9] * 3, 6, pad, npad, skip, nb, True)
    with pytest.raises(GpawAmdError):   # k must be 2, 4, 6 or 8
        TransformerHandle([16] * 3, [8] * 3, 5, pad, npad, skip, nb, True)
    with pytest.raises(GpawAmdError):   # asymmetric exchange is not self-wrap
        TransformerHandle([16] * 3, [8] * 3, 6, pad, pad, skip, nb, True)


def test_transformer_pads_match_reference_formulas():
    """gpaw/transformers.py:30-45 for the configurations used."""
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.transformers import Transformer
    gd = GridDescriptor((64,) * 3, (12.0,) * 3, True)
    r = Transformer(gd, gd.coarsen(), 3)
    assert (r.pad_cd == [[5, 4]] * 3).all() and (r.neighborpad_cd == [[4, 5]] * 3).all()
    i = Transformer(gd.coarsen(), gd, 3)
    assert (i.pad_cd == [[2, 3]] * 3).all() and (i.neighborpad_cd == [[3, 2]] * 3).all()
    gz = GridDescriptor((96,) * 3, (18.0,) * 3, False)
    r = Transformer(gz, gz.coarsen(), 1)
    assert (r.pad_cd == 0).all()
    i = Transformer(gz.coarsen(), gz, 1)
    assert (i.pad_cd == 1).all() and (i.skip_cd == [[1, 0]] * 3).all()


def test_coefficients_match_oracle():
    """The product's stencil builders give the oracle's coefficients."""
    import oracle as O
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Gradient, Laplace, LaplaceA
    for cell in ([3.0, 3.3, 3.1], [(3.0, 0, 0), (1.5, 2.6, 0), (0, 0, 3.3)]):
        gd = GridDescriptor((12, 12, 12), cell)
        og = O.Grid((12, 12, 12), cell)
        for n in (1, 2, 3):
            lap = Laplace(gd, -0.5, n, allocate=False)
            c, o = O.laplace_stencil(og, -0.5, n)
            assert np.array_equal(lap.coef_p, c)
            assert np.array_equal(lap.offset_pc, o)
            g = Gradient(gd, 1, 1.0, n, allocate=False)
            c, o = O.gradient_stencil(og, 1, 1.0, n)
            assert np.array_equal(g.coef_p, c) and np.array_equal(g.offset_pc, o)
    gd = GridDescriptor((12, 12, 12), [3.0, 3.3, 3.1])
    og = O.Grid((12, 12, 12), [3.0, 3.3, 3.1])
    a = LaplaceA(gd, 0.1, allocate=False)
    c, o = O.mehrstellen_a(og, 0.1)
    assert np.array_equal(a.coef_p, c)


def test_madelung_matches_oracle_and_constants():
    """The product's Ewald sum (its own summation to a tolerance) against the
    oracle's fixed-truncation restatement on orthorhombic, hexagonal and
    fcc cells; against the published Madelung constants of the Wigner
    lattices (alpha_M / r_s, r_s the Wigner-Seitz radius: sc 1.7601189,
    bcc 1.7918585, fcc 1.7917472); and independent of the Ewald
    splitting parameter."""
    import oracle as O
    from gpaw_amd.utilities import madelung
    for cell in (np.eye(3) * 5.0, np.diag([4.0, 5.0, 6.0]),
                 [[3.0, 0, 0], [1.5, 2.6, 0], [0, 0, 3.3]],
                 [[0, 2.0, 2.0], [2.0, 0, 2.0], [2.0, 2.0, 0]]):
        assert abs(madelung(cell) - O.madelung(cell)) < 1e-12
    assert abs(madelung(np.eye(3) * 2.0) * 2.0 - 2.837297479480) < 1e-9
    lattices = [(np.eye(3), 1.7601189),
                (0.5 * np.array([[-1, 1, 1], [1, -1, 1], [1, 1, -1.0]]), 1.7918585),
                (0.5 * np.array([[0, 1, 1], [1, 0, 1], [1, 1, 0.0]]), 1.7917472)]
    for cell, aM in lattices:
        for a in (1.0, 4.5):
            c = a * cell
            rs = (3 * abs(np.linalg.det(c)) / (4 * np.pi)) ** (1 / 3)
            assert abs(madelung(c) * rs - aM) < 1e-7
            assert abs(madelung(c, alpha=0.7 / a) - madelung(c, alpha=3.0 / a)) < 1e-12 / a


def test_poisson_hierarchy_shape():
    """gpaw/poisson.py:64-89,111-113 on the C2 grid."""
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.poisson import PoissonSolver
    gd = GridDescriptor((96,) * 3, (96 * 0.2 / 0.52917720859,) * 3, False)
    ps = PoissonSolver(nn=3)
    ps.set_grid_descriptor(gd)
    assert ps.levels == 4
    assert [tuple(o.shape) for o in ps.operators] == \
        [(95,) * 3, (47,) * 3, (23,) * 3, (11,) * 3, (5,) * 3]
    assert ps.presmooths == [2, 4, 4, 4, 4] and ps.postsmooths == [1, 4, 4, 4, 4]
    assert ps.weights == [2.0 / 3.0, 1.0, 1.0, 1.0, 1.0]
    assert [o.npoints for o in ps.operators] == [19, 7, 7, 7, 7]


def test_preconditioner_plan_and_workspace():
    """Levels chain, the step is the reference's, the workspace is one fine
    and six coarse grid functions per band, and bad calls fail before any
    device work."""
    import ctypes
    from gpaw_amd import GpawAmdError, lib
    from gpaw_amd._lib import check, GPAW_C128, GPAW_F32
    from gpaw_amd.grid_descriptor import GridDescriptor
    from gpaw_amd.fd_operators import Laplace
    from gpaw_amd.preconditioner import Preconditioner
    gd = GridDescriptor((32, 24, 16), (6.0, 4.5, 3.0), True)
    kin = Laplace(gd, -0.5, 3, complex)
    pc = Preconditioner(gd, kin, complex, block=4)
    pc.allocate()
    h = pc._plan()
    L = lib()
    assert L.gpaw_preconditioner_step(h) == 0.66666666 / kin.get_diagonal_element()
    nb = ctypes.c_long()
    check(L.gpaw_preconditioner_workspace(h, 3, GPAW_C128, ctypes.byref(nb)))

    def sl(n):
        return (n * 3 * 16 + 255) // 256 * 256
    assert nb.value == sl(32 * 24 * 16) + 3 * sl(16 * 12 * 8) + 3 * sl(8 * 6 * 4)
    with pytest.raises(GpawAmdError):    # FP32 data
        check(L.gpaw_preconditioner_workspace(h, 3, GPAW_F32, ctypes.byref(nb)))
    with pytest.raises(GpawAmdError):    # NULL arrays
        check(L.gpaw_preconditioner_apply(h, None, None, 1, GPAW_C128, None,
                                          None, 0, None))
    # levels that do not chain
    other = Laplace(GridDescriptor((16, 24, 16), (3.0, 4.5, 3.0)), -0.5, 1,
                    complex)
    from gpaw_amd.preconditioner import _handle
    out = ctypes.c_void_p()
    rc = L.gpaw_preconditioner_new(
        _handle(kin), _handle(other), _handle(pc.kin2),
        _handle(pc.restrictor_object0), _handle(pc.restrictor_object1),
        _handle(pc.interpolator_object2), _handle(pc.interpolator_object1),
        ctypes.byref(out))
    assert rc != 0 and b'chain' in L.gpaw_last_error()

"""TEST INFRASTRUCTURE ONLY: CPU restatement of GPAW's FD / multigrid hot path.

The numerical kernels live in ``bmgs_oracle.c`` (built to ``liboracle.so``);
``Kernels('reference')`` instead loads ``_ref/libgpawref.so``, the reference's
own ``c/bmgs/bmgs.c`` + ``c/bc.c`` compiled in place, with identical entry
points.  This module restates the Python layers that drive them:

* grid geometry           gpaw/grid_descriptor.py:65-150,197-237,494-501,
                          gpaw/domain.py:15-150
* stencil coefficients    gpaw/fd_operators.py:20-66 (FDOperator),
                          :123-144 (Gradient), :192-229 (LaplaceA/B),
                          :231-292 (NewGUCLaplace and its table),
                          :317-367 (GUCLaplace; its mixed-derivative
                          weights derived here, not tabulated)
* transformer paddings    gpaw/transformers.py:19-62
* multigrid Poisson       gpaw/poisson.py:21-276
* charged-system helpers  gpaw/utilities/gauss.py:47-83,
                          gpaw/utilities/tools.py:92-124,
                          gpaw/utilities/ewald.py:12-150
* preconditioner          gpaw/preconditioner.py:13-70
* callers of the path     gpaw/wavefunctions/fd.py:50-52,
                          gpaw/hamiltonian.py:254-259,371-398,
                          gpaw/density.py:217-234, gpaw/xc/hybrid.py:148-152,
                          gpaw/lrtddft/omega_matrix.py:383-430

It is imported only by tests, the smoke check and the CPU-baseline leg of
the benchmark.
"""
import ctypes
import math
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

# CODATA-era Bohr radius used by the reference's ASE dependency (ase.units)
bohr_per_angstrom = 1.0 / 0.52917720859

_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)


def _d(a):
    return a.ctypes.data_as(_dp) if a is not None else None


def _l(a):
    return a.ctypes.data_as(_lp)


class Kernels:
    """ctypes front end to the oracle kernels.

    kind='port'          our C restatement (liboracle.so)
    kind='reference'     the reference's own C, built in place (_ref/libgpawref.so)
    kind='reference_mt'  the same with its GPAW_OMP_MONLY threads
                         (_ref/libgpawref_mt.so; OMP_NUM_THREADS threads per
                         bmgs_fd / restrict / interpolate call)
    """

    _FILES = {'port': ('liboracle.so', 'orc_'),
              'reference': (os.path.join('_ref', 'libgpawref.so'), 'ref_'),
              'reference_mt': (os.path.join('_ref', 'libgpawref_mt.so'), 'ref_')}

    def __init__(self, kind='port'):
        if kind not in self._FILES:
            raise ValueError(kind)
        path, prefix = os.path.join(_HERE, self._FILES[kind][0]), self._FILES[kind][1]
        if not os.path.exists(path):
            raise FileNotFoundError(
                '%s not built (run make -C oracle)' % path)
        self.kind = kind
        lib = ctypes.CDLL(path)
        self._apply = getattr(lib, prefix + 'operator_apply')
        self._apply.argtypes = [_dp, _lp, ctypes.c_int, _lp, ctypes.c_int,
                                _lp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                ctypes.c_int, _dp]
        self._relax = getattr(lib, prefix + 'operator_relax')
        self._relax.argtypes = [_dp, _lp, ctypes.c_int, _lp, ctypes.c_int,
                                _lp, ctypes.c_int, ctypes.c_int, _dp, _dp,
                                ctypes.c_int, ctypes.c_double]
        self._trans = getattr(lib, prefix + 'transformer_apply')
        self._trans.argtypes = [_lp, _lp, ctypes.c_int, _lp, _lp, _lp, _lp,
                                ctypes.c_int, ctypes.c_int, _dp, _dp,
                                ctypes.c_int, _dp]
        self._lib = lib

    @staticmethod
    def available(kind):
        return os.path.exists(os.path.join(_HERE, Kernels._FILES[kind][0]))

    def operator_apply(self, coefs, offsets, size, rng, nb, real, cfd,
                       a, b, nin, phases):
        ph = None
        if not real:
            ph = np.ascontiguousarray(phases, complex).view(float)
        rc = self._apply(_d(coefs), _l(offsets), len(coefs), _l(size), rng,
                         _l(nb), int(real), int(cfd), _d(a.view(float)),
                         _d(b.view(float)), nin, _d(ph))
        if rc:
            raise RuntimeError('oracle operator_apply failed (%d)' % rc)

    def operator_relax(self, coefs, offsets, size, rng, nb, cfd, method,
                       f, s, nrelax, w):
        rc = self._relax(_d(coefs), _l(offsets), len(coefs), _l(size), rng,
                         _l(nb), int(cfd), int(method), _d(f), _d(s),
                         int(nrelax), float(w))
        if rc:
            raise RuntimeError('oracle operator_relax failed (%d)' % rc)

    def transformer_apply(self, size_in, size_out, k, pad, npad, skip, nb,
                          real, interpolate, a, b, nin, phases):
        ph = None
        if not real:
            ph = np.ascontiguousarray(phases, complex).view(float)
        rc = self._trans(_l(size_in), _l(size_out), int(k), _l(pad), _l(npad),
                         _l(skip), _l(nb), int(real), int(interpolate),
                         _d(a.view(float)), _d(b.view(float)), nin, _d(ph))
        if rc:
            raise RuntimeError('oracle transformer_apply failed (%d)' % rc)


_default = {}


def _kernels(k):
    if isinstance(k, Kernels):
        return k
    k = k or 'port'
    if k not in _default:
        _default[k] = Kernels(k)
    return _default[k]


# ---------------------------------------------------------------------------
# Geometry: one domain of gpaw/grid_descriptor.py:65-150 + gpaw/domain.py
# ---------------------------------------------------------------------------
class Grid:
    def __init__(self, N_c, cell_cv=(1, 1, 1), pbc_c=True):
        if isinstance(pbc_c, (bool, int, np.bool_)):
            pbc_c = (bool(pbc_c),) * 3
        cell = np.array(cell_cv, float)
        if cell.ndim == 1:
            cell = np.diag(cell)
        self.cell_cv = cell
        self.pbc_c = np.asarray(pbc_c, bool)
        self.N_c = np.array(N_c, int)
        self.icell_cv = np.linalg.inv(cell).T
        cell_c = (cell ** 2).sum(1) ** 0.5
        ucell = np.array([cell[c] / cell_c[c] for c in range(3)])
        self.xxxiucell_cv = np.linalg.inv(ucell.T)
        # non-periodic axes drop global point 0 (domain.py / grid_descriptor.py:116-130)
        self.beg_c = np.where(self.pbc_c, 0, 1)
        self.end_c = self.N_c.copy()
        if (self.end_c - self.beg_c <= 0).any():
            raise ValueError('Grid too small!')
        self.n_c = self.end_c - self.beg_c
        self.h_cv = cell / self.N_c[:, None]
        self.dv = abs(np.linalg.det(cell)) / self.N_c.prod()
        self.orthogonal = not (cell - np.diag(cell.diagonal())).any()
        L_c = (np.linalg.inv(cell) ** 2).sum(0) ** -0.5
        h_c = L_c / self.N_c
        if max(h_c) / min(h_c) > 1.3:
            raise ValueError('Very anisotropic grid spacings: %s' % h_c)
        # single domain: the neighbour is ourselves (0) or nobody (-1)
        self.neighbor_cd = np.where(self.pbc_c[:, None], 0, -1) * np.ones(
            (3, 2), int)
        self.sdisp_cd = np.where(self.pbc_c[:, None], [[1, -1]], 0) * np.ones(
            (3, 2), int)

    def coarsen(self):
        if (self.N_c % 2).any():
            raise ValueError('Grid %s not divisible by 2!' % self.N_c)
        return Grid(self.N_c // 2, self.cell_cv, self.pbc_c)

    def refine(self):
        return Grid(self.N_c * 2, self.cell_cv, self.pbc_c)

    def zeros(self, n=(), dtype=float):
        if isinstance(n, int):
            n = (n,)
        return np.zeros(tuple(n) + tuple(self.n_c), dtype)

    def integrate(self, a):
        if a.ndim == 3:
            return a.sum() * self.dv
        return a.reshape(a.shape[:-3] + (-1,)).sum(-1) * self.dv

    def size_of_global_array(self):
        return self.N_c - 1 + self.pbc_c

    def grid_point_coordinates(self):
        idx = np.indices(self.n_c, float).T + self.beg_c
        return np.dot(idx, self.h_cv).T.copy()

    def phase_cd(self, k_c):
        """phase_cd of gpaw/kpt_descriptor.py:162-166 for scaled k-point k_c."""
        return np.exp(2j * np.pi * self.sdisp_cd *
                      np.asarray(k_c, float)[:, None])


# ---------------------------------------------------------------------------
# Stencil coefficients
# ---------------------------------------------------------------------------
# J. R. Chelikowsky et al., Phys. Rev. B 50, 11355 (1994): central-difference
# second-derivative weights of order 2n (table of gpaw/fd_operators.py:286-292).
LAPLACE_TABLE = [
    [0.0],
    [-2.0, 1.0],
    [-5 / 2, 4 / 3, -1 / 12],
    [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
    [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
    [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
    [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
     -1 / 16632]]


def laplace_stencil(grid, scale=1.0, n=1):
    """NewGUCLaplace coefficients and offsets (gpaw/fd_operators.py:231-277).

    The 13 half-neighbour directions are ordered by |M.u|^2 with a stable
    sort (the reference's numpy sorted 13 elements by insertion sort); up
    to six directions are added until the second-moment equations solve.
    """
    M_ic = np.indices((3, 3, 3)).reshape((3, -1)).T[-13:] - 1
    u_cv = grid.h_cv / (grid.h_cv ** 2).sum(1)[:, None] ** 0.5
    u2_i = (np.dot(M_ic, u_cv) ** 2).sum(1)
    i_d = u2_i.argsort(kind='stable')
    m_mv = np.array([(2, 0, 0), (0, 2, 0), (0, 0, 2),
                     (0, 1, 1), (1, 0, 1), (1, 1, 0)])
    for D in range(3, 7):
        h_dv = np.dot(M_ic[i_d[:D]], grid.h_cv)
        A_md = (h_dv ** m_mv[:, None, :]).prod(2)
        a_d, residual, rank, _ = np.linalg.lstsq(A_md, [1, 1, 1, 0, 0, 0],
                                                 rcond=-1)
        if residual.sum() < 1e-14:
            assert rank == D
            break
    a_d = a_d * scale
    offsets = [(0, 0, 0)]
    coefs = [LAPLACE_TABLE[n][0] * a_d.sum()]
    for d in range(D):
        M_c = M_ic[i_d[d]]
        for sgn in (1, -1):
            for k in range(1, n + 1):
                offsets.append(tuple(int(x) for x in sgn * k * M_c))
                coefs.append(a_d[d] * LAPLACE_TABLE[n][k])
    return np.array(coefs), np.array(offsets, int)


def gradient_stencil(grid, v, scale=1.0, n=1):
    """Gradient along Cartesian v (gpaw/fd_operators.py:123-144)."""
    h = (grid.h_cv ** 2).sum(1) ** 0.5
    d = grid.xxxiucell_cv[:, v]
    A = np.zeros((2 * n + 1, 2 * n + 1))
    for i, io in enumerate(range(-n, n + 1)):
        for j in range(2 * n + 1):
            A[i, j] = io ** j / float(math.factorial(j))
    A[n, 0] = 1.0
    c = np.linalg.inv(A)[1]
    c = np.delete(c, len(c) // 2)
    offs = np.delete(np.arange(-n, n + 1), n)
    coefs, offsets = [], []
    for i in range(3):
        if abs(d[i]) > 1e-11:
            coefs.extend(list(c * d[i] / h[i] * scale))
            o = np.zeros((2 * n, 3), int)
            o[:, i] = offs
            offsets.extend(o)
    return np.array(coefs), np.array(offsets, int)


def _first_derivative_weights(n):
    """Central first-derivative weights w_1..w_n of order 2n as exact
    rationals: f'(0) ~ sum_k w_k (f(k) - f(-k)).  Solves the odd-moment
    conditions sum_k 2 w_k k^(2j-1) = [j == 1] for j = 1..n."""
    from fractions import Fraction
    A = [[Fraction(2 * k ** (2 * j - 1)) for k in range(1, n + 1)]
         for j in range(1, n + 1)]
    b = [Fraction(int(j == 1)) for j in range(1, n + 1)]
    # Gauss-Jordan elimination over the rationals
    for c in range(n):
        p = next(r for r in range(c, n) if A[r][c] != 0)
        A[c], A[p] = A[p], A[c]
        b[c], b[p] = b[p], b[c]
        for r in range(n):
            if r != c and A[r][c] != 0:
                f = A[r][c] / A[c][c]
                A[r] = [x - f * y for x, y in zip(A[r], A[c])]
                b[r] -= f * b[c]
    return [b[k] / A[k][k] for k in range(n)]


def guc_stencil(grid, scale=1.0, n=1):
    """GUCLaplace coefficients and offsets (gpaw/fd_operators.py:317-367).

    In a cell with unit axis vectors u_c and dual vectors d_c (d_c . u_c' =
    delta), nabla^2 = sum_cc' (d_c . d_c') d_s_c d_s_c', s_c the length along
    axis c.  Pure terms use the order-2n second-difference table; each mixed
    pair (c, c+1 mod 3) counts twice and uses the product of two order-2n
    first differences, whose weights (1,1),(1,2),..,(n,n) are derived here
    rather than tabulated.  Terms and offsets follow the reference's order,
    which fixes the summation order of the kernel.
    """
    n = int(n)
    w = _first_derivative_weights(n)
    mixed = [float(w[i] * w[j]) for i in range(n) for j in range(i, n)]
    hlen = np.sqrt((grid.h_cv ** 2).sum(1))
    dual = grid.xxxiucell_cv
    g_diag = (dual ** 2).sum(1)
    offsets = [(0, 0, 0)]
    coefs = [scale * np.sum(g_diag * np.divide(LAPLACE_TABLE[n][0], hlen ** 2))]
    for k in range(1, n + 1):
        ck = scale * g_diag * np.divide(LAPLACE_TABLE[n][k], hlen ** 2)
        for ax in range(3):
            for sgn in (-1, 1):
                o = [0, 0, 0]
                o[ax] = sgn * k
                offsets.append(tuple(o))
                coefs.append(ck[ax])
    idx = 0
    for i in range(1, n + 1):
        for j in range(i, n + 1):
            for ax in range(3):
                nx = (ax + 1) % 3
                cval = (scale * 2.0 * mixed[idx] * np.dot(dual[ax], dual[nx]) /
                        (hlen[ax] * hlen[nx]))
                if abs(cval) <= 1e-11:
                    continue
                # the four corners (+i,+j), (-i,+j), (+i,-j), (-i,-j) of the
                # pair (ax, nx); the z-x pair stores its first step on x
                for (si, sj), sg in (((1, 1), 1), ((-1, 1), -1), ((1, -1), -1),
                                     ((-1, -1), 1)):
                    o = [0, 0, 0]
                    if ax == 2:
                        o[2], o[0] = si * i, sj * j
                    else:
                        o[ax], o[nx] = si * i, sj * j
                    offsets.append(tuple(o))
                    coefs.append(sg * cval)
                if j > i:
                    # the mirror pair (j, i) on the same two axes
                    for (si, sj), sg in (((1, 1), 1), ((-1, 1), -1),
                                         ((1, -1), -1), ((-1, -1), 1)):
                        o = [0, 0, 0]
                        if ax == 2:
                            o[0], o[2] = si * i, sj * j
                        else:
                            o[nx], o[ax] = si * i, sj * j
                        offsets.append(tuple(o))
                        coefs.append(sg * cval)
            idx += 1
    return np.array(coefs), np.array(offsets, int)


def mehrstellen_a(grid, scale):
    """LaplaceA (gpaw/fd_operators.py:192-217)."""
    c = np.divide(-1 / 12, grid.h_cv.diagonal() ** 2) * scale
    c0, c1, c2 = c[1] + c[2], c[0] + c[2], c[1] + c[0]
    a = -16.0 * np.sum(c)
    b = 10.0 * c + 0.125 * a
    coefs = [a, b[0], b[0], b[1], b[1], b[2], b[2],
             c0, c0, c0, c0, c1, c1, c1, c1, c2, c2, c2, c2]
    offsets = [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
               (0, 0, -1), (0, 0, 1),
               (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
               (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
               (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)]
    return np.array(coefs), np.array(offsets, int)


def mehrstellen_b():
    """LaplaceB (gpaw/fd_operators.py:219-229)."""
    a, b = 0.5, 1.0 / 12.0
    return (np.array([a, b, b, b, b, b, b]),
            np.array([(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
                      (0, 0, -1), (0, 0, 1)], int))


def flatten_offsets(offset_pc, n_c):
    """Padded-buffer flat offsets, range and cfd flag
    (gpaw/fd_operators.py:26-42)."""
    offset_pc = np.asarray(offset_pc, int)
    cfd = not any((o != 0).sum() >= 2 for o in offset_pc)
    mp = int(max(offset_pc.max(0)))
    M_c = np.asarray(n_c) + 2 * mp
    stride_c = np.array([M_c[1] * M_c[2], M_c[2], 1])
    return np.dot(offset_pc, stride_c).astype(np.int64), mp, cfd


class Operator:
    """FDOperator driving the oracle kernels (gpaw/fd_operators.py:20-90)."""

    def __init__(self, coef_p, offset_pc, grid, dtype=float, kernels=None):
        self.coefs = np.ascontiguousarray(coef_p, float)
        self.offsets, self.mp, self.cfd = flatten_offsets(offset_pc, grid.n_c)
        self.grid = grid
        self.dtype = dtype
        self.shape = tuple(grid.n_c)
        self.size = np.array(grid.n_c, np.int64)
        self.nb = np.ascontiguousarray(grid.neighbor_cd, np.int64)
        self.k = _kernels(kernels)

    def apply(self, in_xg, out_xg, phase_cd=None):
        assert in_xg.shape == out_xg.shape and in_xg.shape[-3:] == self.shape
        assert in_xg.flags.c_contiguous and out_xg.flags.c_contiguous
        real = in_xg.dtype == np.float64
        nin = in_xg.shape[0] if in_xg.ndim == 4 else 1
        self.k.operator_apply(self.coefs, self.offsets, self.size, self.mp,
                              self.nb, real, self.cfd, in_xg, out_xg, nin,
                              phase_cd)

    def relax(self, method, f_g, s_g, n, w=1.0):
        assert f_g.dtype == np.float64 and f_g.flags.c_contiguous
        self.k.operator_relax(self.coefs, self.offsets, self.size, self.mp,
                              self.nb, self.cfd, method, f_g,
                              np.ascontiguousarray(s_g), n, w)

    def get_diagonal_element(self):
        d = 0.0
        for c, o in zip(self.coefs, self.offsets):
            if o == 0:
                d = c
        return d


def laplace(grid, scale=1.0, n=1, dtype=float, kernels=None):
    c, o = laplace_stencil(grid, scale, n)
    return Operator(c, o, grid, dtype, kernels)


def transformer_plan(gdin, gdout, nn):
    """Paddings / skips of gpaw/transformers.py:19-62."""
    pad = np.empty((3, 2), np.int64)
    npad = np.empty((3, 2), np.int64)
    skip = np.zeros((3, 2), np.int64)
    if (gdin.N_c == 2 * gdout.N_c).all():
        pad[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
        pad[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
        npad[:, 0] = 2 * nn - 2 + 2 * gdout.beg_c - gdin.beg_c
        npad[:, 1] = 2 * nn - 1 - 2 * gdout.end_c + gdin.end_c
        interpolate = False
    else:
        assert (gdout.N_c == 2 * gdin.N_c).all()
        pad[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
        pad[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
        npad[:, 0] = nn + gdout.beg_c // 2 - gdin.beg_c
        npad[:, 1] = nn - 1 - gdout.end_c // 2 + gdin.end_c
        skip[:, 0] = gdout.beg_c % 2
        skip[:, 1] = gdout.end_c % 2
        interpolate = True
        assert ((gdin.n_c[0] + 2 * nn - 1) * (gdin.n_c[1] + 2 * nn - 1) <=
                gdout.n_c[0] * gdout.n_c[1])
    assert (pad >= 0).all()
    return pad, npad, skip, interpolate


class Transformer:
    """_Transformer driving the oracle kernels (gpaw/transformers.py:19-82)."""

    def __init__(self, gdin, gdout, nn=1, dtype=float, kernels=None):
        assert 1 <= nn <= 4
        self.pad, self.npad, self.skip, self.interpolate = transformer_plan(
            gdin, gdout, nn)
        self.nn = nn
        self.size_in = np.array(gdin.n_c, np.int64)
        self.size_out = np.array(gdout.n_c, np.int64)
        self.nb = np.ascontiguousarray(gdin.neighbor_cd, np.int64)
        self.ngpin, self.ngpout = tuple(gdin.n_c), tuple(gdout.n_c)
        self.k = _kernels(kernels)

    def apply(self, a, b, phases=None):
        assert a.shape[-3:] == self.ngpin and b.shape[-3:] == self.ngpout
        real = a.dtype == np.float64
        nin = a.shape[0] if a.ndim == 4 else 1
        self.k.transformer_apply(self.size_in, self.size_out, 2 * self.nn,
                                 self.pad, self.npad, self.skip, self.nb,
                                 real, self.interpolate, a, b, nin, phases)


# ---------------------------------------------------------------------------
# Charged-system helpers
# ---------------------------------------------------------------------------
def _erf(x):
    return np.vectorize(math.erf, otypes=[float])(x)


def _coordinates(grid, tiny=1e-12):
    """gpaw/utilities/tools.py:92-124 (origin at the cell centre)."""
    r0 = 0.5 * grid.cell_cv.sum(0)
    r = grid.grid_point_coordinates() - r0[:, None, None, None]
    for c in range(3):
        if grid.pbc_c[c]:
            L = grid.cell_cv[c, c]
            r[c] = np.where(r[c] > 0.5 * L, r[c] - L, r[c])
            r[c] = np.where(r[c] < -0.5 * L, r[c] + L, r[c])
    r2 = np.sum(r ** 2, axis=0)
    return r, np.where(r2 < tiny, tiny, r2)


def gaussian_monopole(grid, a=19.0):
    """L=0 generalized Gaussian density and its potential
    (gpaw/utilities/gauss.py:61-83 with gauss_L[0], gausspot_L[0])."""
    _, r2 = _coordinates(grid)
    a = 4 * a * (grid.icell_cv ** 2).sum(1).max()
    g = np.sqrt(a ** 3 * 4) / np.pi * np.exp(-a * r2)
    r = np.sqrt(r2)
    v = 2.0 * 1.7724538509055159 * _erf(np.sqrt(a) * r) / r
    return g, v


def madelung(cell, G=5.0, Ng=(9, 9, 9), Nl=(3, 3, 3)):
    """Potential of a unit negative point charge in its own neutralising
    background at its own site, excluding its self term
    (gpaw/utilities/ewald.py:28-150 with r = r_B = 0, q_B = [-1])."""
    cell = np.asarray(cell, float)
    vol = abs(np.linalg.det(cell))
    recip = np.linalg.inv(cell).T
    G = G / np.sqrt((cell ** 2).sum(1).max())
    q = -1.0
    E_g = 0.0
    for i in range(-Ng[0], Ng[0] + 1):
        for j in range(-Ng[1], Ng[1] + 1):
            for k in range(-Ng[2], Ng[2] + 1):
                g_v = 2 * np.pi * (i * recip[0] + j * recip[1] + k * recip[2])
                g2 = np.dot(g_v, g_v)
                if g2 > 1e-10:
                    x = 0.25 * g2 / G ** 2
                    E_g += np.exp(-x) / x
    E_r = 0.0
    for i in range(-Nl[0], Nl[0] + 1):
        for j in range(-Nl[1], Nl[1] + 1):
            for k in range(-Nl[2], Nl[2] + 1):
                l_v = i * cell[0] + j * cell[1] + k * cell[2]
                lr = np.linalg.norm(l_v)
                if lr > 1e-5:
                    E_r += (1.0 - math.erf(G * lr)) / lr
    E0 = q * E_r + q * np.pi / (vol * G ** 2) * E_g
    E0 -= q * 2.0 * G / np.sqrt(np.pi)
    E0 += -q * np.pi / (G ** 2 * vol)
    return E0


def gaussian_charges(grid, centers, widths, charges):
    """Sum of normalised 3-D Gaussians (synthetic densities of the
    benchmark configurations); periodic axes use the minimum image."""
    r_vG = grid.grid_point_coordinates()
    rho = np.zeros(tuple(grid.n_c))
    L = np.diag(grid.cell_cv)
    for c, s, q in zip(centers, widths, charges):
        d = r_vG - np.asarray(c)[:, None, None, None]
        for ax in range(3):
            if grid.pbc_c[ax]:
                d[ax] -= L[ax] * np.round(d[ax] / L[ax])
        r2 = (d ** 2).sum(0)
        rho += q * (2 * np.pi * s * s) ** -1.5 * np.exp(-r2 / (2 * s * s))
    return rho


# ---------------------------------------------------------------------------
# Multigrid Poisson (gpaw/poisson.py:21-276)
# ---------------------------------------------------------------------------
class PoissonConvergenceError(RuntimeError):
    pass


class PoissonSolver:
    def __init__(self, nn=3, relax='J', eps=2e-10, kernels=None):
        self.nn = nn
        self.eps = eps
        self.relax_method = {'GS': 1, 'J': 2}[relax]
        self.maxiter = 1000
        self.k = _kernels(kernels)
        self.charged_periodic_correction = None

    def set_grid_descriptor(self, gd):
        self.gd = gd
        self.dv = gd.dv
        scale = -0.25 / np.pi
        if self.nn == 'M':
            c, o = mehrstellen_a(gd, -scale)
            self.operators = [Operator(c, o, gd, kernels=self.k)]
            cb, ob = mehrstellen_b()
            self.B = Operator(cb, ob, gd, kernels=self.k)
        else:
            self.operators = [laplace(gd, scale, self.nn, kernels=self.k)]
            self.B = None
        self.interpolators, self.restrictors = [], []
        self.presmooths, self.postsmooths = [2], [1]
        self.weights = [2.0 / 3.0]
        level = 0
        while level < 4:
            try:
                gd2 = gd.coarsen()
            except ValueError:
                break
            self.operators.append(laplace(gd2, scale, 1, kernels=self.k))
            self.interpolators.append(Transformer(gd2, gd, kernels=self.k))
            self.restrictors.append(Transformer(gd, gd2, kernels=self.k))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level += 1
            gd = gd2
        self.levels = level

    def initialize(self):
        gd = self.gd
        self.rhos = [gd.zeros()]
        self.phis = [None]
        self.residuals = [gd.zeros()]
        for _ in range(self.levels):
            gd = gd.coarsen()
            self.phis.append(gd.zeros())
            self.rhos.append(gd.zeros())
            self.residuals.append(gd.zeros())
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8

    def solve(self, phi, rho, charge=None, eps=None, maxcharge=1e-6,
              zero_initial_phi=False):
        if eps is None:
            eps = self.eps
        actual_charge = self.gd.integrate(rho)
        background = (actual_charge / self.gd.dv /
                      self.gd.size_of_global_array().prod())
        if charge is None:
            charge = actual_charge
        if abs(charge) <= maxcharge:
            return self.solve_neutral(phi, rho - background, eps=eps)
        if self.gd.pbc_c.all():
            if self.charged_periodic_correction is None:
                self.charged_periodic_correction = madelung(self.gd.cell_cv)
            if zero_initial_phi:
                phi[:] = 0.0
            else:
                phi -= charge * self.charged_periodic_correction
            it = self.solve_neutral(phi, rho - background, eps=eps)
            phi += charge * self.charged_periodic_correction
            return it
        if not self.gd.pbc_c.any():
            g, v = gaussian_monopole(self.gd)
            q = actual_charge / np.sqrt(4 * np.pi)
            rho_neutral = rho - q * g
            if zero_initial_phi:
                phi[:] = 0.0
            else:
                phi += -q * v
            it = self.solve_neutral(phi, rho_neutral, eps=eps)
            phi += q * v
            return it
        raise NotImplementedError

    def solve_neutral(self, phi, rho, eps=2e-10):
        self.phis[0] = phi
        if self.B is None:
            self.rhos[0][:] = rho
        else:
            self.B.apply(rho, self.rhos[0])
        niter = 1
        while self.iterate2(self.step) > eps and niter < self.maxiter:
            niter += 1
        if niter == self.maxiter:
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!' %
                self.maxiter)
        if self.gd.pbc_c.all():
            phi -= np.sum(phi.ravel()) / np.prod(
                self.gd.size_of_global_array())
        return niter

    def run_vcycles(self, phi, rho, count):
        """Exactly `count` V-cycles on the (already neutral) rho; returns the
        error after each cycle.  Used for fixed-cycle parity checks."""
        self.phis[0] = phi
        self.rhos[0][:] = rho
        return [self.iterate2(self.step) for _ in range(count)]

    def iterate2(self, step, level=0):
        residual = self.residuals[level]
        if level < self.levels:
            self.operators[level].relax(self.relax_method, self.phis[level],
                                        self.rhos[level],
                                        self.presmooths[level],
                                        self.weights[level])
            self.operators[level].apply(self.phis[level], residual)
            residual -= self.rhos[level]
            self.restrictors[level].apply(residual, self.rhos[level + 1])
            self.phis[level + 1][:] = 0.0
            self.iterate2(4.0 * step, level + 1)
            self.interpolators[level].apply(self.phis[level + 1], residual)
            self.phis[level] -= residual
        self.operators[level].relax(self.relax_method, self.phis[level],
                                    self.rhos[level],
                                    self.postsmooths[level],
                                    self.weights[level])
        if level == 0:
            self.operators[level].apply(self.phis[level], residual)
            residual -= self.rhos[level]
            return np.dot(residual.ravel(), residual.ravel()) * self.dv


# ---------------------------------------------------------------------------
# Eigensolver preconditioner
# ---------------------------------------------------------------------------
class Preconditioner:
    """Three-grid preconditioner of gpaw/preconditioner.py:13-70 over the
    oracle kernels, with numpy doing the element-wise updates.

    The reference's final BLAS axpy (y <- y + a x, c/blas.c:197-218) is
    evaluated here as y + (a x) with two roundings; a BLAS built with fused
    multiply-add may differ from it in the last bit of that one step.
    """

    def __init__(self, grid0, kin0, dtype=float, kernels=None):
        grid1 = grid0.coarsen()
        grid2 = grid1.coarsen()
        self.kin0 = kin0
        self.kin1 = laplace(grid1, -0.5, 1, dtype, kernels)
        self.kin2 = laplace(grid2, -0.5, 1, dtype, kernels)
        self.grids = (grid0, grid1, grid2)
        self.dtype = dtype
        self.step = 0.66666666 / kin0.get_diagonal_element()
        self.down = (Transformer(grid0, grid1, 1, dtype, kernels),
                     Transformer(grid1, grid2, 1, dtype, kernels))
        self.up = (Transformer(grid2, grid1, 1, dtype, kernels),
                   Transformer(grid1, grid0, 1, dtype, kernels))

    def __call__(self, residuals, phase_cd=None):
        """residuals: (nbands, n0, n1, n2); returns the preconditioned
        residuals as a new array."""
        s = self.step
        nb = len(residuals)
        dt = residuals.dtype
        shape1 = (nb,) + tuple(self.grids[1].n_c)
        shape2 = (nb,) + tuple(self.grids[2].n_c)
        # level 1: restrict the negated residuals
        r1 = np.empty(shape1, dt)
        self.down[0].apply(np.ascontiguousarray(-residuals), r1, phase_cd)
        d1 = (4 * s) * r1
        q1 = np.empty(shape1, dt)
        self.kin1.apply(d1, q1, phase_cd)
        q1 -= r1
        # level 2: one damped step on the restricted defect
        r2 = np.empty(shape2, dt)
        self.down[1].apply(q1, r2, phase_cd)
        d2 = (16 * s) * r2
        q2 = np.empty(shape2, dt)
        self.kin2.apply(d2, q2, phase_cd)
        q2 -= r2
        d2 -= (16 * s) * q2
        # back to level 1: correct, then one more damped step
        self.up[0].apply(d2, q1, phase_cd)
        d1 -= q1
        self.kin1.apply(d1, q1, phase_cd)
        q1 -= r1
        d1 -= (4 * s) * q1
        # level 0
        d0 = np.empty(residuals.shape, dt)
        self.up[1].apply(np.ascontiguousarray(-d1), d0, phase_cd)
        q0 = np.empty(residuals.shape, dt)
        self.kin0.apply(d0, q0, phase_cd)
        q0 -= residuals
        d0 += (-s) * q0
        d0 *= -1.0
        return d0


# ---------------------------------------------------------------------------
# Callers of the hot path (restated call sites)
# ---------------------------------------------------------------------------
def apply_pseudo_hamiltonian(kin, vt_G, psit_xG, Htpsit_xG, phase_cd=None):
    """FDWaveFunctions.apply_pseudo_hamiltonian (gpaw/wavefunctions/fd.py:50-52):
    kin.apply over the bands, then Hamiltonian.apply_local_potential
    (gpaw/hamiltonian.py:393-398), Htpsit_G += psit_G * vt_G band by band."""
    kin.apply(psit_xG, Htpsit_xG, phase_cd)
    if psit_xG.ndim == 3:
        Htpsit_xG += psit_xG * vt_G
    else:
        for psit_G, Htpsit_G in zip(psit_xG, Htpsit_xG):
            Htpsit_G += psit_G * vt_G


def interpolate_density(interpolator, finegd, nt_sG, nt_sg, charge=0.0,
                        comp_charge=0.0):
    """Density.interpolate (gpaw/density.py:217-234) for given charges."""
    for s in range(nt_sG.shape[0]):
        interpolator.apply(nt_sG[s], nt_sg[s])
    if not finegd.pbc_c.all():
        pseudo_charge = -(charge + comp_charge)
        if abs(pseudo_charge) > 1.0e-14:
            x = pseudo_charge / finegd.integrate(nt_sg).sum()
            nt_sg *= x


def restrict_potential(restrictor, vt_sg, vHt_g, vt_sG):
    """The restriction loop of Hamiltonian.update
    (gpaw/hamiltonian.py:254-259): vt_g += vHt_g; restrict(vt_g, vt_G)."""
    for vt_g, vt_G in zip(vt_sg, vt_sG):
        vt_g += vHt_g
        restrictor.apply(vt_g, vt_G)


def pair_potentials(poisson, rhot_pg, charges, eps=1e-12):
    """The exact-exchange pair loop's solves (gpaw/xc/hybrid.py:148-152)."""
    vt_pg = np.zeros_like(rhot_pg)
    niter = []
    for p in range(rhot_pg.shape[0]):
        vt_pg[p] = 0.0
        niter.append(poisson.solve(vt_pg[p], -rhot_pg[p], charge=charges[p],
                                   eps=eps, zero_initial_phi=True))
    return vt_pg, niter


def rpa_coupling(poisson, rhot_pg, gd=None, restrictor=None, rhot_pG=None):
    """The Hartree part of the RPA omega matrix (gpaw/lrtddft/omega_matrix.py:
    383-430), without the pair prefactors and the atomic corrections: for
    every pair density ij, phit = 0, poisson.solve(phit, rhot_ij,
    charge=None); with a restrictor (finegrid == 1) phit is restricted to the
    coarse grid and integrated against the coarse densities rhot_pG; then
    K[ij, kq] = gd.integrate(rhot_kq * phit) for kq >= ij and K[kq, ij] =
    K[ij, kq].  Returns K and the iteration counts."""
    P = rhot_pg.shape[0]
    if gd is None:
        gd = poisson.gd
    rhot_int = rhot_pg if restrictor is None else rhot_pG
    K = np.zeros((P, P))
    niter = []
    for ij in range(P):
        phit_p = np.zeros(rhot_pg.shape[1:])
        niter.append(poisson.solve(phit_p, rhot_pg[ij].copy(), charge=None))
        if restrictor is None:
            phit = phit_p
        else:
            phit = np.zeros(rhot_pG.shape[1:])
            restrictor.apply(phit_p, phit)
        for kq in range(ij, P):
            K[ij, kq] = gd.integrate(rhot_int[kq] * phit)
            K[kq, ij] = K[ij, kq]
    return K, niter

"""Finite-difference operators on the GPU (drop-in for gpaw/fd_operators.py).

Same classes, constructor arguments, stencil construction, coefficient
order and methods as the reference; the compiled object behind them is
ops.OperatorHandle over libgpaw_amd.so instead of _gpaw.Operator.

  FDOperator      gpaw/fd_operators.py:20-121
  Gradient        gpaw/fd_operators.py:123-144
  Laplace         gpaw/fd_operators.py:147-153
  LaplaceA/B      gpaw/fd_operators.py:192-229 (Mehrstellen)
  NewGUCLaplace   gpaw/fd_operators.py:231-282
  GUCLaplace      gpaw/fd_operators.py:317-367
"""
from math import factorial, pi

import numpy as np

from . import config
from .ops import OperatorHandle

# Central second-derivative weights of order 2n, J. R. Chelikowsky et al.,
# Phys. Rev. B 50, 11355 (1994).
laplace = [[0],
           [-2, 1],
           [-5 / 2, 4 / 3, -1 / 12],
           [-49 / 18, 3 / 2, -3 / 20, 1 / 90],
           [-205 / 72, 8 / 5, -1 / 5, 8 / 315, -1 / 560],
           [-5269 / 1800, 5 / 3, -5 / 21, 5 / 126, -5 / 1008, 1 / 3150],
           [-5369 / 1800, 12 / 7, -15 / 56, 10 / 189, -1 / 112, 2 / 1925,
            -1 / 16632]]

# Mixed second-derivative weights, (1,1),(1,2),...,(n,n) order.
cross = [[0],
         [1 / 4],
         [4 / 9, -1 / 18, 1 / 144],
         [9 / 16, -9 / 80, 1 / 80, 9 / 400, -1 / 400, 1 / 3600],
         [16 / 25, -4 / 25, 16 / 525, -1 / 350, 1 / 25, -4 / 525, 1 / 1400,
          16 / 11025, -1 / 7350, 1 / 78400],
         [25 / 36, -25 / 126, 25 / 504, -25 / 3024, 1 / 1512, 25 / 441,
          -25 / 1764, 25 / 10584, -1 / 5292, 25 / 7056, -25 / 42336,
          1 / 21168, 25 / 254016, -1 / 127008, 1 / 1587600],
         [36 / 49, -45 / 196, 10 / 147, -3 / 196, 6 / 2695, -1 / 6468,
          225 / 3136, -25 / 1176, 15 / 3136, -3 / 4312, 5 / 103488,
          25 / 3969, -5 / 3528, 1 / 4851, -5 / 349272, 1 / 3136, -1 / 21560,
          1 / 310464, 1 / 148225, -1 / 2134440, 1 / 30735936]]


class FDOperator:
    def __init__(self, coef_p, offset_pc, gd, dtype=float, allocate=True):
        """FDOperator(coefs, offsets, gd, dtype) -> FDOperator object."""
        offset_pc = np.asarray(offset_pc, int).reshape(-1, 3)
        # a central finite-difference stencil has no mixed offsets
        cfd = True
        for offset_c in offset_pc:
            if sum([offset != 0 for offset in offset_c]) >= 2:
                cfd = False
        maxoffset_c = [max([offset_c[c] for offset_c in offset_pc])
                       for c in range(3)]
        mp = max(maxoffset_c)
        n_c = gd.n_c
        M_c = n_c + 2 * mp
        stride_c = np.array([M_c[1] * M_c[2], M_c[2], 1])
        offset_p = np.dot(offset_pc, stride_c)
        coef_p = np.ascontiguousarray(coef_p, float)
        assert coef_p.ndim == 1
        assert coef_p.shape == offset_p.shape
        assert dtype in [float, complex]
        self.dtype = dtype
        self.shape = tuple(n_c)
        self.operator = None
        self.args = [coef_p, offset_p, n_c, mp, gd.neighbor_cd,
                     dtype == float, None, cfd]
        self.offset_pc = offset_pc
        self.coef_p = coef_p
        self.cfd = cfd
        self.mp = mp
        self.gd = gd
        self.npoints = len(coef_p)
        self.allocated = False
        if allocate:
            self.allocate()

    def __str__(self):
        return '<%d point finite-difference stencil>' % self.npoints

    def allocate(self):
        assert not self.is_allocated()
        self.operator = OperatorHandle(*self.args)
        self.args = None
        self.allocated = True

    def is_allocated(self):
        return self.allocated

    def apply(self, in_xg, out_xg, phase_cd=None):
        if config.debug:
            self._check_apply(in_xg, out_xg, phase_cd)
        self.operator.apply(in_xg, out_xg, phase_cd)

    def relax(self, relax_method, f_g, s_g, n, w=None):
        self.operator.relax(relax_method, f_g, s_g, n, w)

    def get_diagonal_element(self):
        return self.operator.get_diagonal_element()

    def get_async_sizes(self):
        return self.operator.get_async_sizes()

    def estimate_memory(self, mem):
        # no padded buffer: halos are produced inside the kernels
        mem.setsize(0)

    def _check_apply(self, in_xg, out_xg, phase_cd):
        """The debug assertions of gpaw/fd_operators.py:100-111."""
        assert tuple(in_xg.shape) == tuple(out_xg.shape)
        assert tuple(in_xg.shape[-3:]) == self.shape
        if self.dtype == complex:
            ph = np.asarray(phase_cd)
            assert ph.dtype == complex and ph.shape == (3, 2)


class Gradient(FDOperator):
    def __init__(self, gd, v, scale=1.0, n=1, dtype=float, allocate=True):
        h = (gd.h_cv ** 2).sum(1) ** 0.5
        d = gd.xxxiucell_cv[:, v]
        A = np.zeros((2 * n + 1, 2 * n + 1))
        for i, io in enumerate(range(-n, n + 1)):
            for j in range(2 * n + 1):
                A[i, j] = io ** j / float(factorial(j))
        A[n, 0] = 1.
        coefs = np.linalg.inv(A)[1]
        coefs = np.delete(coefs, len(coefs) // 2)
        offs = np.delete(np.arange(-n, n + 1), n)
        coef_p = []
        offset_pc = []
        for i in range(3):
            if abs(d[i]) > 1e-11:
                coef_p.extend(list(coefs * d[i] / h[i] * scale))
                offset = np.zeros((2 * n, 3), int)
                offset[:, i] = offs
                offset_pc.extend(offset)
        FDOperator.__init__(self, coef_p, offset_pc, gd, dtype, allocate)


def Laplace(gd, scale=1.0, n=1, dtype=float, allocate=True):
    if n == 9:
        raise NotImplementedError('FTLaplace (FFT Laplacian) is not part of '
                                  'the finite-difference GPU path')
    if config.extra_parameters.get('newgucstencil', True):
        return NewGUCLaplace(gd, scale, n, dtype, allocate)
    return GUCLaplace(gd, scale, n, dtype, allocate)


class LaplaceA(FDOperator):
    def __init__(self, gd, scale, dtype=float, allocate=True):
        assert gd.orthogonal
        c = np.divide(-1 / 12, gd.h_cv.diagonal() ** 2) * scale
        c0 = c[1] + c[2]
        c1 = c[0] + c[2]
        c2 = c[1] + c[0]
        a = -16.0 * np.sum(c)
        b = 10.0 * c + 0.125 * a
        FDOperator.__init__(
            self,
            [a, b[0], b[0], b[1], b[1], b[2], b[2],
             c0, c0, c0, c0, c1, c1, c1, c1, c2, c2, c2, c2],
            [(0, 0, 0),
             (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0), (0, 0, -1), (0, 0, 1),
             (0, -1, -1), (0, -1, 1), (0, 1, -1), (0, 1, 1),
             (-1, 0, -1), (-1, 0, 1), (1, 0, -1), (1, 0, 1),
             (-1, -1, 0), (-1, 1, 0), (1, -1, 0), (1, 1, 0)],
            gd, dtype, allocate=allocate)


class LaplaceB(FDOperator):
    def __init__(self, gd, dtype=float, allocate=True):
        a = 0.5
        b = 1.0 / 12.0
        FDOperator.__init__(
            self, [a, b, b, b, b, b, b],
            [(0, 0, 0), (-1, 0, 0), (1, 0, 0), (0, -1, 0), (0, 1, 0),
             (0, 0, -1), (0, 0, 1)],
            gd, dtype, allocate=allocate)


class NewGUCLaplace(FDOperator):
    def __init__(self, gd, scale=1.0, n=1, dtype=float, allocate=True):
        """Laplacian for general non orthorhombic grid.

        gd: GridDescriptor; scale: scaling factor (-0.5 for the kinetic
        energy); n: range of stencil, O(h^(2n)) error; dtype: float or
        complex; allocate: create the device object now.
        """
        # the 13 half-neighbour directions, ordered by length after
        # normalising the axes; ties keep their index order (the
        # reference's numpy sorted this 13-element array by insertion sort)
        M_ic = np.indices((3, 3, 3)).reshape((3, -1)).T[-13:] - 1
        u_cv = gd.h_cv / (gd.h_cv ** 2).sum(1)[:, np.newaxis] ** 0.5
        u2_i = (np.dot(M_ic, u_cv) ** 2).sum(1)
        i_d = u2_i.argsort(kind='stable')
        m_mv = np.array([(2, 0, 0), (0, 2, 0), (0, 0, 2),
                         (0, 1, 1), (1, 0, 1), (1, 1, 0)])
        # try 3, 4, 5 and 6 directions
        for D in range(3, 7):
            h_dv = np.dot(M_ic[i_d[:D]], gd.h_cv)
            A_md = (h_dv ** m_mv[:, np.newaxis, :]).prod(2)
            a_d, residual, rank, s = np.linalg.lstsq(A_md, [1, 1, 1, 0, 0, 0],
                                                     rcond=-1)
            if residual.sum() < 1e-14:
                assert rank == D
                break
        a_d = a_d * scale
        offsets = [(0, 0, 0)]
        coefs = [laplace[n][0] * a_d.sum()]
        for d in range(D):
            M_c = M_ic[i_d[d]]
            offsets.extend(np.arange(1, n + 1)[:, np.newaxis] * M_c)
            coefs.extend(a_d[d] * np.array(laplace[n][1:]))
            offsets.extend(np.arange(-1, -n - 1, -1)[:, np.newaxis] * M_c)
            coefs.extend(a_d[d] * np.array(laplace[n][1:]))
        FDOperator.__init__(self, coefs, offsets, gd, dtype, allocate)
        self.n = n

    def __str__(self):
        return ('<%d*%d+1=%d point O(h^%d) finite-difference Laplacian>' %
                ((self.npoints - 1) // self.n, self.n, self.npoints,
                 2 * self.n))


class GUCLaplace(FDOperator):
    def __init__(self, gd, scale=1.0, n=1, dtype=float, allocate=True):
        """Central finite difference Laplacian with cross terms
        (at most 12*n**2 + 6*n neighbours)."""
        n = int(n)
        h = (gd.h_cv ** 2).sum(1) ** 0.5
        h2 = h ** 2
        iucell_cv = gd.xxxiucell_cv
        d2 = (iucell_cv ** 2).sum(1)
        offsets = [(0, 0, 0)]
        coefs = [scale * np.sum(d2 * np.divide(laplace[n][0], h2))]
        for d in range(1, n + 1):
            offsets.extend([(-d, 0, 0), (d, 0, 0), (0, -d, 0), (0, d, 0),
                            (0, 0, -d), (0, 0, d)])
            c = scale * d2 * np.divide(laplace[n][d], h2)
            coefs.extend([c[0], c[0], c[1], c[1], c[2], c[2]])
        ci = 0
        for d1 in range(n):
            for d2_ in range(d1, n):
                offset = [[(d1 + 1, d2_ + 1, 0), (-d1 - 1, d2_ + 1, 0),
                           (d1 + 1, -d2_ - 1, 0), (-d1 - 1, -d2_ - 1, 0)],
                          [(0, d1 + 1, d2_ + 1), (0, -d1 - 1, d2_ + 1),
                           (0, d1 + 1, -d2_ - 1), (0, -d1 - 1, -d2_ - 1)],
                          [(d2_ + 1, 0, d1 + 1), (d2_ + 1, 0, -d1 - 1),
                           (-d2_ - 1, 0, d1 + 1), (-d2_ - 1, 0, -d1 - 1)]]
                for i in range(3):
                    c = scale * 2. * cross[n][ci] * np.dot(
                        iucell_cv[i], iucell_cv[(i + 1) % 3]) / (
                        h[i] * h[(i + 1) % 3])
                    if abs(c) > 1E-11:
                        offsets.extend(offset[i])
                        coefs.extend([c, -c, -c, c])
                        if d2_ > d1:
                            ind = [0, 1, 2]
                            ind[i] = (i + 1) % 3
                            ind[(i + 1) % 3] = i
                            offsets.extend([tuple(np.take(offset[i][i2], ind))
                                            for i2 in range(4)])
                            coefs.extend([c, -c, -c, c])
                ci += 1
        FDOperator.__init__(self, coefs, offsets, gd, dtype, allocate)


def PoissonLaplace(gd, nn=3, allocate=True):
    """The Poisson operator -1/(4 pi) nabla^2 of gpaw/poisson.py:51-59."""
    return Laplace(gd, -0.25 / pi, nn, allocate=allocate)

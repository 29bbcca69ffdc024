"""Uniform 3-D grid geometry for one domain.

Mirrors the part of gpaw/grid_descriptor.py:65-237 and gpaw/domain.py:15-150
that the finite-difference path uses: grid sizes (non-periodic axes store
N_c - 1 points starting at global index 1), spacings, volume element,
neighbours (the domain itself on periodic axes, nobody on zero-boundary
axes), scaled displacements for Bloch phases, coarsen / refine, integrate.
One GPU holds the whole grid, so there is no domain decomposition: the
communicator is a serial stub with the reference's `sum` semantics.
"""
import numpy as np

from . import _arrays


class SerialCommunicator:
    """gpaw/mpi/__init__.py:510-574 for one rank."""
    size = 1
    rank = 0

    def sum(self, a, root=-1):
        if isinstance(a, (int, float, complex)):
            return a
        return a

    def max(self, a, root=-1):
        return a

    def barrier(self):
        pass

    def get_c_object(self):
        raise NotImplementedError('no MPI communicator on one GPU')


serial_comm = SerialCommunicator()


class GridDescriptor:
    def __init__(self, N_c, cell_cv=(1, 1, 1), pbc_c=True, comm=None,
                 parsize=None):
        if isinstance(pbc_c, (bool, int, np.bool_)):
            pbc_c = (bool(pbc_c),) * 3
        if comm is not None and getattr(comm, 'size', 1) != 1:
            raise NotImplementedError(
                'domain decomposition over MPI is not part of the one-GPU path')
        if parsize not in (None, 1, (1, 1, 1)):
            raise NotImplementedError('parsize must be 1 on one GPU')
        self.comm = serial_comm
        cell = np.array(cell_cv, float)
        if cell.ndim == 1:
            self.cell_cv = np.diag(cell)
        else:
            self.cell_cv = cell
        cell_c = (self.cell_cv ** 2).sum(1) ** 0.5
        self.icell_cv = np.linalg.inv(self.cell_cv).T
        self.xxxucell_cv = np.array([self.cell_cv[x] / cell_c[x]
                                     for x in range(3)])
        self.xxxiucell_cv = np.linalg.inv(self.xxxucell_cv.T)
        self.pbc_c = np.asarray(pbc_c, bool)
        self.N_c = np.array(N_c, int)
        self.parsize_c = np.ones(3, int)
        self.parpos_c = np.zeros(3, int)
        self.stride_c = np.array([1, 1, 1])
        self.beg_c = np.where(self.pbc_c, 0, 1)
        self.end_c = self.N_c.copy()
        if (self.end_c - self.beg_c <= 0).any():
            raise ValueError('Grid too small!')
        self.n_c = self.end_c - self.beg_c
        self.h_cv = self.cell_cv / self.N_c[:, np.newaxis]
        self.dv = abs(np.linalg.det(self.cell_cv)) / self.N_c.prod()
        self.orthogonal = not (self.cell_cv -
                               np.diag(self.cell_cv.diagonal())).any()
        L_c = (np.linalg.inv(self.cell_cv) ** 2).sum(0) ** -0.5
        h_c = L_c / self.N_c
        if max(h_c) / min(h_c) > 1.3:
            raise ValueError('Very anisotropic grid spacings: %s' % h_c)
        # find_neighbor_processors (gpaw/domain.py:122-150) for one domain
        self.neighbor_cd = np.zeros((3, 2), int)
        self.sdisp_cd = np.zeros((3, 2), int)
        for c in range(3):
            if self.pbc_c[c]:
                self.sdisp_cd[c] = (1, -1)
            else:
                self.neighbor_cd[c] = -1
        self.use_fixed_bc = False

    def __repr__(self):
        return 'GridDescriptor(%s, pbc=%s)' % (list(self.N_c),
                                               list(self.pbc_c.astype(int)))

    def get_size_of_global_array(self, pad=False):
        if pad:
            return self.N_c
        return self.N_c - 1 + self.pbc_c

    def flat_index(self, G_c):
        g1, g2, g3 = np.asarray(G_c) - self.beg_c
        return g3 + self.n_c[2] * (g2 + g1 * self.n_c[1])

    def get_slice(self):
        return [slice(b - 1 + p, e - 1 + p) for b, e, p in
                zip(self.beg_c, self.end_c, self.pbc_c)]

    # -- arrays: device tensors by default, numpy on request ------------------
    def _shape(self, n):
        if isinstance(n, int):
            n = (n,)
        return tuple(n) + tuple(self.n_c)

    def zeros(self, n=(), dtype=float, global_array=False, pad=False,
              device=True):
        if global_array or pad:
            raise NotImplementedError('global/padded arrays are the local '
                                      'array on one domain')
        if not device:
            return np.zeros(self._shape(n), dtype)
        return _arrays.zeros(self._shape(n), dtype)

    def empty(self, n=(), dtype=float, global_array=False, pad=False,
              device=True):
        if global_array or pad:
            raise NotImplementedError('global/padded arrays are the local '
                                      'array on one domain')
        if not device:
            return np.empty(self._shape(n), dtype)
        return _arrays.empty(self._shape(n), dtype)

    def integrate(self, a_xg, b_xg=None, global_integral=True):
        """gpaw/grid_descriptor.py:197-215.  Device tensors are reduced on
        the GPU with a deterministic FP64 sum; numpy arrays with numpy."""
        if _arrays.is_tensor(a_xg):
            from .ops import integrate as _dev_integrate
            return _dev_integrate(self, a_xg, b_xg)
        shape = a_xg.shape
        if len(shape) == 3:
            if b_xg is None:
                return a_xg.sum() * self.dv
            return np.vdot(a_xg, b_xg) * self.dv
        A_x = np.sum(np.reshape(a_xg, shape[:-3] + (-1,)), axis=-1)
        return A_x * self.dv

    def coarsen(self):
        if (self.N_c % 2).any():
            raise ValueError('Grid %s not divisible by 2!' % self.N_c)
        gd = GridDescriptor(self.N_c // 2, self.cell_cv, self.pbc_c)
        gd.use_fixed_bc = self.use_fixed_bc
        return gd

    def refine(self):
        gd = GridDescriptor(self.N_c * 2, self.cell_cv, self.pbc_c)
        gd.use_fixed_bc = self.use_fixed_bc
        return gd

    def get_grid_point_coordinates(self, dtype=float, global_array=False):
        idx = np.indices(self.n_c, dtype).T + self.beg_c
        return np.dot(idx, self.h_cv).T.copy()

    def bytecount(self, dtype=float):
        return int(np.prod(self.n_c)) * np.array(1, dtype).itemsize

    def phase_cd(self, k_c):
        """Bloch phases of scaled k-point k_c (gpaw/kpt_descriptor.py:162-166)."""
        return np.exp(2j * np.pi * self.sdisp_cd *
                      np.asarray(k_c, float)[:, np.newaxis])

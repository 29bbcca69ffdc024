"""Eigensolver preconditioner on the GPU (drop-in for gpaw/preconditioner.py).

Same constructor, allocate() and __call__(residuals, kpt) as the reference
class (gpaw/preconditioner.py:13-70): two coarser grids with nn=1 kinetic
operators and nn=1 transformers, one damped Jacobi-like step per level.
The whole sequence for all bands runs inside libgpaw_amd.so
(gpaw_preconditioner_apply): ten launches per call instead of the
reference's per-block operator calls and numpy temporaries.

As in the reference, the returned array is the preconditioner's own
buffer; the next call overwrites it.
"""
import ctypes

from . import _arrays, _lib
from ._arrays import Arg
from .fd_operators import Laplace
from .ops import OperatorHandle, TransformerHandle, _phases
from .transformers import Transformer


def _handle(obj):
    """The C handle behind an FDOperator / Transformer (or their debug
    wrappers)."""
    while not isinstance(obj, (OperatorHandle, TransformerHandle)):
        nxt = getattr(obj, 'operator', None)
        if nxt is None:
            nxt = getattr(obj, 'transformer', None)
        if nxt is None:
            raise TypeError('%r is not an allocated operator or transformer'
                            % (obj,))
        obj = nxt
    return obj.h


class Preconditioner:
    def __init__(self, gd0, kin0, dtype=float, block=1):
        gd1 = gd0.coarsen()
        gd2 = gd1.coarsen()
        self.gd0 = gd0
        self.kin0 = kin0
        self.kin1 = Laplace(gd1, -0.5, 1, dtype)
        self.kin2 = Laplace(gd2, -0.5, 1, dtype)
        self.dtype = dtype
        self.block = block
        self.step = 0.66666666 / kin0.get_diagonal_element()
        self.restrictor_object0 = Transformer(gd0, gd1, 1, dtype, False)
        self.restrictor_object1 = Transformer(gd1, gd2, 1, dtype, False)
        self.interpolator_object2 = Transformer(gd2, gd1, 1, dtype, False)
        self.interpolator_object1 = Transformer(gd1, gd0, 1, dtype, False)
        self.allocated = False
        self._pc = None
        self._work = None
        self._out = None

    def allocate(self):
        assert not self.allocated
        for transformer in [self.restrictor_object0,
                            self.restrictor_object1,
                            self.interpolator_object2,
                            self.interpolator_object1]:
            transformer.allocate()
        self.allocated = True

    def _plan(self):
        if self._pc is None:
            h = ctypes.c_void_p()
            _lib.check(_lib.lib().gpaw_preconditioner_new(
                _handle(self.kin0), _handle(self.kin1), _handle(self.kin2),
                _handle(self.restrictor_object0),
                _handle(self.restrictor_object1),
                _handle(self.interpolator_object2),
                _handle(self.interpolator_object1), ctypes.byref(h)))
            self._pc = h
        return self._pc

    def __del__(self):
        pc = getattr(self, '_pc', None)
        if pc is not None and pc.value:
            try:
                _lib.lib().gpaw_preconditioner_free(pc)
            except Exception:
                pass
            self._pc = None

    def _buffers(self, nb, shape, tdtype, code):
        torch = _arrays.torch()
        L = _lib.lib()
        nbytes = ctypes.c_long()
        _lib.check(L.gpaw_preconditioner_workspace(self._plan(), nb, code,
                                                   ctypes.byref(nbytes)))
        # torch's caching allocator returns blocks aligned to >= 512 bytes
        if self._work is None or self._work.numel() < nbytes.value:
            self._work = torch.empty(max(nbytes.value, 256), dtype=torch.uint8,
                                     device='cuda')
        if (self._out is None or self._out.dtype != tdtype or
                self._out.shape[0] < nb or
                tuple(self._out.shape[1:]) != tuple(shape[1:])):
            self._out = torch.empty((max(nb, self.block),) + tuple(shape[1:]),
                                    dtype=tdtype, device='cuda')
        return self._work, self._out[:nb]

    def __call__(self, residuals, kpt):
        """Preconditioned residuals of all bands (gpaw/preconditioner.py:
        43-70).  residuals: (nbands, n0, n1, n2) GPU tensor or numpy array
        (a 3-D array is one band); kpt: any object with phase_cd (or None
        for a real, Gamma-point calculation)."""
        if not self.allocated:
            self.allocate()
        phases = getattr(kpt, 'phase_cd', None) if kpt is not None else None
        r = Arg(residuals, name='residuals')
        if r.code not in (_lib.GPAW_F64, _lib.GPAW_C128):
            raise TypeError('preconditioner: float64 or complex128 residuals')
        shape = r.shape if len(r.shape) == 4 else (1,) + r.shape
        if len(shape) != 4 or tuple(shape[1:]) != tuple(self.gd0.n_c):
            raise ValueError('preconditioner: residuals of shape %s for grid '
                             '%s' % (r.shape, tuple(self.gd0.n_c)))
        nb = shape[0]
        work, out = self._buffers(nb, shape, r.tensor.dtype, r.code)
        _, pp = _phases(phases)
        _lib.check(_lib.lib().gpaw_preconditioner_apply(
            self._pc, r.ptr, out.data_ptr(), nb, r.code, pp,
            work.data_ptr(), work.numel(), _arrays.stream()))
        if len(r.shape) == 3:
            out = out[0]
        if r.host is None:
            return out
        return out.cpu().numpy()

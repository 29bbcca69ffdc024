"""Batched FD apply on the GPU: one persistent launch over many grids.

The reference applies an operator to the bands of one grid and one
k-point per call (FDOperator.apply -> Operator_apply,
c/operators.c:316-405).  BatchApply takes a list of such calls -- bands
of different k-points (each with its own phase_cd), grids of uneven
size, operators with different coefficients -- and runs them as one
persistent kernel per stencil range (gpaw_batch_apply).  Each group's
result is bit-identical to its own FDOperator.apply.
"""
import ctypes

import numpy as np

from . import _arrays, _lib
from ._arrays import Arg
from .ops import OperatorHandle


def _operator_handle(op):
    while not isinstance(op, OperatorHandle):
        op = op.operator
    return op


class BatchApply:
    def __init__(self, max_groups=256):
        self.max_groups = max_groups
        h = ctypes.c_void_p()
        _lib.check(_lib.lib().gpaw_batch_new(int(max_groups), ctypes.byref(h)))
        self.h = h

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                _lib.lib().gpaw_batch_free(h)
            except Exception:
                pass
            self.h = None

    def __call__(self, groups):
        """groups: sequence of (operator, in_xg, out_xg, phase_cd or None).
        Inputs and outputs are GPU tensors of one dtype, 3-D (one band) or
        4-D (bands first) with the operator's grid as the last three axes."""
        groups = list(groups)
        n = len(groups)
        if n > self.max_groups:
            raise ValueError('%d groups, the batch holds %d' % (n, self.max_groups))
        if n == 0:
            return
        ops = (ctypes.c_void_p * n)()
        ins = (ctypes.c_void_p * n)()
        outs = (ctypes.c_void_p * n)()
        nins = np.zeros(n, np.int64)
        phases = np.zeros((n, 3, 2), np.complex128)
        phases[:] = 1.0
        code = None
        args = []
        for i, (op, a, b, ph) in enumerate(groups):
            h = _operator_handle(op)
            x = Arg(a, name='input')
            y = Arg(b, write=True, name='output')
            if x.host is not None or y.host is not None:
                raise TypeError('batched apply takes GPU tensors')
            if x.shape != y.shape or tuple(x.shape[-3:]) != h.shape:
                raise ValueError('group %d: shapes %s -> %s for grid %s' %
                                 (i, x.shape, y.shape, h.shape))
            if code is None:
                code = x.code
            if x.code != code or y.code != code:
                raise TypeError('all groups must have one dtype')
            ops[i] = h.h.value
            ins[i] = x.ptr
            outs[i] = y.ptr
            nins[i] = 1 if len(x.shape) == 3 else int(np.prod(x.shape[:-3]))
            if ph is not None:
                phases[i] = np.asarray(ph, np.complex128).reshape(3, 2)
            args.append((x, y))
        flat = np.ascontiguousarray(phases).view(np.float64).ravel()
        _lib.check(_lib.lib().gpaw_batch_apply(
            self.h, n, ops, ins, outs, nins.ctypes.data_as(_lib._lp),
            flat.ctypes.data_as(_lib._dp), code, _arrays.stream()))

"""Device buffers for grid functions.

torch (ROCm) tensors are only used as HBM allocations and for their HIP
stream; all arithmetic on them is done by libgpaw_amd.so.  Public entry
points accept either torch tensors already on the GPU (no copies, work is
ordered on torch's current stream) or numpy arrays (copied to the device
and back; the call is then synchronous like the reference C extension).
"""
import numpy as np

from . import _lib

_NP = {np.dtype(np.float64): _lib.GPAW_F64,
       np.dtype(np.complex128): _lib.GPAW_C128,
       np.dtype(np.float32): _lib.GPAW_F32,
       np.dtype(np.complex64): _lib.GPAW_C64}


_CODE_NAMES = {_lib.GPAW_F64: 'float64', _lib.GPAW_C128: 'complex128',
               _lib.GPAW_F32: 'float32', _lib.GPAW_C64: 'complex64'}


def torch():
    import torch as _t
    return _t


def _torch_codes():
    t = torch()
    return {t.float64: _lib.GPAW_F64, t.complex128: _lib.GPAW_C128,
            t.float32: _lib.GPAW_F32, t.complex64: _lib.GPAW_C64}


def is_tensor(a):
    try:
        return isinstance(a, torch().Tensor)
    except ImportError:
        return False


def dtype_code(a):
    if is_tensor(a):
        return _torch_codes()[a.dtype]
    return _NP[np.dtype(a.dtype)]


def stream():
    t = torch()
    return t.cuda.current_stream().cuda_stream


class Arg:
    """One array argument of a library call.

    ptr   device pointer to pass to C
    code  GPAW_* dtype code
    """

    def __init__(self, a, write=False, name='array', shape=None, code=None):
        t = torch()
        self.host = None
        self.write = write
        if is_tensor(a):
            if not a.is_cuda:
                raise ValueError('%s: tensor must live on the GPU' % name)
            if not a.is_contiguous():
                raise ValueError('%s: tensor must be C-contiguous' % name)
            self.tensor = a
        else:
            a = np.asarray(a) if not isinstance(a, np.ndarray) else a
            if not a.flags.c_contiguous:
                raise ValueError('%s: array must be C-contiguous' % name)
            if np.dtype(a.dtype) not in _NP:
                raise TypeError('%s: unsupported dtype %s' % (name, a.dtype))
            self.host = a
            self.tensor = t.from_numpy(a).to('cuda')
        self.code = dtype_code(self.tensor)
        self.shape = tuple(self.tensor.shape)
        if code is not None and self.code != code:
            raise TypeError('%s: dtype %s, expected %s' % (
                name, self.tensor.dtype, _CODE_NAMES[code]))
        if shape is not None and self.shape != tuple(shape):
            raise ValueError('%s: shape %s, expected %s' % (
                name, self.shape, tuple(shape)))
        self.ptr = self.tensor.data_ptr()

    def finish(self):
        """Copy a written numpy argument back (synchronises)."""
        if self.host is not None and self.write:
            self.host[...] = self.tensor.cpu().numpy()


def empty(shape, dtype=np.float64):
    t = torch()
    conv = {np.dtype(np.float64): t.float64, np.dtype(np.complex128): t.complex128,
            np.dtype(np.float32): t.float32, np.dtype(np.complex64): t.complex64}
    return t.empty(tuple(shape), dtype=conv[np.dtype(dtype)], device='cuda')


def zeros(shape, dtype=np.float64):
    return empty(shape, dtype).zero_()

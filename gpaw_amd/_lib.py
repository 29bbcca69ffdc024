"""ctypes binding of libgpaw_amd.so (the C ABI of include/gpaw_amd.h).

The library is the only compute path: if it is missing this module raises
ImportError-derived errors at first use instead of falling back to anything
else.
"""
import ctypes
import os

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, 'libgpaw_amd.so')

GPAW_F64, GPAW_C128, GPAW_F32, GPAW_C64 = 0, 1, 2, 3
GPAW_OK, GPAW_EINVAL, GPAW_EUNSUPPORTED = 0, -1, -2
GPAW_EDEVICE, GPAW_ENOMEM, GPAW_ENOCONV = -3, -4, -5

_vp = ctypes.c_void_p
_dp = ctypes.POINTER(ctypes.c_double)
_lp = ctypes.POINTER(ctypes.c_long)
_ip = ctypes.POINTER(ctypes.c_int)
_i, _l, _d = ctypes.c_int, ctypes.c_long, ctypes.c_double

# name: (restype, argtypes)
PROTOTYPES = {
    'gpaw_last_error': (ctypes.c_char_p, []),
    'gpaw_operator_new': (_i, [_dp, _lp, _i, _lp, _i, _lp, _i, _i,
                               ctypes.POINTER(_vp)]),
    'gpaw_operator_apply': (_i, [_vp, _vp, _vp, _l, _i, _dp, _vp]),
    'gpaw_operator_relax': (_i, [_vp, _i, _vp, _vp, _i, _d, _vp]),
    'gpaw_operator_get_diagonal_element': (_d, [_vp]),
    'gpaw_operator_residual': (_i, [_vp, _vp, _vp, _vp, _vp, _d, _vp]),
    'gpaw_operator_free': (None, [_vp]),
    'gpaw_transformer_new': (_i, [_lp, _lp, _i, _lp, _lp, _lp, _lp, _i, _i,
                                  ctypes.POINTER(_vp)]),
    'gpaw_transformer_apply': (_i, [_vp, _vp, _vp, _l, _i, _dp, _vp]),
    'gpaw_transformer_free': (None, [_vp]),
    'gpaw_poisson_new': (_i, [_i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                              ctypes.POINTER(_vp), _ip, _ip, _dp, _i, _d,
                              ctypes.POINTER(_vp)]),
    'gpaw_poisson_solve_neutral': (_i, [_vp, _vp, _vp, _d, _i, _ip, _dp,
                                        _vp]),
    'gpaw_poisson_vcycles': (_i, [_vp, _vp, _vp, _i, _dp, _vp]),
    'gpaw_poisson_mixed_weights': (_i, [_vp, _i, _dp, _ip, _dp, _ip]),
    'gpaw_poisson_solve': (_i, [_vp, _vp, _vp, _d, _i, _d, _dp, _ip, _ip, _dp, _vp]),
    'gpaw_poisson_neutralize': (_i, [_vp, _vp, _d, _dp, _ip, _vp]),
    'gpaw_poisson_solve_mixed': (_i, [_vp, _vp, _vp, _d, _i, _i, _i, _ip, _dp,
                                      _vp]),
    'gpaw_poisson_precondition': (_i, [_vp, _vp, _vp, _i, _vp]),
    'gpaw_poisson_free': (None, [_vp]),
    'gpaw_preconditioner_new': (_i, [_vp, _vp, _vp, _vp, _vp, _vp, _vp,
                                     ctypes.POINTER(_vp)]),
    'gpaw_preconditioner_step': (_d, [_vp]),
    'gpaw_preconditioner_workspace': (_i, [_vp, _l, _i, ctypes.POINTER(_l)]),
    'gpaw_preconditioner_apply': (_i, [_vp, _vp, _vp, _l, _i, _dp, _vp, _l,
                                       _vp]),
    'gpaw_preconditioner_free': (None, [_vp]),
    'gpaw_batch_new': (_i, [_i, ctypes.POINTER(_vp)]),
    'gpaw_batch_apply': (_i, [_vp, _i, ctypes.POINTER(_vp), ctypes.POINTER(_vp),
                              ctypes.POINTER(_vp), _lp, _dp, _i, _vp]),
    'gpaw_batch_free': (None, [_vp]),
    'gpaw_autotune': (_i, [_i]),
    'gpaw_autotune_report': (_l, [ctypes.c_char_p, _l]),
    'gpaw_stencil_kernel': (_i, [_i]),
    'gpaw_tile_shape': (_i, [_i, _i, _i]),
    'gpaw_sum': (_i, [_vp, _l, _d, _vp, _vp]),
    'gpaw_dots': (_i, [_vp, _vp, _l, _l, _d, _vp, _vp]),
    'gpaw_dots_noalloc': (_i, [_vp, _vp, _l, _l, _d, _vp, _vp]),
    'gpaw_axpby': (_i, [_vp, _vp, _d, _d, _l, _vp]),
    'gpaw_sub_scalar': (_i, [_vp, _vp, _d, _l, _vp]),
    'gpaw_cast': (_i, [_vp, _i, _vp, _i, _l, _vp]),
    'gpaw_copy': (_i, [_vp, _vp, _l, _vp]),
    'gpaw_add_potential': (_i, [_vp, _vp, _vp, _l, _l, _i, _vp]),
    'gpaw_operator_apply_potential': (_i, [_vp, _vp, _vp, _l, _i, _dp, _vp, _vp]),
    'gpaw_scale': (_i, [_vp, ctypes.c_double, _l, _vp]),
}


class GpawAmdError(RuntimeError):
    def __init__(self, code, msg):
        RuntimeError.__init__(self, '%s (code %d)' % (msg, code))
        self.code = code


class LibraryNotBuilt(ImportError):
    pass


_lib = None


def lib():
    """Load libgpaw_amd.so once and return it (raises if it is absent)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LibraryNotBuilt(
                '%s is not built; run `python -m gpaw_amd.build` '
                '(there is no CPU fallback)' % LIB_PATH)
        # torch ships its own libamdhip64 (soname libamdhip64.so.7).  Load it
        # first so that our NEEDED libamdhip64.so.7 binds to that copy: two
        # HIP runtimes in one process cannot both open the GPU.
        import torch  # noqa: F401
        L = ctypes.CDLL(LIB_PATH)
        for name, (res, args) in PROTOTYPES.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def check(rc, exc=None):
    if rc != GPAW_OK:
        msg = lib().gpaw_last_error().decode(errors='replace')
        if exc is not None:
            raise exc(msg)
        raise GpawAmdError(rc, msg)


def longs(a):
    """C long array (kept alive by the caller) from an int sequence."""
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.int64).ravel())
    return arr, arr.ctypes.data_as(_lp)


def doubles(a):
    import numpy as np
    arr = np.ascontiguousarray(np.asarray(a, dtype=np.float64).ravel())
    return arr, arr.ctypes.data_as(_dp)

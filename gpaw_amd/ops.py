"""Thin Python handles over the C ABI (include/gpaw_amd.h).

These are the drop-ins for the reference's compiled objects:
  OperatorHandle     <- _gpaw.Operator    (c/operators.c:465-498)
  TransformerHandle  <- _gpaw.Transformer (c/transformers.c:222-265)
Their methods take the same arguments as the reference objects
(apply(in, out, phases), relax(method, f, s, n, w), get_diagonal_element(),
get_async_sizes()), accepting GPU tensors or numpy arrays.
"""
import ctypes

import numpy as np

from . import _arrays, _lib
from ._arrays import Arg


def _phases(phases):
    if phases is None:
        return None, None
    ph = np.ascontiguousarray(np.asarray(phases, np.complex128).reshape(3, 2))
    flat = ph.view(np.float64).ravel().copy()
    return flat, flat.ctypes.data_as(_lib._dp)


def _nin(shape, size):
    if len(shape) == 3:
        return 1
    if len(shape) == 4:
        return shape[0]
    # the reference reads dims[0] only for 4-D input; other ranks flatten
    return int(np.prod(shape[:-3]))


class OperatorHandle:
    def __init__(self, coef_p, offset_p, size_c, mp, neighbor_cd, real,
                 comm=None, cfd=True):
        if comm is not None:
            raise NotImplementedError('MPI communicators are not supported')
        L = _lib.lib()
        self._coefs, cp = _lib.doubles(coef_p)
        self._offs, op = _lib.longs(offset_p)
        self._size, sp = _lib.longs(size_c)
        self._nb, nbp = _lib.longs(neighbor_cd)
        h = ctypes.c_void_p()
        _lib.check(L.gpaw_operator_new(cp, op, len(self._coefs), sp, int(mp),
                                       nbp, int(bool(real)), int(bool(cfd)),
                                       ctypes.byref(h)))
        self.h = h
        self.real = bool(real)
        self.shape = tuple(int(x) for x in size_c)

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                _lib.lib().gpaw_operator_free(h)
            except Exception:
                pass
            self.h = None

    def apply(self, in_xg, out_xg, phases=None):
        a = Arg(in_xg, name='input')
        b = Arg(out_xg, write=True, name='output')
        if a.shape != b.shape or a.shape[-3:] != self.shape:
            raise ValueError('operator apply: shapes %s -> %s for grid %s' %
                             (a.shape, b.shape, self.shape))
        if a.code != b.code:
            raise TypeError('operator apply: input and output dtypes differ')
        _, pp = _phases(phases)
        _lib.check(_lib.lib().gpaw_operator_apply(
            self.h, a.ptr, b.ptr, _nin(a.shape, self.shape), a.code, pp,
            _arrays.stream()))
        b.finish()

    def apply_potential(self, in_xg, out_xg, v_g, phases=None):
        """out = L in + in * v in one pass (v: float64 grid function shared
        by the bands; gpaw/wavefunctions/fd.py:51-52)."""
        a = Arg(in_xg, name='input')
        b = Arg(out_xg, write=True, name='output')
        v = Arg(v_g, name='potential')
        if a.shape != b.shape or a.shape[-3:] != self.shape or tuple(v.shape) != self.shape:
            raise ValueError('operator apply_potential: shapes %s -> %s, potential %s for grid %s'
                             % (a.shape, b.shape, v.shape, self.shape))
        if a.code != b.code:
            raise TypeError('operator apply_potential: input and output dtypes differ')
        if v.code != _lib.GPAW_F64:
            raise TypeError('operator apply_potential: the potential is float64')
        _, pp = _phases(phases)
        _lib.check(_lib.lib().gpaw_operator_apply_potential(
            self.h, a.ptr, b.ptr, _nin(a.shape, self.shape), a.code, pp, v.ptr,
            _arrays.stream()))
        b.finish()

    def relax(self, relax_method, f_g, s_g, nrelax, w=1.0):
        if w is None:   # PyArg_ParseTuple("iOOi|d") rejects None
            raise TypeError('a float is required')
        f = Arg(f_g, write=True, name='f')
        s = Arg(s_g, name='s')
        if f.shape != self.shape or s.shape != self.shape:
            raise ValueError('relax: shapes %s, %s for grid %s' %
                             (f.shape, s.shape, self.shape))
        if f.code != _lib.GPAW_F64 or s.code != _lib.GPAW_F64:
            raise TypeError('relax works on float64 arrays')
        _lib.check(_lib.lib().gpaw_operator_relax(
            self.h, int(relax_method), f.ptr, s.ptr, int(nrelax), float(w),
            _arrays.stream()))
        f.finish()

    def residual(self, phi, rho, r=None, norm_out=None, scale=1.0):
        """r = L phi - rho (and scale * sum r^2 into the 1-element tensor
        norm_out), fused."""
        F64 = _lib.GPAW_F64
        p = Arg(phi, name='phi', shape=self.shape, code=F64)
        q = Arg(rho, name='rho', shape=self.shape, code=F64)
        rr = Arg(r, write=True, name='r', shape=self.shape, code=F64) \
            if r is not None else None
        nn = Arg(norm_out, write=True, name='norm', shape=(1,), code=F64) \
            if norm_out is not None else None
        _lib.check(_lib.lib().gpaw_operator_residual(
            self.h, p.ptr, q.ptr, rr.ptr if rr else None,
            nn.ptr if nn else None, float(scale), _arrays.stream()))
        if rr:
            rr.finish()
        if nn:
            nn.finish()

    def get_diagonal_element(self):
        return _lib.lib().gpaw_operator_get_diagonal_element(self.h)

    def get_async_sizes(self):
        return (0, 1, 1)


class TransformerHandle:
    def __init__(self, size_in, size_out, k, pad_cd, npad_cd, skip_cd,
                 neighbor_cd, real, comm=None, interpolate=False):
        if comm is not None:
            raise NotImplementedError('MPI communicators are not supported')
        L = _lib.lib()
        keep = []

        def lp(a):
            arr, p = _lib.longs(a)
            keep.append(arr)
            return p
        h = ctypes.c_void_p()
        _lib.check(L.gpaw_transformer_new(
            lp(size_in), lp(size_out), int(k), lp(pad_cd), lp(npad_cd),
            lp(skip_cd), lp(neighbor_cd), int(bool(real)), int(bool(interpolate)),
            ctypes.byref(h)))
        self.h = h
        self.ngpin = tuple(int(x) for x in size_in)
        self.ngpout = tuple(int(x) for x in size_out)

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                _lib.lib().gpaw_transformer_free(h)
            except Exception:
                pass
            self.h = None

    def apply(self, input, output, phases=None):
        a = Arg(input, name='input')
        b = Arg(output, write=True, name='output')
        if a.shape[-3:] != self.ngpin or b.shape[-3:] != self.ngpout or \
                a.shape[:-3] != b.shape[:-3]:
            raise ValueError('transformer apply: shapes %s -> %s for %s -> %s'
                             % (a.shape, b.shape, self.ngpin, self.ngpout))
        if a.code != b.code:
            raise TypeError('transformer apply: input and output dtypes differ')
        _, pp = _phases(phases)
        _lib.check(_lib.lib().gpaw_transformer_apply(
            self.h, a.ptr, b.ptr, _nin(a.shape, self.ngpin), a.code, pp,
            _arrays.stream()))
        b.finish()

    def get_async_sizes(self):
        return (0, 1)


def integrate(gd, a_xg, b_xg=None):
    """Device version of GridDescriptor.integrate for float64 tensors."""
    t = _arrays.torch()
    if b_xg is not None:
        raise NotImplementedError('device integrate of a product')
    if a_xg.dtype != t.float64:
        raise TypeError('device integrate needs float64')
    a = Arg(a_xg)
    lead = a.shape[:-3]
    ng = int(np.prod(a.shape[-3:]))
    out = t.empty(max(1, int(np.prod(lead))), dtype=t.float64, device='cuda')
    L = _lib.lib()
    st = _arrays.stream()
    for i in range(out.numel()):
        _lib.check(L.gpaw_sum(a.ptr + 8 * ng * i, ng, float(gd.dv),
                              out.data_ptr() + 8 * i, st))
    vals = out.cpu().numpy()
    if not lead:
        return float(vals[0])
    return vals.reshape(lead)


def dots(a_g, b_xg, scale=1.0):
    """scale * sum_g a_g * b_xg[j] for every leading index j of b_xg, as an
    FP64 device tensor of b_xg's leading shape (one row of the pair-density
    integrals gd.integrate(rhot * phit), gpaw/lrtddft/omega_matrix.py:427)."""
    t = _arrays.torch()
    a, b = Arg(a_g), Arg(b_xg)
    if a.tensor.dtype != t.float64 or b.tensor.dtype != t.float64:
        raise TypeError('dots needs float64 arrays')
    n = a.tensor.numel()
    lead = tuple(b.shape[:b.tensor.dim() - a.tensor.dim()])
    if tuple(b.shape[len(lead):]) != tuple(a.shape):
        raise ValueError('dots: trailing shape %s of b differs from %s'
                         % (tuple(b.shape), tuple(a.shape)))
    nb = int(np.prod(lead)) if lead else 1
    out = t.empty(nb, dtype=t.float64, device='cuda')
    _lib.check(_lib.lib().gpaw_dots(a.ptr, b.ptr, nb, n, float(scale), out.data_ptr(),
                                    _arrays.stream()))
    return out.reshape(lead)


def sub_scalar(y, x, b):
    """y = x - b on float64 device tensors."""
    yy = Arg(y, write=True, name='y', code=_lib.GPAW_F64)
    xx = Arg(x, name='x', shape=yy.shape, code=_lib.GPAW_F64)
    _lib.check(_lib.lib().gpaw_sub_scalar(yy.ptr, xx.ptr, float(b),
                                          int(np.prod(yy.shape)),
                                          _arrays.stream()))
    yy.finish()


def axpby(y, x, a, b):
    """y = a*y + b*x on float64 device tensors."""
    yy = Arg(y, write=True, name='y', code=_lib.GPAW_F64)
    xx = Arg(x, name='x', shape=yy.shape, code=_lib.GPAW_F64)
    _lib.check(_lib.lib().gpaw_axpby(yy.ptr, xx.ptr, float(a), float(b),
                                     int(np.prod(yy.shape)), _arrays.stream()))
    yy.finish()


def cast(y, x):
    """Precision conversion (float64 <-> float32, complex128 <-> complex64)."""
    yy = Arg(y, write=True)
    xx = Arg(x)
    _lib.check(_lib.lib().gpaw_cast(yy.ptr, yy.code, xx.ptr, xx.code,
                                    int(np.prod(yy.shape)), _arrays.stream()))
    yy.finish()


def copy(dst, src):
    """dst[...] = src on the device with the library's 16-byte copy kernel
    (both C-contiguous tensors of the same size in bytes)."""
    d = Arg(dst, write=True, name='dst')
    a = Arg(src, name='src')
    nb = a.tensor.numel() * a.tensor.element_size()
    if d.tensor.numel() * d.tensor.element_size() != nb:
        raise ValueError('copy: sizes differ')
    _lib.check(_lib.lib().gpaw_copy(d.ptr, a.ptr, nb, _arrays.stream()))
    d.finish()


def add_potential(out, psi, v):
    """out += psi * v on the device: psi / out of shape (..., n0, n1, n2)
    (float64 or complex128 bands), v the float64 grid function shared by
    the bands (gpaw/hamiltonian.py:393-398)."""
    o = Arg(out, write=True, name='out')
    a = Arg(psi, name='psi')
    p = Arg(v, name='v')
    if p.code != _lib.GPAW_F64:
        raise TypeError('add_potential: the potential is float64')
    if o.code != a.code or tuple(o.shape) != tuple(a.shape):
        raise ValueError('add_potential: out and psi differ in dtype or shape')
    ng = int(np.prod(p.shape))
    if tuple(a.shape[-len(p.shape):]) != tuple(p.shape):
        raise ValueError('add_potential: potential shape %s, bands %s' % (p.shape, a.shape))
    nin = int(np.prod(a.shape)) // max(ng, 1)
    _lib.check(_lib.lib().gpaw_add_potential(o.ptr, a.ptr, p.ptr, nin, ng, a.code,
                                             _arrays.stream()))
    o.finish()


def scale(y, a):
    """y *= a on float64 device tensors."""
    yy = Arg(y, write=True)
    _lib.check(_lib.lib().gpaw_scale(yy.ptr, float(a), int(np.prod(yy.shape)),
                                     _arrays.stream()))
    yy.finish()


def autotune(enable=True):
    """Switch the launch-shape autotuner of the stencil kernels on or off;
    returns the previous setting."""
    return bool(_lib.lib().gpaw_autotune(int(bool(enable))))


STENCIL_KERNELS = {'default': 0, 'tile': 1, 'box': 2, 'queue': 3, 'generic': 4}


def tile_shape(threads=0, tz=32, planes=1):
    """Force the tile kernel's launch shape (with stencil_kernel('tile')) for
    launches without partial sums; threads=0 restores the default shape."""
    _lib.check(_lib.lib().gpaw_tile_shape(int(threads), int(tz), int(planes)))


def stencil_kernel(kind='default'):
    """Kernel family for star-stencil launches without partial sums
    ('default', 'tile', 'box', 'queue' or 'generic'; every family gives the
    same bits); returns the previous family."""
    prev = _lib.lib().gpaw_stencil_kernel(STENCIL_KERNELS[kind])
    return {v: k for k, v in STENCIL_KERNELS.items()}[prev]


def autotune_report():
    """The tuner's choices so far, one line per (dtype, mode, range, grid,
    bands) combination."""
    L = _lib.lib()
    n = L.gpaw_autotune_report(None, 0)
    buf = ctypes.create_string_buffer(n + 1)
    L.gpaw_autotune_report(buf, n + 1)
    return buf.value.decode()

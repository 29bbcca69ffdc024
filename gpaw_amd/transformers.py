"""Grid transformers on the GPU (drop-in for gpaw/transformers.py).

_Transformer computes the reference's paddings and skips
(gpaw/transformers.py:19-62) and drives ops.TransformerHandle over
libgpaw_amd.so in place of _gpaw.Transformer.
"""
import numpy as np

from . import config
from .ops import TransformerHandle


class _Transformer:
    def __init__(self, gdin, gdout, nn=1, dtype=float, allocate=True):
        self.gdin = gdin
        self.gdout = gdout
        self.nn = nn
        assert 1 <= nn <= 4
        self.dtype = dtype
        pad_cd = np.empty((3, 2), int)
        neighborpad_cd = np.empty((3, 2), int)
        skip_cd = np.zeros((3, 2), int)
        if (gdin.N_c == 2 * gdout.N_c).all():
            # restriction
            pad_cd[:, 0] = 2 * nn - 1 - 2 * gdout.beg_c + gdin.beg_c
            pad_cd[:, 1] = 2 * nn - 2 + 2 * gdout.end_c - gdin.end_c
            neighborpad_cd[:, 0] = 2 * nn - 2 + 2 * gdout.beg_c - gdin.beg_c
            neighborpad_cd[:, 1] = 2 * nn - 1 - 2 * gdout.end_c + gdin.end_c
            self.interpolate = False
        else:
            assert (gdout.N_c == 2 * gdin.N_c).all()
            # interpolation
            pad_cd[:, 0] = nn - 1 - gdout.beg_c // 2 + gdin.beg_c
            pad_cd[:, 1] = nn + gdout.end_c // 2 - gdin.end_c
            neighborpad_cd[:, 0] = nn + gdout.beg_c // 2 - gdin.beg_c
            neighborpad_cd[:, 1] = nn - 1 - gdout.end_c // 2 + gdin.end_c
            skip_cd[:, 0] = gdout.beg_c % 2
            skip_cd[:, 1] = gdout.end_c % 2
            self.interpolate = True
            assert ((gdin.n_c[0] + 2 * nn - 1) * (gdin.n_c[1] + 2 * nn - 1) <=
                    gdout.n_c[0] * gdout.n_c[1])
        assert (pad_cd.ravel() >= 0).all()
        self.ngpin = tuple(gdin.n_c)
        self.ngpout = tuple(gdout.n_c)
        assert dtype in [float, complex]
        self.transformer = None
        self.pad_cd = pad_cd
        self.neighborpad_cd = neighborpad_cd
        self.skip_cd = skip_cd
        self.allocated = False
        if allocate:
            self.allocate()

    def allocate(self):
        assert not self.allocated
        self.transformer = TransformerHandle(
            self.gdin.n_c, self.gdout.n_c, 2 * self.nn, self.pad_cd,
            self.neighborpad_cd, self.skip_cd, self.gdin.neighbor_cd,
            self.dtype == float, None, self.interpolate)
        self.allocated = True

    def apply(self, input, output, phases=None):
        self.transformer.apply(input, output, phases)

    def get_async_sizes(self):
        return self.transformer.get_async_sizes()

    def estimate_memory(self, mem):
        # the kernels read the input in place: no padded or scratch buffers
        mem.subnode('buf', 0)


class TransformerWrapper:
    """Argument checks of gpaw/transformers.py:110-139 (debug mode)."""

    def __init__(self, transformer):
        self.transformer = transformer
        self.allocated = transformer.allocated
        self.dtype = transformer.dtype
        self.ngpin = transformer.ngpin
        self.ngpout = transformer.ngpout

    def allocate(self):
        assert not self.allocated
        self.transformer.allocate()
        self.allocated = True

    def apply(self, input, output, phases=None):
        assert tuple(input.shape[-3:]) == self.ngpin
        assert tuple(output.shape[-3:]) == self.ngpout
        assert (self.dtype == float or
                (np.asarray(phases).dtype == complex and
                 np.asarray(phases).shape == (3, 2)))
        assert self.allocated
        self.transformer.apply(input, output, phases)

    def get_async_sizes(self):
        return self.transformer.get_async_sizes()

    def estimate_memory(self, mem):
        self.transformer.estimate_memory(mem)


def Transformer(gdin, gdout, nn=1, dtype=float, allocate=True):
    if nn != 9:
        t = _Transformer(gdin, gdout, nn, dtype, allocate)
        if config.debug:
            t = TransformerWrapper(t)
        return t

    class T:
        def allocate(self):
            pass

        def apply(self, input, output, phases=None):
            output[:] = input

        def estimate_memory(self, mem):
            mem.set('Nulltransformer', 0)
    return T()

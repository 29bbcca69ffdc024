"""Multigrid Poisson solver on the GPU (drop-in for gpaw/poisson.py:21-304).

Same constructor, hierarchy (up to four coarsenings, nn-point operator on
the finest grid, 7-point operators and K=2 transformers below it), smoothing
counts (2/1 on level 0, 4/4 below, 8 on the coarsest), Jacobi weights
(2/3, then 1), stopping rule (dv * sum r^2 <= eps, at most maxiter V-cycles)
and charged-system branches.  The V-cycle itself runs in libgpaw_amd.so as
one captured hipGraph per cycle (gpaw_poisson_solve_neutral); Python only
prepares the neutral density and reads back one number per cycle.
"""
import ctypes
from math import pi, sqrt

import numpy as np

from . import _arrays, _lib, ops
from ._arrays import Arg
from .fd_operators import Laplace, LaplaceA, LaplaceB
from .transformers import Transformer
from .utilities import Gaussian, madelung


class PoissonConvergenceError(RuntimeError):
    pass


def _handle(obj):
    t = getattr(obj, 'transformer', None)
    if t is not None:
        if hasattr(t, 'transformer'):       # debug wrapper
            t = t.transformer
        return t.h
    return obj.operator.h


class PoissonSolver:
    def __init__(self, nn=3, relax='J', eps=2e-10):
        self.relax = relax
        self.nn = nn
        self.eps = eps
        self.charged_periodic_correction = None
        self.maxiter = 1000
        if relax == 'GS':
            self.relax_method = 1
        elif relax == 'J':
            self.relax_method = 2
        else:
            raise NotImplementedError('Relaxation method %s' % relax)
        self._plan = None

    def get_method(self):
        return ['Gauss-Seidel', 'Jacobi'][self.relax_method - 1]

    def get_stencil(self):
        return self.nn

    def set_grid_descriptor(self, gd):
        self.gd = gd
        self.dv = gd.dv
        scale = -0.25 / pi
        if self.nn == 'M':
            if not gd.orthogonal:
                raise RuntimeError('Cannot use Mehrstellen stencil with '
                                   'non orthogonal cell.')
            self.operators = [LaplaceA(gd, -scale, allocate=False)]
            self.B = LaplaceB(gd, allocate=False)
        else:
            self.operators = [Laplace(gd, scale, self.nn, allocate=False)]
            self.B = None
        self.interpolators = []
        self.restrictors = []
        level = 0
        self.presmooths = [2]
        self.postsmooths = [1]
        self.weights = [2.0 / 3.0]
        while level < 4:
            try:
                gd2 = gd.coarsen()
            except ValueError:
                break
            self.operators.append(Laplace(gd2, scale, 1, allocate=False))
            self.interpolators.append(Transformer(gd2, gd, allocate=False))
            self.restrictors.append(Transformer(gd, gd2, allocate=False))
            self.presmooths.append(4)
            self.postsmooths.append(4)
            self.weights.append(1.0)
            level += 1
            gd = gd2
        self.levels = level

    def initialize(self, load_gauss=False):
        gd = self.gd
        self.rhos = [gd.empty()]
        self.phis = [None]
        self.residuals = [gd.empty()]
        for level in range(self.levels):
            gd = gd.coarsen()
            self.phis.append(gd.empty())
            self.rhos.append(gd.empty())
            self.residuals.append(gd.empty())
        for obj in self.operators + self.interpolators + self.restrictors:
            obj.allocate()
        if self.B is not None:
            self.B.allocate()
        self.step = 0.66666666 / self.operators[0].get_diagonal_element()
        self.presmooths[self.levels] = 8
        self.postsmooths[self.levels] = 8
        self._make_plan()
        if load_gauss:
            self.load_gauss()

    def _make_plan(self):
        L = self.levels
        VP = ctypes.c_void_p
        opa = (VP * (L + 1))(*[_handle(o) for o in self.operators])
        ra = (VP * max(L, 1))(*[_handle(t) for t in self.restrictors])
        ia = (VP * max(L, 1))(*[_handle(t) for t in self.interpolators])
        pre = (ctypes.c_int * (L + 1))(*self.presmooths)
        post = (ctypes.c_int * (L + 1))(*self.postsmooths)
        w = (ctypes.c_double * (L + 1))(*self.weights)
        h = VP()
        _lib.check(_lib.lib().gpaw_poisson_new(
            L, opa, ra, ia, pre, post, w, self.relax_method, float(self.dv),
            ctypes.byref(h)))
        self._plan = h

    def __del__(self):
        p = getattr(self, '_plan', None)
        if p is not None and p.value:
            try:
                _lib.lib().gpaw_poisson_free(p)
            except Exception:
                pass
            self._plan = None

    def load_gauss(self):
        if not hasattr(self, 'rho_gauss'):
            gauss = Gaussian(self.gd)
            self.rho_gauss = _arrays.torch().from_numpy(
                gauss.get_gauss(0)).to('cuda')
            self.phi_gauss = _arrays.torch().from_numpy(
                gauss.get_gauss_pot(0)).to('cuda')

    # ------------------------------------------------------------------ solve
    def solve(self, phi, rho, charge=None, eps=None, maxcharge=1e-6,
              zero_initial_phi=False):
        if eps is None:
            eps = self.eps
        P, R = self._args(phi, rho)
        phi_d, rho_d = P.tensor, R.tensor
        if charge is None and self._device_branch():
            # the charge integral and the neutral branch's mean subtraction
            # on the device, the solve from the density left in the plan
            # (same kernels, same bits)
            niter, actual_charge = self._solve_device(P, R, eps, maxcharge)
            if niter is not None:
                return niter
        else:
            actual_charge = self.gd.integrate(rho_d)
        background = (actual_charge / self.gd.dv /
                      self.gd.get_size_of_global_array().prod())
        if charge is None:
            charge = actual_charge
        if abs(charge) <= maxcharge:
            rho_n = _arrays.torch().empty_like(rho_d)
            ops.sub_scalar(rho_n, rho_d, background)
            niter = self.solve_neutral(phi_d, rho_n, eps=eps)
        elif abs(charge) > maxcharge and self.gd.pbc_c.all():
            if self.charged_periodic_correction is None:
                self.charged_periodic_correction = madelung(self.gd.cell_cv)
            shift = charge * self.charged_periodic_correction
            if zero_initial_phi:
                phi_d.zero_()
            else:
                ops.sub_scalar(phi_d, phi_d, shift)
            rho_n = _arrays.torch().empty_like(rho_d)
            ops.sub_scalar(rho_n, rho_d, background)
            niter = self.solve_neutral(phi_d, rho_n, eps=eps)
            ops.sub_scalar(phi_d, phi_d, -shift)
        elif abs(charge) > maxcharge and not self.gd.pbc_c.any():
            self.load_gauss()
            q = actual_charge / sqrt(4 * pi)
            rho_n = rho_d.clone()
            ops.axpby(rho_n, self.rho_gauss, 1.0, -q)
            if zero_initial_phi:
                phi_d.zero_()
            else:
                ops.axpby(phi_d, self.phi_gauss, 1.0, -q)
            niter = self.solve_neutral(phi_d, rho_n, eps=eps)
            ops.axpby(phi_d, self.phi_gauss, 1.0, q)
        else:
            raise NotImplementedError
        P.finish()
        return niter

    def _device_branch(self):
        """Whether solve() may run its neutral branch on the device: a
        Laplacian (no Mehrstellen right-hand side) and this class's own
        solve_neutral."""
        return self.B is None and type(self).solve_neutral is PoissonSolver.solve_neutral

    def _solve_device(self, P, R, eps, maxcharge):
        """gpaw_poisson_solve: (niter, charge) when rho was neutral and the
        solve ran, else (None, charge) for the charged branches."""
        self.phis[0] = P.tensor
        q = ctypes.c_double(0.0)
        neutral = ctypes.c_int(0)
        niter = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        rc = _lib.lib().gpaw_poisson_solve(
            self._plan, P.ptr, R.ptr, float(eps), int(self.maxiter), float(maxcharge),
            ctypes.byref(q), ctypes.byref(neutral), ctypes.byref(niter), ctypes.byref(err),
            _arrays.stream())
        if not neutral.value:
            _lib.check(rc)
            return None, q.value
        self.error = err.value
        if rc == _lib.GPAW_ENOCONV:
            P.finish()
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!' %
                self.maxiter)
        _lib.check(rc)
        return self._finish_neutral(P, niter.value), q.value

    def _finish_neutral(self, P, niter):
        """The end of solve_neutral: phi's mean removed on periodic grids."""
        if np.all(self.gd.pbc_c):
            phi_ave = self.gd.integrate(P.tensor) / self.gd.dv
            phi_ave /= np.prod(self.gd.get_size_of_global_array())
            ops.sub_scalar(P.tensor, P.tensor, phi_ave)
        P.finish()
        return niter

    def _args(self, phi, rho):
        """phi (written) and rho as device arguments; both must be float64
        arrays of this grid's shape (the plan reads and writes exactly that
        many points)."""
        n = tuple(int(x) for x in self.gd.n_c)
        return (Arg(phi, write=True, name='phi', shape=n, code=_lib.GPAW_F64),
                Arg(rho, name='rho', shape=n, code=_lib.GPAW_F64))

    def solve_neutral(self, phi, rho, eps=2e-10):
        P, R = self._args(phi, rho)
        if self.B is not None:
            # Mehrstellen: the right-hand side is B rho (gpaw/poisson.py:196-199)
            brho = _arrays.torch().empty_like(R.tensor)
            self.B.apply(R.tensor, brho)
            R = Arg(brho, name='B rho')
        self.phis[0] = P.tensor
        niter = ctypes.c_int(0)
        err = ctypes.c_double(0.0)
        rc = _lib.lib().gpaw_poisson_solve_neutral(
            self._plan, P.ptr, R.ptr, float(eps), int(self.maxiter),
            ctypes.byref(niter), ctypes.byref(err), _arrays.stream())
        self.error = err.value
        if rc == _lib.GPAW_ENOCONV:
            P.finish()
            raise PoissonConvergenceError(
                'Poisson solver did not converge in %d iterations!' %
                self.maxiter)
        _lib.check(rc)
        if np.all(self.gd.pbc_c):
            phi_ave = self.gd.integrate(P.tensor) / self.gd.dv
            phi_ave /= np.prod(self.gd.get_size_of_global_array())
            ops.sub_scalar(P.tensor, P.tensor, phi_ave)
        P.finish()
        return niter.value

    def iterate2(self, step, level=0):
        """One V-cycle on (phis[0], rhos[0]); returns dv * sum r^2."""
        assert level == 0
        errs = (ctypes.c_double * 1)()
        P, R = self._args(self.phis[0], self.rhos[0])
        _lib.check(_lib.lib().gpaw_poisson_vcycles(
            self._plan, P.ptr, R.ptr, 1, errs, _arrays.stream()))
        P.finish()
        return errs[0]

    def vcycles(self, phi, rho, count):
        """Exactly `count` V-cycles on an already neutral rho; returns the
        error after each cycle (fixed-cycle comparisons)."""
        errs = (ctypes.c_double * max(count, 1))()
        P, R = self._args(phi, rho)
        _lib.check(_lib.lib().gpaw_poisson_vcycles(
            self._plan, P.ptr, R.ptr, int(count), errs, _arrays.stream()))
        P.finish()
        return list(errs[:count])

    def estimate_memory(self, mem):
        gdbytes = self.gd.bytecount()
        mem.subnode('V-cycle chunk buffers (phi after each of up to 32 cycles '
                    'and a sweep buffer)', 33 * gdbytes)
        nbytes = 0
        for level in range(self.levels + 1):
            nbytes += 4 * gdbytes        # phi pair, rho, residual
            gdbytes //= 8
        mem.subnode('phi, rho, residual [%d levels]' % self.levels, nbytes)

    def __repr__(self):
        template = 'PoissonSolver(relax=\'%s\', nn=%s, eps=%e)'
        return template % (self.relax, repr(self.nn), self.eps)


class MixedPoissonSolver(PoissonSolver):
    """Mixed-precision multigrid Poisson solver.

    Same hierarchy and charge handling as PoissonSolver (the charge
    integral, background and monopole / Madelung terms stay FP64).  The
    neutral problem is solved by an FP64 outer loop -- residual L phi - rho,
    correction accumulation, dot products -- around inner V-cycles (FP32
    smoothing, residuals, restriction, interpolation; the smallest levels in
    FP64).  outer='pcg': flexible preconditioned CG, one V-cycle per
    iteration; outer='refine': iterative refinement with `inner` V-cycles per
    correction.  inner_precision='fp64' runs the same inner V-cycle in FP64
    (same outer loop), which separates what FP32 contributes from what the
    outer iteration does.  Converged when ||L phi - rho||_2 <= rel_tol *
    ||rho||_2 for the neutral density (checked on the true residual); `eps`
    is accepted for signature compatibility but the relative criterion is
    the one applied.  reuse_history (CG): the first chunk of iterations is
    as long as the previous converged solve took, so repeated solves of
    similar densities (an SCF loop) read a norm back once; the step count,
    and with it the last bits of phi beyond the tolerance, then depend on
    that history.  False makes every result depend on its inputs only.
    """

    def __init__(self, nn=3, relax='J', eps=2e-10, rel_tol=1e-10, outer='pcg',
                 inner=1, maxiter=200, inner_precision='fp32',
                 reuse_history=True):
        PoissonSolver.__init__(self, nn, relax, eps)
        if outer not in ('pcg', 'refine'):
            raise ValueError('outer must be "pcg" or "refine"')
        if inner_precision not in ('fp32', 'fp64'):
            raise ValueError('inner_precision must be "fp32" or "fp64"')
        self.rel_tol = rel_tol
        self.outer = outer
        self.inner = inner
        self.inner_precision = inner_precision
        self.maxiter = maxiter
        self.reuse_history = reuse_history

    def _device_branch(self):
        return self.relax_method == 2 and self.B is None

    def _solve_device(self, P, R, eps, maxcharge):
        """gpaw_poisson_neutralize, then the mixed solve of the density it
        left in the plan: (niter, charge), or (None, charge) when charged."""
        q = ctypes.c_double(0.0)
        neutral = ctypes.c_int(0)
        _lib.check(_lib.lib().gpaw_poisson_neutralize(
            self._plan, R.ptr, float(maxcharge), ctypes.byref(q), ctypes.byref(neutral),
            _arrays.stream()))
        if not neutral.value:
            return None, q.value
        return self._solve_mixed(P, None), q.value

    def solve_neutral(self, phi, rho, eps=None):
        if self.relax_method != 2 or self.B is not None:
            raise NotImplementedError('mixed precision needs Jacobi and a '
                                      'Laplace operator')
        P, R = self._args(phi, rho)
        return self._solve_mixed(P, R.ptr)

    def _solve_mixed(self, P, rho_ptr):
        """gpaw_poisson_solve_mixed on phi P and rho at rho_ptr (None: the
        density gpaw_poisson_neutralize left in the plan)."""
        niter = ctypes.c_int(0)
        rel = ctypes.c_double(0.0)
        rc = _lib.lib().gpaw_poisson_solve_mixed(
            self._plan, P.ptr, rho_ptr, float(self.rel_tol), int(self.maxiter),
            (1 if self.outer == 'pcg' else 0) |
            (2 if self.inner_precision == 'fp64' else 0) |
            (4 if self.reuse_history else 0), int(self.inner),
            ctypes.byref(niter), ctypes.byref(rel), _arrays.stream())
        self.relres = rel.value
        if rc == _lib.GPAW_ENOCONV:
            P.finish()
            raise PoissonConvergenceError(
                'mixed-precision Poisson solver did not converge in %d '
                'iterations' % self.maxiter)
        _lib.check(rc)
        return self._finish_neutral(P, niter.value)

    def smoothing_weights(self, level=0):
        """(pre, post) Jacobi weights of level `level` in the inner V-cycle,
        one per sweep in the order they run."""
        pre = (ctypes.c_double * 16)()
        post = (ctypes.c_double * 16)()
        npre, npost = ctypes.c_int(0), ctypes.c_int(0)
        _lib.check(_lib.lib().gpaw_poisson_mixed_weights(
            self._plan, level, pre, ctypes.byref(npre), post, ctypes.byref(npost)))
        return list(pre[:npre.value]), list(post[:npost.value])

    def precondition(self, r, z, precision=None):
        """z = one inner V-cycle applied to r from z = 0 (the multigrid
        preconditioner of the outer CG): float64 grid functions of the
        solver's grid, on the device or as numpy arrays.  precision 'fp32'
        or 'fp64' (default: the solver's inner_precision)."""
        precision = precision or self.inner_precision
        if precision not in ('fp32', 'fp64'):
            raise ValueError('precision must be "fp32" or "fp64"')
        if self.relax_method != 2:
            raise NotImplementedError('the mixed V-cycle uses Jacobi smoothing')
        shape = tuple(self.gd.n_c)
        Rr = Arg(r, name='r', shape=shape, code=_lib.GPAW_F64)
        Z = Arg(z, write=True, name='z', shape=shape, code=_lib.GPAW_F64)
        _lib.check(_lib.lib().gpaw_poisson_precondition(
            self._plan, Rr.ptr, Z.ptr, 1 if precision == 'fp64' else 0,
            _arrays.stream()))
        Z.finish()

    def __repr__(self):
        return ('MixedPoissonSolver(nn=%r, outer=%r, rel_tol=%e, '
                'inner_precision=%r)' % (self.nn, self.outer, self.rel_tol,
                                         self.inner_precision))

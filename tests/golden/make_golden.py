"""Write the golden fixtures under tests/golden/ from the reference's own C.

The reference's c/bmgs/bmgs.c and c/bc.c are compiled in place into
oracle/_ref/libgpawref.so (``make -C oracle ref``); this script drives them
through oracle.Kernels('reference') on small seeded inputs and stores inputs
and outputs as .npz data.  Run from the repository root:

    python tests/golden/make_golden.py

Each fixture holds the stencil/transformer description, the input arrays and
the reference output.  No reference source is copied: only numbers.
"""
import os
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.dirname(os.path.abspath(
    __file__))))
sys.path.insert(0, ROOT)

import oracle as O  # noqa: E402
from oracle.gpaw_oracle import Operator, Transformer  # noqa: E402

OUT = os.path.dirname(os.path.abspath(__file__))


def _rand(rng, shape, dtype):
    a = rng.uniform(-1, 1, shape)
    if dtype == complex:
        a = a + 1j * rng.uniform(-1, 1, shape)
    return np.ascontiguousarray(a)


def operator_cases():
    """(name, grid, coefs, offsets, dtype, nin, k_c)"""
    cases = []
    for nn in (1, 2, 3):
        for pbc in (True, False):
            g = O.Grid((10, 12, 8), (3.0, 3.6, 2.4), pbc)
            c, o = O.laplace_stencil(g, -0.25 / np.pi, nn)
            cases.append(('lap_nn%d_%s_real' % (nn, 'pbc' if pbc else 'zero'),
                          g, c, o, float, 2, None))
    g = O.Grid((8, 8, 8), (2.4, 2.4, 2.4), True)
    c, o = O.laplace_stencil(g, -0.5, 3)
    cases.append(('lap_nn3_bloch_complex', g, c, o, complex, 3,
                  (0.13, -0.37, 0.41)))
    g = O.Grid((8, 10, 8), (2.4, 3.0, 2.4), (True, False, True))
    c, o = O.laplace_stencil(g, 1.0, 2)
    cases.append(('lap_nn2_mixed_complex', g, c, o, complex, 2,
                  (0.25, 0.0, -0.11)))
    # non-orthogonal cell -> more than 3 directions, cfd = 0, corner halos
    g = O.Grid((8, 8, 8), [(2.0, 0, 0), (1.0, 1.7, 0), (0, 0, 2.1)], True)
    c, o = O.laplace_stencil(g, 1.0, 2)
    cases.append(('lap_nn2_hex_complex', g, c, o, complex, 2,
                  (0.2, -0.3, 0.05)))
    cases.append(('lap_nn2_hex_real', g, c, o, float, 1, None))
    g = O.Grid((8, 10, 12), (2.4, 3.0, 3.6), False)
    c, o = O.mehrstellen_a(g, 0.25 / np.pi)
    cases.append(('mehrstellenA_zero_real', g, c, o, float, 1, None))
    g = O.Grid((10, 8, 8), (3.0, 2.4, 2.4), True)
    for v in range(3):
        c, o = O.gradient_stencil(g, v, 1.0, 3)
        cases.append(('grad%d_nn3_pbc_complex' % v, g, c, o, complex, 1,
                      (0.3, 0.1, -0.2)))
    return cases


def transformer_cases():
    cases = []
    for nn in (1, 2, 3, 4):
        for pbc in (True, False):
            g = O.Grid((12, 12, 16), (3.0, 3.0, 4.0), pbc)
            tag = 'pbc' if pbc else 'zero'
            cases.append(('restrict_k%d_%s_real' % (2 * nn, tag), g,
                          g.coarsen(), nn, float, 2, None))
            gc = O.Grid((8, 8, 12), (3.0, 3.0, 4.5), pbc)
            cases.append(('interp_k%d_%s_real' % (2 * nn, tag), gc,
                          gc.refine(), nn, float, 2, None))
    g = O.Grid((16, 12, 12), (4.0, 3.0, 3.0), True)
    cases.append(('restrict_k6_bloch_complex', g, g.coarsen(), 3, complex, 2,
                  (0.21, -0.44, 0.07)))
    gc = O.Grid((8, 6, 10), (4.0, 3.0, 5.0), True)
    cases.append(('interp_k6_bloch_complex', gc, gc.refine(), 3, complex, 2,
                  (-0.12, 0.33, 0.48)))
    g = O.Grid((12, 16, 12), (3.0, 4.0, 3.0), (False, True, True))
    cases.append(('restrict_k4_mixed_complex', g, g.coarsen(), 2, complex, 1,
                  (0.0, 0.3, -0.25)))
    gc = O.Grid((8, 8, 6), (3.0, 3.0, 2.25), (True, False, True))
    cases.append(('interp_k8_mixed_complex', gc, gc.refine(), 4, complex, 1,
                  (0.4, 0.0, 0.15)))
    return cases


def main():
    K = O.Kernels('reference')
    rng = np.random.default_rng(20240611)
    written = []
    for name, g, c, o, dtype, nin, k_c in operator_cases():
        op = Operator(c, o, g, dtype, kernels=K)
        a = _rand(rng, (nin,) + tuple(g.n_c), dtype)
        b = np.zeros_like(a)
        ph = g.phase_cd(k_c) if dtype == complex else None
        op.apply(a, b, ph)
        d = dict(kind='operator', coefs=op.coefs, offsets_pc=o, N_c=g.N_c,
                 cell_cv=g.cell_cv, pbc_c=g.pbc_c, a=a, b=b)
        if ph is not None:
            d['phase_cd'] = ph
        np.savez_compressed(os.path.join(OUT, 'op_%s.npz' % name), **d)
        written.append(name)
    # smoother: Jacobi and Gauss-Seidel sweeps on real data
    for method, tag in ((2, 'jacobi'), (1, 'gs')):
        for pbc in (True, False):
            for nn in (1, 3):
                g = O.Grid((10, 8, 12), (3.0, 2.4, 3.6), pbc)
                c, o = O.laplace_stencil(g, -0.25 / np.pi, nn)
                op = Operator(c, o, g, kernels=K)
                f = _rand(rng, tuple(g.n_c), float)
                s = _rand(rng, tuple(g.n_c), float)
                w = 2.0 / 3.0 if nn == 3 else 1.0
                f0 = f.copy()
                op.relax(method, f, s, 3, w)
                name = 'relax_%s_nn%d_%s' % (tag, nn, 'pbc' if pbc else 'zero')
                np.savez_compressed(
                    os.path.join(OUT, 'op_%s.npz' % name), kind='relax',
                    coefs=op.coefs, offsets_pc=o, N_c=g.N_c, cell_cv=g.cell_cv,
                    pbc_c=g.pbc_c, f0=f0, s=s, f=f, method=method, nrelax=3,
                    w=w)
                written.append(name)
    for name, gi, go, nn, dtype, nin, k_c in transformer_cases():
        t = Transformer(gi, go, nn, dtype, kernels=K)
        a = _rand(rng, (nin,) + tuple(gi.n_c), dtype)
        b = np.zeros((nin,) + tuple(go.n_c), dtype)
        ph = gi.phase_cd(k_c) if dtype == complex else None
        t.apply(a, b, ph)
        d = dict(kind='transformer', nn=nn, N_in=gi.N_c, N_out=go.N_c,
                 cell_cv=gi.cell_cv, pbc_c=gi.pbc_c, a=a, b=b)
        if ph is not None:
            d['phase_cd'] = ph
        np.savez_compressed(os.path.join(OUT, 'tr_%s.npz' % name), **d)
        written.append(name)
    # Poisson V-cycles with the reference kernels under the restated driver
    for pbc, N in ((False, 16), (True, 16)):
        g = O.Grid((N,) * 3, (N * 0.35,) * 3, pbc)
        L = np.diag(g.cell_cv)
        centers = [L * f for f in ((0.3, 0.4, 0.5), (0.65, 0.55, 0.4))]
        rho = O.gaussian_charges(g, centers, (0.9, 1.1), (1.0, -1.0))
        # neutralise exactly as PoissonSolver.solve does before the cycles
        rho = rho - g.integrate(rho) / g.dv / g.size_of_global_array().prod()
        ps = O.PoissonSolver(nn=3, relax='J', kernels=K)
        ps.set_grid_descriptor(g)
        ps.initialize()
        phi = g.zeros()
        errs = ps.run_vcycles(phi, rho, 4)
        tag = 'pbc' if pbc else 'zero'
        np.savez_compressed(os.path.join(OUT, 'poisson_vcycle4_%s.npz' % tag),
                            kind='poisson', N_c=g.N_c, cell_cv=g.cell_cv,
                            pbc_c=g.pbc_c, rho=rho, phi=phi,
                            errors=np.array(errs))
        written.append('poisson_vcycle4_' + tag)
    print('wrote %d fixtures' % len(written))


if __name__ == '__main__':
    main()

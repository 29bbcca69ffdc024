"""Host-side helpers of the charged Poisson branches and synthetic inputs.

  coordinates        gpaw/utilities/tools.py:92-124
  Gaussian (L = 0)   gpaw/utilities/gauss.py:47-83 (gauss_L[0], gausspot_L[0])
  madelung           the Madelung shift of gpaw/poisson.py:160-174 (the
                     reference evaluates it with gpaw/utilities/ewald.py),
                     summed here by Ewald's method to a stated tolerance
These run once per solver on the host (numpy) and their results are moved
to the device.
"""
from math import erf, erfc, pi, sqrt

import numpy as np

# Bohr radius in Angstrom of the ASE version the reference used (ase.units)
Bohr = 0.52917720859


def coordinates(gd, origin=None, tiny=1e-12):
    """Cartesian coordinates relative to `origin` (default: cell centre) and
    r^2 floored at `tiny`; minimum image on periodic orthorhombic axes."""
    if origin is None:
        origin = 0.5 * gd.cell_cv.sum(0)
    r0_v = np.array(origin)
    r_vG = gd.get_grid_point_coordinates() - r0_v[:, None, None, None]
    for c in range(3):
        if gd.pbc_c[c]:
            L = gd.cell_cv[c, c]
            r_vG[c] = np.where(r_vG[c] > 0.5 * L, r_vG[c] - L, r_vG[c])
            r_vG[c] = np.where(r_vG[c] < -0.5 * L, r_vG[c] + L, r_vG[c])
    r2_G = np.sum(r_vG ** 2, axis=0)
    r2_G = np.where(r2_G < tiny, tiny, r2_G)
    return r_vG, r2_G


class Gaussian:
    """Normalised L=0 generalised Gaussian at the cell centre and its
    potential; exponent a = 4 * 19 * max |icell|^2 (the reference default)."""

    def __init__(self, gd, a=19., center=None):
        self.gd = gd
        self.xyz, self.r2 = coordinates(gd, center)
        self.a = 4 * a * (gd.icell_cv ** 2).sum(1).max()

    def get_gauss(self, L=0):
        if L != 0:
            raise NotImplementedError('only the monopole is used by the solver')
        a = self.a
        return np.sqrt(a ** 3 * 4) / pi * np.exp(-a * self.r2)

    def get_gauss_pot(self, L=0):
        if L != 0:
            raise NotImplementedError('only the monopole is used by the solver')
        a = self.a
        r = np.sqrt(self.r2)
        return 2.0 * 1.7724538509055159 * np.vectorize(erf)(np.sqrt(a) * r) / r


def _lattice_points(basis, radius):
    """All integer combinations n . basis with |n . basis| <= radius (rows of
    `basis` are the generating vectors), as an array of vectors."""
    dual = np.linalg.inv(basis).T
    nmax = [int(np.ceil(radius * np.linalg.norm(d))) for d in dual]
    rng = [np.arange(-m, m + 1) for m in nmax]
    n = np.stack(np.meshgrid(*rng, indexing='ij'), -1).reshape(-1, 3)
    v = n @ basis
    return v[(v ** 2).sum(1) <= radius * radius]


def madelung(cell, alpha=None, tol=1e-16):
    """Electrostatic potential at a lattice site of a periodic array of unit
    NEGATIVE point charges in a uniform neutralising background, without the
    site's own charge: the potential shift of a charged periodic system
    (gpaw/poisson.py:160-174 multiplies it by the net charge).

    Ewald's splitting of 1/r into erfc(alpha r)/r + erf(alpha r)/r: with
    psi = sum_{L != 0} erfc(alpha|L|)/|L|
          + (4 pi / V) sum_{g != 0} exp(-g^2 / 4 alpha^2) / g^2
          - 2 alpha / sqrt(pi) - pi / (alpha^2 V)
    the potential of unit positive charges at a site, the result is -psi.
    Both sums are taken over every lattice / reciprocal vector whose term
    exceeds `tol` relative to the leading one (erfc(x) < tol beyond
    alpha|L| = x_cut, the Gaussian factor below tol beyond g = 2 alpha
    x_cut), so the value does not depend on alpha.  alpha defaults to
    sqrt(pi) / V^(1/3), which balances the two sums."""
    cell = np.asarray(cell, float)
    vol = abs(np.linalg.det(cell))
    if alpha is None:
        alpha = sqrt(pi) / vol ** (1.0 / 3.0)
    x_cut = 1.0
    while erfc(x_cut) > tol:
        x_cut += 0.25
    L = _lattice_points(cell, x_cut / alpha)
    d = np.sqrt((L ** 2).sum(1))
    d = d[d > 0]
    real = sum(erfc(alpha * x) / x for x in np.sort(d)[::-1])
    G = _lattice_points(2 * pi * np.linalg.inv(cell).T, 2 * alpha * x_cut)
    g2 = (G ** 2).sum(1)
    g2 = np.sort(g2[g2 > 0])[::-1]
    recip = 4 * pi / vol * np.sum(np.exp(-g2 / (4 * alpha * alpha)) / g2)
    psi = real + recip - 2 * alpha / sqrt(pi) - pi / (alpha * alpha * vol)
    return -psi


def gaussian_density(gd, centers, widths, charges, grid_normalized=True):
    """Sum of 3-D Gaussians q exp(-r^2 / 2 s^2) on the grid points (minimum
    image on periodic axes): the synthetic charge density of the benchmark
    configurations.  Each Gaussian is normalised so that its grid integral
    (sum * dv) is exactly its charge q when grid_normalized (a Gaussian near
    a wall otherwise loses part of its charge outside the box and a set of
    +-1 charges is no longer neutral on the grid), else analytically by
    (2 pi s^2)^-3/2."""
    r_vG = gd.get_grid_point_coordinates()
    rho = np.zeros(tuple(gd.n_c))
    L = np.diag(gd.cell_cv)
    for c, s, q in zip(centers, widths, charges):
        d = r_vG - np.asarray(c, float)[:, None, None, None]
        for ax in range(3):
            if gd.pbc_c[ax]:
                d[ax] -= L[ax] * np.round(d[ax] / L[ax])
        r2 = (d ** 2).sum(0)
        g = np.exp(-r2 / (2 * s * s))
        if grid_normalized:
            g *= 1.0 / (g.sum() * gd.dv)
        else:
            g *= (2 * pi * s * s) ** -1.5
        rho += q * g
    return rho


def random_gaussian_charges(gd, rng, count, wmin, wmax, margin):
    """Seeded neutral set of Gaussian charges: `count` centres at least
    `margin` (bohr) from the walls, widths uniform in [wmin, wmax] (bohr),
    charges +1/-1 alternating (sum zero for even counts)."""
    L = np.diag(gd.cell_cv)
    centers = [rng.uniform(margin, L - margin) for _ in range(count)]
    widths = rng.uniform(wmin, wmax, count)
    charges = np.array([1.0 if i % 2 == 0 else -1.0 for i in range(count)])
    return centers, widths, charges

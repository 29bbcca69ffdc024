"""Host-side helpers of the charged Poisson branches and synthetic inputs.

  coordinates        gpaw/utilities/tools.py:92-124
  Gaussian (L = 0)   gpaw/utilities/gauss.py:47-83 (gauss_L[0], gausspot_L[0])
  madelung / Ewald   gpaw/utilities/ewald.py:12-150
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


class Ewald:
    """Ewald sums for a lattice of point charges (gpaw/utilities/ewald.py)."""

    def __init__(self, cell, G=5, Ng=(9, 9, 9), Nl=(3, 3, 3)):
        self.cell = np.asarray(cell, float)
        self.Vcell = abs(np.linalg.det(self.cell))
        self.recip_cell = np.linalg.inv(self.cell).T
        self.Ng = Ng
        self.Nl = Nl
        self.G = G / np.sqrt((self.cell ** 2).sum(1).max())

    def get_sum_recip_ij(self, r_v, eps=1e-10):
        N = self.Ng
        b1, b2, b3 = self.recip_cell
        E_g = 0.
        for i in range(-N[0], N[0] + 1):
            for j in range(-N[1], N[1] + 1):
                for k in range(-N[2], N[2] + 1):
                    g_v = 2 * pi * (i * b1 + j * b2 + k * b3)
                    g2 = np.dot(g_v, g_v)
                    if g2 > eps:
                        x = .25 * g2 / self.G ** 2
                        E_g += np.exp(1.j * np.dot(r_v, g_v) - x) / x
        return pi / (self.Vcell * self.G ** 2) * E_g.real

    def get_sum_real_ij(self, r_v, eps=1e-5):
        N = self.Nl
        a1, a2, a3 = self.cell
        E_r = 0.
        for i in range(-N[0], N[0] + 1):
            for j in range(-N[1], N[1] + 1):
                for k in range(-N[2], N[2] + 1):
                    l_v = i * a1 + j * a2 + k * a3
                    if np.linalg.norm(l_v) > eps:
                        lr = np.linalg.norm(l_v - r_v)
                        E_r += erfc(self.G * lr) / lr
        return E_r

    def get_electrostatic_potential(self, r, r_B, q_B, excludefroml0=None):
        E0 = 0.
        if excludefroml0 is None:
            excludefroml0 = np.zeros(len(q_B), dtype=int)
        if excludefroml0 in range(len(q_B)):
            i = excludefroml0
            excludefroml0 = np.zeros(len(q_B), dtype=int)
            excludefroml0[i] = 1
        assert sum(excludefroml0) <= 1
        for i, q in enumerate(q_B):
            rprime = r - r_B[i]
            absr = np.linalg.norm(rprime)
            E0 += q * self.get_sum_real_ij(rprime)
            E0 += q * self.get_sum_recip_ij(rprime)
            if excludefroml0[i]:
                if absr < 1e-14:
                    E0 -= q * 2. * self.G / np.sqrt(pi)
                else:
                    E0 -= q * erf(absr * self.G) / absr
            else:
                E0 += q * erfc(absr * self.G) / absr
        q_hom = -sum(q_B)
        E0 += q_hom * pi / (self.G ** 2 * self.Vcell)
        return E0


def madelung(cell):
    return Ewald(cell).get_electrostatic_potential(np.zeros(3), np.zeros(3),
                                                   [-1], 0)


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

"""The FP64 quotient of the fused coarse-level kernel (csrc/coarse.hip,
div_rn): a / d from r = 1 / d as q = a r, e = fma(-q, d, a),
result = q if e == 0 else fma(e, r, q).  The reference divides
(c/bmgs/relax.c:62, `w * (*src - x) / s->coefs[0]`), so the kernel is
bit-identical only if this sequence is the correctly rounded quotient.
Here it is evaluated with exact rational arithmetic (each fma rounded once,
as the hardware does) and compared with IEEE division, bit for bit, on
Jacobi centre coefficients and random operands; the two plausible
shortcuts (no correction step, no zero-remainder guard) each fail."""
import random
import struct
from fractions import Fraction

import pytest


def _fma(x, y, z):
    return float(Fraction(x) * Fraction(y) + Fraction(z))


def _div_rn(a, d, r):
    q = a * r
    e = _fma(-q, d, a)
    return q if e == 0.0 else _fma(e, r, q)


def _bits(x):
    return struct.pack('<d', x)


def _operands(rng, n):
    out = [0.0, -0.0, 1.0, -1.0, 3.0, 1e-300 * 1e200]
    for _ in range(n):
        m = rng.random() + 1.0
        out.append(rng.choice((-1, 1)) * m * 2.0 ** rng.randint(-60, 60))
    return out


def _divisors(rng):
    # Jacobi centre coefficients of the 7-point Laplacian (scale -1/(4 pi))
    # on coarse levels of spacings 0.2 .. 3.2 bohr, and random ones
    ds = []
    for h in (0.2, 0.37, 0.4, 0.8, 1.6, 3.2):
        for hy in (h, 1.1 * h):
            ds.append(-0.25 / 3.141592653589793 * -2 * (2 / h ** 2 + 1 / hy ** 2))
    ds += [rng.choice((-1, 1)) * (rng.random() + 1.0) * 2.0 ** rng.randint(-20, 20)
           for _ in range(8)]
    return ds


def test_quotient_is_correctly_rounded():
    rng = random.Random(11)
    for d in _divisors(rng):
        r = 1.0 / d
        for a in _operands(rng, 400):
            assert _bits(_div_rn(a, d, r)) == _bits(a / d), (a, d)
        # exact quotients: a = d k
        for k in range(1, 40):
            a = d * k
            assert _bits(_div_rn(a, d, r)) == _bits(a / d), (a, d)


def test_planted_shortcuts_are_caught():
    rng = random.Random(12)
    bad_product = bad_guard = 0
    for d in _divisors(rng):
        r = 1.0 / d
        for a in _operands(rng, 200):
            if _bits(a * r) != _bits(a / d):
                bad_product += 1
            q = a * r
            unguarded = _fma(_fma(-q, d, a), r, q)
            if _bits(unguarded) != _bits(a / d):
                bad_guard += 1
    assert bad_product > 0     # the product alone is not the quotient
    assert bad_guard > 0       # without the guard, -0 / d comes out +0 for d > 0


@pytest.mark.parametrize('a', [-0.0, 0.0])
def test_signed_zero(a):
    for d in (0.75, -0.75):
        assert _bits(_div_rn(a, d, 1.0 / d)) == _bits(a / d)

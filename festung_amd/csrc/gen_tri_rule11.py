"""Writes tri_rule11.inc: a fully symmetric 28-point quadrature rule on the
reference triangle {x1, x2 >= 0, x1 + x2 <= 1} that is exact for every
polynomial of degree <= 11, with all points inside the triangle and all
weights positive.

Why degree 11: at p = 4 the element term integrates grad(phi_i) (degree 3)
times the discrete concentration (degree 4) times the discrete velocity
(degree 4), a polynomial of degree 11.  The reference integrates it with
quadRule2D(8) (core/integrateRefElemDphiPhiPhi.m:55-82 with
core/quadRule2D.m:104-108: the 64-point conical product rule, exact to
degree 15), so its hatG holds the exact integrals; any rule exact to degree
11 gives the same numbers up to rounding, with fewer than half the points.

Construction.  Orbits of the symmetry group of the triangle in barycentric
coordinates (l1, l2, l3), (x1, x2) = (l2, l3): the centroid, five orbits
(a, a, 1 - 2a) of three points and two orbits (a, b, 1 - a - b) of six
points: 1 + 5 * 2 + 2 * 3 = 17 parameters.  A symmetric rule is exact to
degree 11 iff it integrates the 16 invariants s2^i s3^j (s2 = l1 l2 + l2 l3
+ l3 l1, s3 = l1 l2 l3, 2i + 3j <= 11) exactly; their integrals follow from
the integral of l1^a l2^b l3^c over the triangle, a! b! c! / (a + b + c + 2)!.
That leaves a one-parameter family.  A multistart least-squares search over
this orbit structure found it; along it, the smallest barycentric coordinate
of any point peaks at about 0.0079.  Here the smallest coordinate of the
first six-point orbit is fixed at 1/128 (next to that peak), and the other
16 parameters solve the 16 moment conditions by Newton's method in 60-digit
arithmetic from the starting values below.  The points and weights are
written with 17 significant digits (each rounds to the nearest double).

    python3 festung_amd/csrc/gen_tri_rule11.py   (needs sympy and mpmath)
"""
import os

import mpmath as mp
import sympy as sp

mp.mp.dps = 60

# starting values: w0; (a, w) x 5; (a, b, w) x 2
START = [0.042780211, 0.49601137, 0.0082036685, 0.10327327, 0.019346890, 0.028649162, 0.0052754693,
         0.43841031, 0.033563908, 0.21045715, 0.035188090, 0.0078125, 0.84224742, 0.0052656622,
         0.046104365, 0.66324276, 0.020148623]
FIXED = 11  # index of the fixed parameter (1/128)
INV = [(i, j) for j in range(4) for i in range(6) if 2 * i + 3 * j <= 11]


def exact_invariant_integrals():
    l1, l2, l3 = sp.symbols('l1 l2 l3')
    s2, s3 = l1 * l2 + l2 * l3 + l3 * l1, l1 * l2 * l3
    out = []
    for i, j in INV:
        tot = sp.Rational(0)
        for (a, b, c), co in sp.Poly(sp.expand(s2 ** i * s3 ** j), l1, l2, l3).terms():
            tot += co * sp.factorial(a) * sp.factorial(b) * sp.factorial(c) / sp.factorial(a + b + c + 2)
        out.append(mp.mpf(tot.p) / tot.q)
    return out


def invariants(a, b, c):
    s2, s3 = a * b + b * c + c * a, a * b * c
    return [s2 ** i * s3 ** j for i, j in INV]


def residual(z, target):
    third = mp.mpf(1) / 3
    r = [z[0] * v for v in invariants(third, third, third)]
    for k in range(5):
        a, w = z[1 + 2 * k], z[2 + 2 * k]
        r = [x + 3 * w * v for x, v in zip(r, invariants(a, a, 1 - 2 * a))]
    for k in range(2):
        a, b, w = z[11 + 3 * k], z[12 + 3 * k], z[13 + 3 * k]
        r = [x + 6 * w * v for x, v in zip(r, invariants(a, b, 1 - a - b))]
    return [x / t - 1 for x, t in zip(r, target)]


def solve():
    target = exact_invariant_integrals()
    z = [mp.mpf(v) for v in START]
    z[FIXED] = mp.mpf(1) / 128
    free = [i for i in range(len(z)) if i != FIXED]
    h = mp.mpf(10) ** -30
    for _ in range(50):
        r = residual(z, target)
        if max(abs(x) for x in r) < mp.mpf(10) ** -50:
            return z
        J = mp.matrix(len(r), len(free))
        for col, i in enumerate(free):
            z2 = list(z)
            z2[i] += h
            r2 = residual(z2, target)
            for row in range(len(r)):
                J[row, col] = (r2[row] - r[row]) / h
        d = mp.lu_solve(J, mp.matrix(r))
        for col, i in enumerate(free):
            z[i] -= d[col]
    raise RuntimeError('Newton did not converge')


def points(z):
    """(x1, x2, w) of the 28 points: centroid, then the orbits in order."""
    third = mp.mpf(1) / 3
    pts = [(third, third, z[0])]
    for k in range(5):
        a, w = z[1 + 2 * k], z[2 + 2 * k]
        for l in ((a, a, 1 - 2 * a), (a, 1 - 2 * a, a), (1 - 2 * a, a, a)):
            pts.append((l[1], l[2], w))
    for k in range(2):
        a, b, w = z[11 + 3 * k], z[12 + 3 * k], z[13 + 3 * k]
        c = 1 - a - b
        for l in ((a, b, c), (a, c, b), (b, a, c), (b, c, a), (c, a, b), (c, b, a)):
            pts.append((l[1], l[2], w))
    return pts


def main():
    pts = points(solve())
    assert len(pts) == 28
    assert all(x1 > 0 and x2 > 0 and x1 + x2 < 1 and w > 0 for x1, x2, w in pts)
    lines = ['// Generated by gen_tri_rule11.py: {x1, x2, w} of a fully symmetric 28-point rule',
             '// on the reference triangle, exact to polynomial degree 11 (weights sum to 1/2).']
    for x1, x2, w in pts:
        lines.append('{%s, %s, %s},' % tuple(mp.nstr(v, 17, min_fixed=-5, max_fixed=5) for v in (x1, x2, w)))
    path = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'tri_rule11.inc')
    with open(path, 'w') as f:
        f.write('\n'.join(lines) + '\n')
    print('wrote', path)


if __name__ == '__main__':
    main()

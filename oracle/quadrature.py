"""Quadrature rules (oracle).  Restates core/quadRule1D.m, core/quadRule2D.m
and core/triquad.m.

Pinned by: exactness on monomials up to the stated order (tests), the 1-D
nodes against ``numpy.polynomial.legendre.leggauss``, ``triquad`` against the
textbook Duffy-collapsed Gauss rule built from ``scipy.special.roots_jacobi``
and ``roots_legendre``, and the tabulated constants of the reference files
(known-answer values copied into the tests as data).
"""
import numpy as np


def quadRule1D(qOrd):
    """Gauss-Legendre rule on [0, 1]; core/quadRule1D.m:48-95.

    Returns (Q, W) as 1-D arrays; the weights sum to 1.
    """
    s = np.sqrt
    if qOrd in (0, 1):
        Q = np.array([0.0]); W = np.array([2.0])
    elif qOrd in (2, 3):
        Q = s(1 / 3) * np.array([-1.0, 1.0]); W = np.array([1.0, 1.0])
    elif qOrd in (4, 5):
        Q = s(3 / 5) * np.array([-1.0, 0.0, 1.0]); W = 1 / 9 * np.array([5.0, 8.0, 5.0])
    elif qOrd in (6, 7):
        Q = np.array([-1, -1, 1, 1.0]) * s(3 / 7 + np.array([1, -1, -1, 1.0]) * 2 / 7 * s(6 / 5))
        W = 1 / 36 * (18 + s(30) * np.array([-1, 1, 1, -1.0]))
    elif qOrd in (8, 9):
        Q = np.array([-1, -1, 0, 1, 1.0]) * s(5 + np.array([2, -2, 0, -2, 2.0]) * s(10 / 7)) / 3
        W = 1 / 900 * (322 + 13 * s(70) * np.array([-1, 1, 0, 1, -1.0]) + np.array([0, 0, 190, 0, 0.0]))
    elif qOrd in (10, 11):
        Q = np.array([0.6612093864662645, -0.6612093864662645, -0.2386191860831969,
                      0.2386191860831969, -0.9324695142031521, 0.9324695142031521])
        W = np.array([0.3607615730481386, 0.3607615730481386, 0.4679139345726910,
                      0.4679139345726910, 0.1713244923791704, 0.171324492379170])
    elif qOrd in (12, 13):
        Q = np.array([0.0000000000000000, 0.4058451513773972, -0.4058451513773972,
                      -0.7415311855993945, 0.7415311855993945, -0.9491079123427585,
                      0.9491079123427585])
        W = np.array([0.4179591836734694, 0.3818300505051189, 0.3818300505051189,
                      0.2797053914892766, 0.2797053914892766, 0.1294849661688697,
                      0.1294849661688697])
    elif qOrd in (14, 15):
        Q = np.array([-0.1834346424956498, 0.1834346424956498, -0.5255324099163290,
                      0.5255324099163290, -0.7966664774136267, 0.7966664774136267,
                      -0.9602898564975363, 0.9602898564975363])
        W = np.array([0.3626837833783620, 0.3626837833783620, 0.3137066458778873,
                      0.3137066458778873, 0.2223810344533745, 0.2223810344533745,
                      0.1012285362903763, 0.1012285362903763])
    elif qOrd in (16, 17):
        Q = np.array([0.0000000000000000, -0.8360311073266358, 0.8360311073266358,
                      -0.9681602395076261, 0.9681602395076261, -0.3242534234038089,
                      0.3242534234038089, -0.6133714327005904, 0.6133714327005904])
        W = np.array([0.3302393550012598, 0.1806481606948574, 0.1806481606948574,
                      0.0812743883615744, 0.0812743883615744, 0.3123470770400029,
                      0.3123470770400029, 0.2606106964029354, 0.2606106964029354])
    else:
        raise ValueError('Order %d not implemented' % qOrd)
    return (Q + 1) / 2, W / 2


def triquad(N, v=((0.0, 0.0), (0.0, 1.0), (1.0, 0.0))):
    """Collapsed Gauss-Jacobi x Gauss-Legendre rule on a triangle;
    core/triquad.m:30-76 (Golub-Welsch eigen-solve of the Jacobi matrix for
    alpha=0, beta=1 at :56-60, Newton iteration for the Legendre nodes at
    :62-71, collapse at :73-76).

    Returns X, Y (N x N, Jacobi index fastest in column-major order) and the
    1-D weight vectors Wx, Wy.
    """
    v = np.asarray(v, dtype=float)
    n = np.arange(1, N + 1, dtype=float)
    nnk = 2 * n + 1
    A = np.concatenate([[1 / 3], np.ones(N) / (nnk * (nnk + 2))])
    n = np.arange(2, N + 1, dtype=float)
    nnk = nnk[1:]
    B1 = 2 / 9
    nk = n + 1
    nnk2 = nnk * nnk
    B = 4 * (n * nk) ** 2 / (nnk2 * nnk2 - nnk2)
    ab = np.column_stack([A, np.concatenate([[2.0, B1], B])])
    s = np.sqrt(ab[1:N, 1])
    J = np.diag(ab[:N, 0]) + np.diag(s, -1) + np.diag(s, 1)
    Xe, V = np.linalg.eigh(J)
    I = np.argsort(Xe, kind='stable')
    Xe = Xe[I]
    x = (Xe + 1) / 2
    wx = ab[0, 1] * V[0, I] ** 2 / 4

    N = N - 1
    N1 = N + 1
    N2 = N + 2
    y = np.cos((2 * np.arange(N, -1, -1, dtype=float) + 1) * np.pi / (2 * N + 2))
    L = np.zeros((N1, N2))
    y0 = 2.0
    Lp = None
    while np.max(np.abs(y - y0)) > np.finfo(float).eps:
        L[:, 0] = 1
        L[:, 1] = y
        for k in range(2, N1 + 1):
            L[:, k] = ((2 * k - 1) * y * L[:, k - 1] - (k - 1) * L[:, k - 2]) / k
        Lp = N2 * (L[:, N1 - 1] - y * L[:, N2 - 1]) / (1 - y ** 2)
        y0 = y
        y = y0 - L[:, N2 - 1] / Lp

    cd = np.array([[1, 0, 0], [-1, 0, 1], [0, 1, -1]], dtype=float) @ v
    t = (1 + y) / 2
    Wx = abs(np.linalg.det(cd[1:3, :])) * wx
    Wy = 1 / ((1 - y ** 2) * Lp ** 2) * (N2 / N1) ** 2
    tt, xx = np.meshgrid(t, x)
    yy = tt * xx
    X = cd[0, 0] + cd[1, 0] * xx + cd[2, 0] * yy
    Y = cd[0, 1] + cd[1, 1] * xx + cd[2, 1] * yy
    return X, Y, Wx, Wy


def quadRule2D(qOrd):
    """Quadrature rule on the reference triangle {(0,0),(1,0),(0,1)};
    core/quadRule2D.m:64-109.  Weights sum to 1/2."""
    if qOrd in (0, 1):
        Q1 = np.array([1 / 3]); Q2 = np.array([1 / 3]); W = np.array([1 / 2])
    elif qOrd == 2:
        Q1 = np.array([1 / 6, 2 / 3, 1 / 6]); Q2 = np.array([1 / 6, 1 / 6, 2 / 3])
        W = np.array([1 / 6, 1 / 6, 1 / 6])
    elif qOrd == 3:
        a1 = (3 + np.sqrt(3)) / 6; a2 = (3 - np.sqrt(3)) / 6
        b1 = (6 + np.sqrt(6)) / 10; b2 = (6 - np.sqrt(6)) / 10
        w1 = 1 / 8 + np.sqrt(6) / 72; w2 = 1 / 8 - np.sqrt(6) / 72
        Q1 = np.array([a1 * b1, a2 * b1, a1 * b2, a2 * b2])
        Q2 = np.array([a2 * b1, a1 * b1, a2 * b2, a1 * b2])
        W = np.array([w1, w1, w2, w2])
    elif qOrd == 4:
        Q1 = np.array([0.445948490915965, 0.108103018168070, 0.445948490915965,
                       0.091576213509771, 0.816847572980458, 0.091576213509771])
        Q2 = np.array([0.108103018168070, 0.445948490915965, 0.445948490915965,
                       0.816847572980458, 0.091576213509771, 0.091576213509771])
        W = np.array([0.111690794839005, 0.111690794839005, 0.111690794839005,
                      0.054975871827661, 0.054975871827661, 0.054975871827661])
    elif qOrd == 5:
        a1 = (6 - np.sqrt(15)) / 21; a2 = (6 + np.sqrt(15)) / 21
        w1 = (155 - np.sqrt(15)) / 2400; w2 = (155 + np.sqrt(15)) / 2400
        Q1 = np.array([1 / 3, a1, 1 - 2 * a1, a1, a2, 1 - 2 * a2, a2])
        Q2 = np.array([1 / 3, 1 - 2 * a1, a1, a1, 1 - 2 * a2, a2, a2])
        W = np.array([9 / 80, w1, w1, w1, w2, w2, w2])
    elif qOrd == 6:
        Q1 = np.array([0.063089014491502, 0.873821971016996, 0.063089014491502,
                       0.249286745170910, 0.501426509658179, 0.249286745170910,
                       0.310352451033785, 0.053145049844816, 0.636502499121399,
                       0.053145049844816, 0.636502499121399, 0.310352451033785])
        Q2 = np.array([0.063089014491502, 0.063089014491502, 0.873821971016996,
                       0.249286745170910, 0.249286745170910, 0.501426509658179,
                       0.053145049844816, 0.310352451033785, 0.053145049844816,
                       0.636502499121399, 0.310352451033785, 0.636502499121399])
        W = np.array([0.025422453185103, 0.025422453185103, 0.025422453185103,
                      0.058393137863189, 0.058393137863189, 0.058393137863189,
                      0.041425537809187, 0.041425537809187, 0.041425537809187,
                      0.041425537809187, 0.041425537809187, 0.041425537809187])
    else:
        X, Y, Wx, Wy = triquad(qOrd)
        Q1 = X.flatten(order='F')
        Q2 = Y.flatten(order='F')
        W = np.outer(Wx, Wy).flatten(order='F')
    return Q1, Q2, W

"""Basis functions, edge maps, basis tables and reference-element blocks
(oracle).  Restates core/phi.m, core/gradPhi.m, core/gammaMap.m,
core/theta.m, core/computeBasesOnQuad.m, core/integrateRefElemPhiPhi.m,
core/integrateRefElemDphiPhiPhi.m, core/integrateRefEdgePhiIntPhiIntPerQuad.m
and core/integrateRefEdgePhiIntPhiExtPerQuad.m.

Pinned by: orthonormality of phi under exact rules (M-hat = I), gradPhi
against central finite differences of phi, theta/gammaMap against the
geometric identity "same physical point seen from the neighbour", and
R-hat / M-hat against independent high-order quadrature.  Ghat for p >= 2 is
integrated with the reference's deliberately inexact rule of order 2p; it is
checked against an independent evaluation of that same rule.
"""
import numpy as np

from .quadrature import quadRule1D, quadRule2D

S2, S3, S5, S6 = np.sqrt(2), np.sqrt(3), np.sqrt(5), np.sqrt(6)
S10, S14, S30, S70 = np.sqrt(10), np.sqrt(14), np.sqrt(30), np.sqrt(70)


def phi(i, X1, X2):
    """Orthonormal Dubiner basis function i (1-based) on the reference
    triangle; core/phi.m:40-61."""
    X1 = np.asarray(X1, dtype=float); X2 = np.asarray(X2, dtype=float)
    if i == 1: return S2 * np.ones_like(X1)
    if i == 2: return 2 - 6 * X1
    if i == 3: return 2 * S3 * (1 - X1 - 2 * X2)
    if i == 4: return S6 * ((10 * X1 - 8) * X1 + 1)
    if i == 5: return S3 * ((5 * X1 - 4) * X1 + (-15 * X2 + 12) * X2 - 1)
    if i == 6: return 3 * S5 * ((3 * X1 + 8 * X2 - 4) * X1 + (3 * X2 - 4) * X2 + 1)
    if i == 7: return 2 * S2 * (-1 + (15 + (-45 + 35 * X1) * X1) * X1)
    if i == 8: return 2 * S6 * (-1 + (13 + (-33 + 21 * X1) * X1) * X1 + (2 + (-24 + 42 * X1) * X1) * X2)
    if i == 9: return 2 * S10 * (-1 + (9 + (-15 + 7 * X1) * X1) * X1 + (6 + (-48 + 42 * X1) * X1 + (-6 + 42 * X1) * X2) * X2)
    if i == 10: return 2 * S14 * (-1 + (3 + (-3 + X1) * X1) * X1 + (12 + (-24 + 12 * X1) * X1 + (-30 + 30 * X1 + 20 * X2) * X2) * X2)
    if i == 11: return S10 * (1 + (-24 + (126 + (-224 + 126 * X1) * X1) * X1) * X1)
    if i == 12: return S30 * (1 + (-22 + (105 + (-168 + 84 * X1) * X1) * X1) * X1 + (-2 + (42 + (-168 + 168 * X1) * X1) * X1) * X2)
    if i == 13: return 5 * S2 * (1 + (-18 + (69 + (-88 + 36 * X1) * X1) * X1) * X1 + (-6 + (102 + (-312 + 216 * X1) * X1) * X1
                                 + (6 + (-96 + 216 * X1) * X1) * X2) * X2)
    if i == 14: return S70 * (1 + (-12 + (30 + (-28 + 9 * X1) * X1) * X1) * X1 + (-12 + (132 + (-228 + 108 * X1) * X1) * X1
                              + (30 + (-300 + 270 * X1) * X1 + (-20 + 180 * X1) * X2) * X2) * X2)
    if i == 15: return 3 * S10 * (1 + (-4 + (6 + (-4 + X1) * X1) * X1) * X1 + (-20 + (60 + (-60 + 20 * X1) * X1) * X1
                                  + (90 + (-180 + 90 * X1) * X1 + (-140 + 140 * X1 + 70 * X2) * X2) * X2) * X2)
    raise ValueError(i)


def gradPhi(i, m, X1, X2):
    """Derivative d/dX_m of basis function i (both 1-based);
    core/gradPhi.m:42-87."""
    X1 = np.asarray(X1, dtype=float); X2 = np.asarray(X2, dtype=float)
    z = np.zeros_like(X1); o = np.ones_like(X1)
    if m == 1:
        if i == 1: return z
        if i == 2: return -6 * o
        if i == 3: return -2 * S3 * o
        if i == 4: return S6 * (20 * X1 - 8)
        if i == 5: return S3 * (10 * X1 - 4)
        if i == 6: return 6 * S5 * (3 * X1 + 4 * X2 - 2)
        if i == 7: return 2 * S2 * (15 + (-90 + 105 * X1) * X1)
        if i == 8: return 2 * S6 * (13 + (-66 + 63 * X1) * X1 + (-24 + 84 * X1) * X2)
        if i == 9: return 2 * S10 * (9 + (-30 + 21 * X1) * X1 + (-48 + 84 * X1 + 42 * X2) * X2)
        if i == 10: return 2 * S14 * (3 + (-6 + 3 * X1) * X1 + (-24 + 24 * X1 + 30 * X2) * X2)
        if i == 11: return S10 * (-24 + (252 + (-672 + 504 * X1) * X1) * X1)
        if i == 12: return S30 * (-22 + (210 + (-504 + 336 * X1) * X1) * X1 + (42 + (-336 + 504 * X1) * X1) * X2)
        if i == 13: return 5 * S2 * (-18 + (138 + (-264 + 144 * X1) * X1) * X1
                                     + (102 + (-624 + 648 * X1) * X1 + (-96 + 432 * X1) * X2) * X2)
        if i == 14: return S70 * (-12 + (60 + (-84 + 36 * X1) * X1) * X1
                                  + (132 + (-456 + 324 * X1) * X1 + (-300 + 540 * X1 + 180 * X2) * X2) * X2)
        if i == 15: return 3 * S10 * (-4 + (12 + (-12 + 4 * X1) * X1) * X1
                                      + (60 + (-120 + 60 * X1) * X1 + (-180 + 180 * X1 + 140 * X2) * X2) * X2)
    else:
        if i == 1: return z
        if i == 2: return z
        if i == 3: return -4 * S3 * o
        if i == 4: return z
        if i == 5: return 2 * S3 * (-15 * X2 + 6)
        if i == 6: return 6 * S5 * (4 * X1 + 3 * X2 - 2)
        if i == 7: return z
        if i == 8: return 2 * S6 * (2 + (-24 + 42 * X1) * X1)
        if i == 9: return 2 * S10 * (6 + (-48 + 42 * X1) * X1 + (-12 + 84 * X1) * X2)
        if i == 10: return 2 * S14 * (12 + (-24 + 12 * X1) * X1 + (-60 + 60 * X1 + 60 * X2) * X2)
        if i == 11: return z
        if i == 12: return S30 * (-2 + (42 + (-168 + 168 * X1) * X1) * X1)
        if i == 13: return 5 * S2 * (-6 + (102 + (-312 + 216 * X1) * X1) * X1
                                     + (12 + (-192 + 432 * X1) * X1) * X2)
        if i == 14: return S70 * (-12 + (132 + (-228 + 108 * X1) * X1) * X1
                                  + (60 + (-600 + 540 * X1) * X1 + (-60 + 540 * X1) * X2) * X2)
        if i == 15: return 3 * S10 * (-20 + (60 + (-60 + 20 * X1) * X1) * X1
                                      + (180 + (-360 + 180 * X1) * X1 + (-420 + 420 * X1 + 280 * X2) * X2) * X2)
    raise ValueError((i, m))


def gammaMap(n, S):
    """Parametrisation of local edge n (1-based) of the reference triangle;
    core/gammaMap.m:50-57."""
    S = np.asarray(S, dtype=float).ravel()
    if n == 1: return 1 - S, S
    if n == 2: return np.zeros_like(S), 1 - S
    if n == 3: return S, np.zeros_like(S)
    raise ValueError(n)


def theta(nn, np_, X1, X2):
    """Map a point on edge nn of the reference triangle to the matching point
    on edge np of the neighbour's reference triangle; core/theta.m:109-133."""
    X1 = np.asarray(X1, dtype=float); X2 = np.asarray(X2, dtype=float)
    z = np.zeros_like(X1)
    if nn == 1:
        if np_ == 1: return 1 - X1, 1 - X2
        if np_ == 2: return z, X2
        if np_ == 3: return X1, z
    if nn == 2:
        if np_ == 1: return 1 - X2, X2
        if np_ == 2: return z, 1 - X2
        if np_ == 3: return X2, z
    if nn == 3:
        if np_ == 1: return X1, 1 - X1
        if np_ == 2: return z, X1
        if np_ == 3: return 1 - X1, z
    raise ValueError((nn, np_))


def p_of_N(N):
    p = int(round((np.sqrt(8 * N + 1) - 3) / 2))
    assert (p + 1) * (p + 2) // 2 == N, 'N does not match a polynomial order'
    return p


def computeBasesOnQuad(N, requiredOrders=None):
    """Tables of basis values on quadrature points; core/computeBasesOnQuad.m:72-134.

    Returns a dict with phi2D[ord] (R x N), gradPhi2D[ord] (R x N x 2),
    phi1D[ord] (R x N x 3), thetaPhi1D[ord] (R x N x 3 x 3), gradPhi1D and
    thetaGradPhi1D, for ord in [2p, 2p+1] (1 for p = 0)."""
    p = p_of_N(N)
    if requiredOrders is None:
        requiredOrders = [2 * p, 2 * p + 1] if p > 0 else [1]
    b = {'phi2D': {}, 'gradPhi2D': {}, 'phi1D': {}, 'thetaPhi1D': {},
         'gradPhi1D': {}, 'thetaGradPhi1D': {}}
    for ord_ in requiredOrders:
        Q1, Q2, _ = quadRule2D(ord_)
        b['phi2D'][ord_] = np.stack([phi(i, Q1, Q2) for i in range(1, N + 1)], axis=1)
        b['gradPhi2D'][ord_] = np.stack(
            [np.stack([gradPhi(i, m, Q1, Q2) for i in range(1, N + 1)], axis=1) for m in (1, 2)], axis=2)
        Q, _ = quadRule1D(ord_)
        R = len(Q)
        p1 = np.zeros((R, N, 3)); g1 = np.zeros((R, N, 3, 2))
        tp = np.zeros((R, N, 3, 3)); tg = np.zeros((R, N, 3, 3, 2))
        for nn in range(1, 4):
            A1, A2 = gammaMap(nn, Q)
            for i in range(1, N + 1):
                p1[:, i - 1, nn - 1] = phi(i, A1, A2)
                for m in (1, 2):
                    g1[:, i - 1, nn - 1, m - 1] = gradPhi(i, m, A1, A2)
            for np_ in range(1, 4):
                P1, P2 = theta(nn, np_, A1, A2)
                for i in range(1, N + 1):
                    tp[:, i - 1, nn - 1, np_ - 1] = phi(i, P1, P2)
                    for m in (1, 2):
                        tg[:, i - 1, nn - 1, np_ - 1, m - 1] = gradPhi(i, m, P1, P2)
        b['phi1D'][ord_] = p1; b['gradPhi1D'][ord_] = g1
        b['thetaPhi1D'][ord_] = tp; b['thetaGradPhi1D'][ord_] = tg
    return b


def integrateRefElemPhiPhi(N, basesOnQuad, qOrd=None):
    """M-hat(i,j) = sum_q W phi_i phi_j; core/integrateRefElemPhiPhi.m:49-59."""
    if qOrd is None:
        qOrd = max(2 * p_of_N(N), 1)
    _, _, W = quadRule2D(qOrd)
    P = basesOnQuad['phi2D'][qOrd]
    ret = np.zeros((N, N))
    for i in range(N):
        for j in range(N):
            ret[i, j] = np.sum(W * P[:, i] * P[:, j])
    return ret


def integrateRefElemDphiPhiPhi(N, basesOnQuad, qOrd=None):
    """G-hat(i,j,l,m) = sum_q W d_m phi_i phi_j phi_l with qOrd = max(2p, 1);
    core/integrateRefElemDphiPhiPhi.m:55-82 (inexact for p >= 2 by design)."""
    ret = np.zeros((N, N, N, 2))
    if N == 1:
        return ret
    if qOrd is None:
        qOrd = max(2 * p_of_N(N), 1)
    _, _, W = quadRule2D(qOrd)
    P = basesOnQuad['phi2D'][qOrd]; G = basesOnQuad['gradPhi2D'][qOrd]
    for i in range(N):
        for j in range(N):
            for l in range(N):
                for m in range(2):
                    ret[i, j, l, m] = np.sum(W * G[:, i, m] * P[:, j] * P[:, l])
    return ret


def integrateRefEdgePhiIntPhiIntPerQuad(N, basesOnQuad, qOrd=None):
    """R-hat-diag(i,j,n,r) = phi_i(gamma_n(q_r)) phi_j(gamma_n(q_r)) w_r;
    core/integrateRefEdgePhiIntPhiIntPerQuad.m:52-66."""
    if qOrd is None:
        qOrd = max(2 * p_of_N(N) + 1, 1)
    _, W = quadRule1D(qOrd)
    P = basesOnQuad['phi1D'][qOrd]
    ret = np.zeros((N, N, 3, len(W)))
    for n in range(3):
        for i in range(N):
            for j in range(N):
                ret[i, j, n, :] = P[:, i, n] * P[:, j, n] * W
    return ret


def integrateRefEdgePhiIntPhiExtPerQuad(N, basesOnQuad, qOrd=None):
    """R-hat-offdiag(i,j,nn,np,r) = phi_i(gamma_nn(q_r)) phi_j(theta(...)) w_r;
    core/integrateRefEdgePhiIntPhiExtPerQuad.m:54-70."""
    if qOrd is None:
        qOrd = 2 * p_of_N(N) + 1
    _, W = quadRule1D(qOrd)
    P = basesOnQuad['phi1D'][qOrd]; T = basesOnQuad['thetaPhi1D'][qOrd]
    ret = np.zeros((N, N, 3, 3, len(W)))
    for nn in range(3):
        for np_ in range(3):
            for i in range(N):
                for j in range(N):
                    ret[i, j, nn, np_, :] = P[:, i, nn] * T[:, j, nn, np_] * W
    return ret

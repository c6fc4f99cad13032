"""Reference gradient for the discrete adjoint (oracle, test infrastructure).

Built from the oracle's own assembled operators: the forward SSP-RK steps
use c_{s+1} = w_s c_0 + (1 - w_s)(c_s + tau M^-1 A c_s) with
A = G1 + G2 - R assembled as in core/assembleMatElemDphiPhiFuncDisc.m and
core/assembleMatEdgePhiPhiValUpwind.m; the reverse sweep applies the sparse
transposes A^T and M^-1; the velocity derivative dA/da_m is assembled with
the same routines from mode m's fields, with the upwind selection of the
edge term frozen at the current velocity (the derivative of v [v > 0] and
v [v < 0] away from v = 0).

The reference has no adjoint; this checks the device adjoint against the
transposes of the reference-structured operators.
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from . import assembly as A
from .advection import rungeKuttaExplicit


def assembleMatEdgePhiPhiValUpwindDerivative(g, markE0Tbdr, hatRdiag, hatRoff, vn, dvn):
    """dR for a change dvn of the normal velocity with the upwind masks of
    vn: core/assembleMatEdgePhiPhiValUpwind.m:261-299 with valOnQuad -> dvn
    in the values and vn in the [v > 0] / [v < 0] selections."""
    K, _, R = vn.shape
    ret = sp.csr_matrix((K * hatRdiag.shape[0], K * hatRdiag.shape[0]))
    for nn in range(3):
        anyNb = (g.markE0TE0T[(nn, 0)] + g.markE0TE0T[(nn, 1)] + g.markE0TE0T[(nn, 2)])
        markElemEdges = (np.asarray(anyNb.sum(axis=1)).ravel() > 0) | (g.idE0T[:, nn] != 0)
        Rkn = markE0Tbdr[:, nn] * g.areaE0T[:, nn]
        for r in range(R):
            d = Rkn * dvn[:, nn, r] * (markElemEdges & (vn[:, nn, r] > 0))
            ret = ret + sp.kron(sp.diags(d), hatRdiag[:, :, nn, r], format='csr')
        for np_ in range(3):
            RknTimesVal = np.zeros((K * hatRdiag.shape[0], hatRdiag.shape[0]))
            for r in range(R):
                RknTimesVal = RknTimesVal + A.kron_col_block(Rkn * dvn[:, nn, r] * (vn[:, nn, r] < 0),
                                                             hatRoff[:, :, nn, np_, r])
            ret = ret + A.kronVec(g.markE0TE0T[(nn, np_)], RknTimesVal)
    return ret.tocsr()


def gradient(pd, u1Disc, u2Disc, vn, modes, c0, nsteps, tau, cObs):
    """J, dJ/dc0 and dJ/da for the oracle problem pd (preprocessed, limiter
    off).  modes: list of (u1Disc_m, u2Disc_m, vn_m)."""
    K, N = pd.K, pd.N
    G = A.assembleMatElemDphiPhiFuncDisc(pd.g, pd.hatG, [u1Disc, u2Disc])
    R = A.assembleMatEdgePhiPhiValUpwind(pd.g, ~pd.g.markE0TbdrN, pd.hatRdiagOnQuad, pd.hatRoffdiagOnQuad, vn)
    Amat = (G[0] + G[1] - R).tocsr()
    Mlu = spla.splu(pd.globM.tocsc())
    _, omega = rungeKuttaExplicit(pd.ordRK, tau, 0.0)
    S = len(omega)
    states = []
    c = c0.reshape(-1).copy()
    for n in range(nsteps):
        st = [c]
        cs = c
        for s in range(S):
            cs = omega[s] * c + (1 - omega[s]) * (cs + tau * Mlu.solve(Amat @ cs))
            if s + 1 < S:
                st.append(cs)
        states.append(st)
        c = cs
    d = c - cObs.reshape(-1)
    J = float(d @ (pd.globM @ d))
    lam = 2 * (pd.globM @ d)
    dA = []
    for (m1, m2, mv) in modes:
        Gm = A.assembleMatElemDphiPhiFuncDisc(pd.g, pd.hatG, [m1, m2])
        Rm = assembleMatEdgePhiPhiValUpwindDerivative(pd.g, ~pd.g.markE0TbdrN, pd.hatRdiagOnQuad,
                                                      pd.hatRoffdiagOnQuad, vn, mv)
        dA.append((Gm[0] + Gm[1] - Rm).tocsr())
    ga = np.zeros(len(modes))
    AT = Amat.T.tocsr()
    for n in range(nsteps - 1, -1, -1):
        st = states[n]
        lam_s = [None] * (S + 1)
        lam_s[S] = lam
        lam0 = np.zeros_like(lam)
        for s in range(S - 1, -1, -1):
            mu = Mlu.solve(lam_s[s + 1])
            for m in range(len(modes)):
                ga[m] += (1 - omega[s]) * tau * (mu @ (dA[m] @ st[s]))
            y = (1 - omega[s]) * (lam_s[s + 1] + tau * (AT @ mu))
            lam0 = lam0 + omega[s] * lam_s[s + 1]
            if s > 0:
                lam_s[s] = y
            else:
                lam0 = lam0 + y
        lam = lam0
    return J, lam.reshape(K, N), ga

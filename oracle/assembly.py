"""Global sparse assembly, projections and evaluation (oracle).  Restates
core/kronVec.m, core/assembleMatElemPhiPhi.m,
core/assembleMatElemDphiPhiFuncDisc.m (vector variant),
core/assembleMatEdgePhiPhiValUpwind.m (no-area variant),
core/assembleVecEdgePhiIntFuncContVal.m, core/assembleVecEdgePhiIntFuncCont.m,
core/projectFuncCont2DataDisc.m, core/computeFuncContNuOnQuadEdge.m,
core/computeFuncContV0T.m, core/computeFuncDiscAtPoints.m and
core/computeL2Error.m.

The operators are built exactly as the reference builds them: KN x KN sparse
matrices from ``kron`` of per-element coefficient columns with reference
blocks, ``kronVec`` with the neighbour masks, and element-major KN vectors.

Pinned by: brute-force per-edge / per-element quadrature on small meshes
against the assembled matrices, a second (matrix-free) oracle path, the
discrete mass balance of the i = 1 row, free-stream preservation (c = 1 with
c_D = 1 and divergence-free linear u stays 1), and the L2 projection /
L2 error of polynomials of degree <= p (exact to rounding).
"""
import numpy as np
import scipy.sparse as sp

from .basis import gammaMap, p_of_N
from .grid import mapRef2Phy
from .quadrature import quadRule1D, quadRule2D


def kronVec(A, B):
    """Block Kronecker of a K x K (sparse) pattern A with a KN x N stack of
    per-row blocks B: block (k, k') = A(k, k') * B(k-th block); core/kronVec.m:78-104
    (sparse branch :94-102)."""
    A = sp.csr_matrix(A)
    ma, na = A.shape
    mb, nb = B.shape
    mc = mb // ma
    Ac = A.tocoo()
    # find(kron(A, ones(mc, 1))): every nonzero (i, j) of A expands to rows
    # i*mc .. i*mc+mc-1 in column j
    i2 = (Ac.row[:, None] * mc + np.arange(mc)[None, :]).ravel()
    j2 = np.repeat(Ac.col, mc)
    v2 = np.repeat(Ac.data, mc)
    ik = np.repeat(i2, nb)
    jk = (nb * j2[:, None] + np.arange(nb)[None, :]).ravel()
    Bd = B.toarray() if sp.issparse(B) else np.asarray(B)
    sk = (v2[:, None] * Bd[i2, :]).ravel()
    return sp.csr_matrix((sk, (ik, jk)), shape=(mb, na * nb))


def kron_col_block(col, blk):
    """kron(col, blk) for a K x 1 column and an N x N block -> KN x N dense."""
    return (col[:, None, None] * blk[None, :, :]).reshape(-1, blk.shape[1])


def assembleMatElemPhiPhi(g, refElemPhiPhi):
    """globM = kron(diag(detJ0T), M-hat); core/assembleMatElemPhiPhi.m:88-100."""
    return sp.kron(sp.diags(g.detJ0T), refElemPhiPhi, format='csr')


def assembleMatElemDphiPhiFuncDisc(g, refElemDphiPhiPhi, dataDisc):
    """{G1, G2} for a velocity given as {u1Disc, u2Disc} (K x N each);
    vector variant core/assembleMatElemDphiPhiFuncDisc.m:181-202."""
    K = g.numT
    N = refElemDphiPhiPhi.shape[0]
    dataN = dataDisc[0].shape[1]
    J = g.B
    ret = []
    for m in range(2):
        Rs = np.zeros((K * N, N))
        mo = 1 - m   # Matlab 3-m, 0-based
        for l in range(dataN):
            Rs = Rs + kron_col_block(dataDisc[m][:, l] * J[:, mo, mo], refElemDphiPhiPhi[:, :, l, m]) \
                    - kron_col_block(dataDisc[m][:, l] * J[:, mo, m], refElemDphiPhiPhi[:, :, l, mo])
        ret.append(kronVec(sp.identity(K, format='csr'), Rs))
    return ret


def assembleMatEdgePhiPhiValUpwind(g, markE0Tbdr, refEdgePhiIntPhiIntOnQuad,
                                    refEdgePhiIntPhiExtOnQuad, valOnQuad):
    """Upwind edge operator R; core/assembleMatEdgePhiPhiValUpwind.m:194-217 and
    the no-area variant :261-299 (markElem all true)."""
    K, _, R = valOnQuad.shape
    N = refEdgePhiIntPhiIntOnQuad.shape[0]
    ret = sp.csr_matrix((K * N, K * N))
    for nn in range(3):
        anyNb = (g.markE0TE0T[(nn, 0)] + g.markE0TE0T[(nn, 1)] + g.markE0TE0T[(nn, 2)])
        markElemEdges = (np.asarray(anyNb.sum(axis=1)).ravel() > 0) | (g.idE0T[:, nn] != 0)
        Rkn = markE0Tbdr[:, nn] * g.areaE0T[:, nn]
        for r in range(R):
            v = valOnQuad[:, nn, r]
            d = Rkn * v * (markElemEdges & (v > 0))
            ret = ret + sp.kron(sp.diags(d), refEdgePhiIntPhiIntOnQuad[:, :, nn, r], format='csr')
        for np_ in range(3):
            RknTimesVal = np.zeros((K * N, N))
            for r in range(R):
                v = valOnQuad[:, nn, r]
                RknTimesVal = RknTimesVal + kron_col_block(Rkn * v * (v < 0), refEdgePhiIntPhiExtOnQuad[:, :, nn, np_, r])
            ret = ret + kronVec(g.markE0TE0T[(nn, np_)], RknTimesVal)
    return ret.tocsr()


def assembleVecEdgePhiIntFuncContVal(g, markE0Tbdr, funcCont, valOnQuad, N, basesOnQuad, qOrd=None):
    """Dirichlet inflow vector K_D (element-major KN); no-area variant of
    core/assembleVecEdgePhiIntFuncContVal.m:162-185."""
    if qOrd is None:
        qOrd = 2 * p_of_N(N) + 1
    Q, W = quadRule1D(qOrd)
    ret = np.zeros((g.numT, N))
    for n in range(3):
        Q1, Q2 = gammaMap(n + 1, Q)
        funcOnQuad = funcCont(mapRef2Phy(g, 0, Q1, Q2), mapRef2Phy(g, 1, Q1, Q2))
        Kkn = markE0Tbdr[:, n] * g.areaE0T[:, n]
        v = valOnQuad[:, n, :]
        for i in range(N):
            integral = (funcOnQuad * ((v < 0) * v)) @ (W * basesOnQuad['phi1D'][qOrd][:, i, n])
            ret[:, i] = ret[:, i] + Kkn * integral
    return ret.reshape(-1)


def assembleVecEdgePhiIntFuncCont(g, markE0T, funcCont, N, basesOnQuad, qOrd=None):
    """Neumann vector K_N (element-major KN); core/assembleVecEdgePhiIntFuncCont.m:86-136."""
    if qOrd is None:
        qOrd = 2 * p_of_N(N) + 1
    Q, W = quadRule1D(qOrd)
    ret = np.zeros((g.numT, N))
    for n in range(3):
        Q1, Q2 = gammaMap(n + 1, Q)
        funcQ0E = funcCont(mapRef2Phy(g, 0, Q1, Q2), mapRef2Phy(g, 1, Q1, Q2))
        markCoefE0T = markE0T[:, n] * g.areaE0T[:, n]
        for i in range(N):
            ret[:, i] = ret[:, i] + markCoefE0T * (funcQ0E @ (W * basesOnQuad['phi1D'][qOrd][:, i, n]))
    return ret.reshape(-1)


def projectFuncCont2DataDisc(g, funcCont, qOrd, refElemPhiPhi, basesOnQuad):
    """L2 projection of a continuous function; core/projectFuncCont2DataDisc.m:94-102.
    Returns K x N."""
    qOrd = max(qOrd, 1)
    Q1, Q2, W = quadRule2D(qOrd)
    N = refElemPhiPhi.shape[0]
    rhs = funcCont(mapRef2Phy(g, 0, Q1, Q2), mapRef2Phy(g, 1, Q1, Q2)) @ (W[:, None] * basesOnQuad['phi2D'][qOrd][:, :N])
    # rhs / refElemPhiPhi  ==  rhs * inv(refElemPhiPhi)
    return np.linalg.solve(refElemPhiPhi.T, rhs.T).T


def computeFuncContNuOnQuadEdge(g, funcCont1, funcCont2, qOrd):
    """u . nu at the edge quadrature points of every element, evaluated from
    each element's own side; core/computeFuncContNuOnQuadEdge.m:46-67.
    Returns K x 3 x R."""
    Q, W = quadRule1D(qOrd)
    ret = np.zeros((g.numT, 3, len(W)))
    for n in range(3):
        Q1, Q2 = gammaMap(n + 1, Q)
        X1 = mapRef2Phy(g, 0, Q1, Q2); X2 = mapRef2Phy(g, 1, Q1, Q2)
        ret[:, n, :] = g.nuE0T[:, n, 0][:, None] * funcCont1(X1, X2) + g.nuE0T[:, n, 1][:, None] * funcCont2(X1, X2)
    return ret


def computeFuncContV0T(g, funcCont):
    """Function values at the three vertices of every element;
    core/computeFuncContV0T.m:41-51."""
    return np.column_stack([funcCont(g.coordV0T[:, n, 0], g.coordV0T[:, n, 1]) for n in range(3)])


def computeFuncDiscAtPoints(funcDisc, phiAtPoints):
    """ret(k, i) = sum_j funcDisc(k, j) phiAtPoints(k, i, j);
    core/computeFuncDiscAtPoints.m:43-55."""
    return np.einsum('kj,kij->ki', funcDisc, phiAtPoints)


def computeL2Error(g, dataDisc, funcCont, qOrd, basesOnQuad):
    """sqrt(2 sum_k |T_k| sum_r W_r (c_h - c)^2); core/computeL2Error.m:97-114."""
    qOrd = max(qOrd, 1)
    Q1, Q2, W = quadRule2D(qOrd)
    N = dataDisc.shape[1]
    X1 = np.outer(g.B[:, 0, 0], Q1) + np.outer(g.B[:, 0, 1], Q2) + g.coordV0T[:, 0, 0][:, None]
    X2 = np.outer(g.B[:, 1, 0], Q1) + np.outer(g.B[:, 1, 1], Q2) + g.coordV0T[:, 0, 1][:, None]
    cEx = funcCont(X1, X2)
    cAp = dataDisc @ basesOnQuad['phi2D'][qOrd][:, :N].T
    return float(np.sqrt(2 * np.dot(((cAp - cEx) ** 2) @ W, g.areaT)))

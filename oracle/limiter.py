"""Taylor-basis vertex-based slope limiter (oracle).  Restates
core/phiTaylorPhy.m, core/phiTaylorRef.m, core/computeTaylorBasesV0T.m,
core/assembleMatElemPhiTaylorPhiTaylor.m, core/assembleMatElemPhiDiscPhiTaylor.m,
core/projectDataDisc2DataTaylor.m, core/projectDataTaylor2DataDisc.m,
core/computeMinMaxV0TElementPatch.m, core/computeVertexBasedLimiter.m,
core/applySlopeLimiterTaylorLinear.m,
core/applySlopeLimiterTaylorHierarchicalVertex.m,
core/applySlopeLimiterTaylorStrict.m, core/multiindex.m, core/mult2ind.m,
core/applySlopeLimiterTaylor.m, core/applySlopeLimiterDisc.m,
core/projectDataDisc2DataLagr.m and core/applySlopeLimiterDiscLinear.m.

Pinned by: brute-force patch min/max on small meshes (dense markV0TV0T path),
the invariants of the limiter (cell means unchanged; alpha in [0, 1]; linear
data that respects the patch bounds is untouched), the exactness of the
Disc <-> Taylor change of basis on polynomials (round trip is the identity to
rounding, and Taylor coefficients equal centroid derivatives scaled by dX),
and the mean-free property of the Taylor basis.  The hierarchical and strict
types are pinned by their reduction to the linear limiter at p = 1, by a
loop-based evaluation on small meshes that forms every patch, vertex value
and Taylor basis value directly from coordinates, and by their invariants
(cell means unchanged, data inside its bounds untouched).  The limiter of
the DG basis' linear part is pinned by its vertex values reproducing a
linear function at the mesh vertices from coordinates, its centroid value
reproducing the function at the centroid, its agreement with the Taylor
'linear' limiter at p = 1 (where both limit the same linear polynomial) and
its invariants (c_1 kept, higher coefficients dropped exactly where alpha < 1).
"""
import numpy as np
import scipy.sparse as sp
import scipy.sparse.linalg as spla

from .basis import p_of_N, phi
from .grid import mapRef2Phy
from .quadrature import quadRule2D
from .assembly import computeFuncDiscAtPoints

# (a, b) exponents of Taylor basis function i (1-based), core/phiTaylorPhy.m:83-131
TAYLOR_EXP = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3),
              (4, 0), (3, 1), (2, 2), (1, 3), (0, 4)]


def phiTaylorPhy(g, i, X1, X2):
    """Physical Taylor basis function i at K x numP physical points;
    core/phiTaylorPhy.m:70-135 (mean-free monomials about the centroid,
    scaled by dX = 2 / sqrt(2 |T|))."""
    X1 = np.asarray(X1, dtype=float); X2 = np.asarray(X2, dtype=float)
    K = g.numT
    numP = X1.shape[1]
    qOrd = int(np.ceil((np.sqrt(8 * i + 1) - 3) / 2))
    Q1, Q2, W = quadRule2D(qOrd)
    dX = np.repeat((2 / np.sqrt(2 * g.areaT))[:, None], numP, axis=1)
    dx = X1 - g.baryT[:, 0][:, None]
    dy = X2 - g.baryT[:, 1][:, None]
    if i == 1:
        return np.ones((K, numP))
    if i == 2:
        return dx * dX
    if i == 3:
        return dy * dX
    qx = mapRef2Phy(g, 0, Q1, Q2) - g.baryT[:, 0][:, None]
    qy = mapRef2Phy(g, 1, Q1, Q2) - g.baryT[:, 1][:, None]
    if i == 4: return (0.5 * dx ** 2 - ((qx ** 2) @ W)[:, None]) * dX ** 2
    if i == 5: return (dx * dy - (2 * (qx * qy) @ W)[:, None]) * dX ** 2
    if i == 6: return (0.5 * dy ** 2 - ((qy ** 2) @ W)[:, None]) * dX ** 2
    if i == 7: return (dx ** 3 / 6 - ((qx ** 3) @ W / 3)[:, None]) * dX ** 3
    if i == 8: return (0.5 * dx ** 2 * dy - ((qx ** 2 * qy) @ W)[:, None]) * dX ** 3
    if i == 9: return (0.5 * dx * dy ** 2 - ((qx * qy ** 2) @ W)[:, None]) * dX ** 3
    if i == 10: return (dy ** 3 / 6 - ((qy ** 3) @ W / 3)[:, None]) * dX ** 3
    if i == 11: return (dx ** 4 / 24 - ((qx ** 4) @ W / 12)[:, None]) * dX ** 4
    if i == 12: return (dx ** 3 * dy / 6 - ((qx ** 3 * qy) @ W / 3)[:, None]) * dX ** 4
    if i == 13: return (dx ** 2 * dy ** 2 / 4 - ((qx ** 2 * qy ** 2) @ W / 2)[:, None]) * dX ** 4
    if i == 14: return (dx * dy ** 3 / 6 - ((qx * qy ** 3) @ W / 3)[:, None]) * dX ** 4
    if i == 15: return (dy ** 4 / 24 - ((qy ** 4) @ W / 12)[:, None]) * dX ** 4
    raise ValueError('Taylor basis functions for p>4 not implemented')


def phiTaylorRef(g, i, hatX1, hatX2):
    """core/phiTaylorRef.m:51-53."""
    return phiTaylorPhy(g, i, mapRef2Phy(g, 0, hatX1, hatX2), mapRef2Phy(g, 1, hatX1, hatX2))


def computeTaylorBasesV0T(g, N, basesOnQuad):
    """phiTaylorV0T (K x 3 x N): Taylor basis at the element vertices;
    core/computeTaylorBasesV0T.m:50-65."""
    out = np.zeros((g.numT, 3, N))
    for n in range(3):
        for i in range(1, N + 1):
            out[:, n, i - 1] = phiTaylorPhy(g, i, g.coordV0T[:, n, 0][:, None], g.coordV0T[:, n, 1][:, None])[:, 0]
    basesOnQuad['phiTaylorV0T'] = out
    return basesOnQuad


def _blockdiag_from_entries(K, N, ent):
    """ent[i][j] is a K-vector: block-diagonal KN x KN with block(k)(i, j)."""
    rows = []; cols = []; vals = []
    base = np.arange(K) * N
    for i in range(N):
        for j in range(N):
            rows.append(base + i); cols.append(base + j); vals.append(ent[i][j])
    return sp.csr_matrix((np.concatenate(vals), (np.concatenate(rows), np.concatenate(cols))), shape=(K * N, K * N))


def assembleMatElemPhiTaylorPhiTaylor(g, N):
    """Block-diagonal Taylor mass matrix; core/assembleMatElemPhiTaylorPhiTaylor.m:50-64."""
    p = p_of_N(N)
    Q1, Q2, W = quadRule2D(max(2 * p, 1))
    T = [phiTaylorRef(g, i, Q1, Q2) for i in range(1, N + 1)]
    ent = [[2 * g.areaT * ((T[i] * T[j]) @ W) for j in range(N)] for i in range(N)]
    return _blockdiag_from_entries(g.numT, N, ent)


def assembleMatElemPhiDiscPhiTaylor(g, N, basesOnQuad):
    """Block-diagonal mixed mass matrix (test: DG basis, trial: Taylor basis);
    core/assembleMatElemPhiDiscPhiTaylor.m:68-78."""
    p = p_of_N(N)
    qOrd = max(2 * p + 1, 1)
    Q1, Q2, W = quadRule2D(qOrd)
    T = [phiTaylorRef(g, j, Q1, Q2) for j in range(1, N + 1)]
    P = basesOnQuad['phi2D'][qOrd]
    ent = [[2 * g.areaT * ((P[:, i][None, :] * T[j]) @ W) for j in range(N)] for i in range(N)]
    return _blockdiag_from_entries(g.numT, N, ent)


def projectDataDisc2DataTaylor(dataDisc, globMDisc, globMDiscTaylor):
    """core/projectDataDisc2DataTaylor.m:56-61 (sparse direct solve)."""
    K, N = dataDisc.shape
    x = spla.spsolve(globMDiscTaylor.tocsc(), globMDisc @ dataDisc.reshape(-1))
    return x.reshape(K, N)


def projectDataTaylor2DataDisc(dataTaylor, globMDisc, globMDiscTaylor):
    """core/projectDataTaylor2DataDisc.m:56-61 (sparse direct solve)."""
    K, N = dataTaylor.shape
    x = spla.spsolve(globMDisc.tocsc(), globMDiscTaylor @ dataTaylor.reshape(-1))
    return x.reshape(K, N)


def _sparse_row_min(M, vals):
    """min over the stored entries of each row of M .* vals' (all entries of
    a row of markV0TT0V are ones and the shifted values are < 0, so the
    implicit zeros never win; core/computeMinMaxV0TElementPatch.m:107)."""
    M = M.tocsr()
    data = vals[M.indices]
    return np.minimum.reduceat(data, M.indptr[:-1])


def _sparse_row_max(M, vals):
    M = M.tocsr()
    data = vals[M.indices]
    return np.maximum.reduceat(data, M.indptr[:-1])


def computeMinMaxV0TElementPatch(g, valCentroid, markV0TbdrD, dataV0T):
    """Per (k, n): min / max of the centroid values over the patch of vertex
    V0T(k, n), with the reference's global shift trick and the Dirichlet
    vertex value joined NaN-ignoring; core/computeMinMaxV0TElementPatch.m:66-112."""
    shiftPos = abs(np.min(valCentroid)) + 1
    shiftNeg = abs(np.max(valCentroid)) + 1
    vPos = valCentroid + shiftPos
    vNeg = valCentroid - shiftNeg
    valD = np.full((g.numT, 3), np.nan)
    valD[markV0TbdrD] = dataV0T[markV0TbdrD]
    mn = np.zeros((g.numT, 3)); mx = np.zeros((g.numT, 3))
    for i in range(3):
        mn[:, i] = np.fmin(_sparse_row_min(g.markV0TT0V[i], vNeg) + shiftNeg, valD[:, i])
        mx[:, i] = np.fmax(_sparse_row_max(g.markV0TT0V[i], vPos) - shiftPos, valD[:, i])
    return mn, mx


def computeVertexBasedLimiter(g, valCentroid, valV0T, markV0TbdrD, dataV0T):
    """alpha per element; core/computeVertexBasedLimiter.m:78-106."""
    mn, mx = computeMinMaxV0TElementPatch(g, valCentroid, markV0TbdrD, dataV0T)
    d = valV0T - valCentroid[:, None]
    dMin = mn - valCentroid[:, None]
    dMax = mx - valCentroid[:, None]
    tol = 1.e-8
    markNeg = d < dMin + tol
    markPos = d > dMax - tol
    a = np.ones((g.numT, 3))
    with np.errstate(divide='ignore', invalid='ignore'):
        a[markNeg] = np.fmax(0, np.fmin(1, dMin[markNeg] / (d[markNeg] - tol)))
        a[markPos] = np.fmax(0, np.fmin(1, dMax[markPos] / (d[markPos] + tol)))
    return np.min(a, axis=1), (mn, mx)


def applySlopeLimiterTaylorLinear(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad):
    """core/applySlopeLimiterTaylorLinear.m:108-122."""
    valV0T = computeFuncDiscAtPoints(dataTaylor[:, 0:3], basesOnQuad['phiTaylorV0T'][:, :, 0:3])
    alphaE, minMax = computeVertexBasedLimiter(g, dataTaylor[:, 0], valV0T, markV0TbdrD, dataV0T)
    lim = np.column_stack([dataTaylor[:, 0], alphaE[:, None] * dataTaylor[:, 1:3],
                           (alphaE == 1)[:, None] * dataTaylor[:, 3:]])
    return lim, minMax


def multiindex(p):
    """Multi-indices (a, b) of total degree <= p in the Taylor order;
    core/multiindex.m:52-61."""
    mult = [(0, 0)]
    for o in range(1, p + 1):
        for i in range(1, o + 2):
            mult.append((o - i + 1, i - 1))
    return np.array(mult, dtype=int)


def mult2ind(a):
    """1-based linear Taylor index of multi-indices; core/mult2ind.m:49-53."""
    a = np.atleast_2d(a)
    p = a.sum(axis=1)
    return p * (p + 1) // 2 + 1 + a[:, 1]


def applySlopeLimiterTaylorHierarchicalVertex(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad):
    """core/applySlopeLimiterTaylorHierarchicalVertex.m:96-131: for ord = p..1
    the linear reconstructions of the degree-(ord-1) coefficients give
    alphaOrd; alpha = max(alpha, alphaOrd) scales the degree-ord block."""
    K, N = dataTaylor.shape
    p = p_of_N(N)
    lim = np.zeros_like(dataTaylor)
    lim[:, 0] = dataTaylor[:, 0]
    alpha = np.zeros(K)
    minMax = None
    for o in range(p, 0, -1):
        alphaOrd = np.ones(K)
        indDOF = np.arange(o * (o + 1) // 2, (o + 1) * (o + 2) // 2)
        for i in range(1, o + 1):
            ind = mult2ind(np.array([o - i, i - 1]) + multiindex(1)) - 1
            valV0T = computeFuncDiscAtPoints(dataTaylor[:, ind], basesOnQuad['phiTaylorV0T'][:, :, 0:3])
            dV = valV0T if o > 1 else dataV0T
            alphaTmp, minMax = computeVertexBasedLimiter(g, dataTaylor[:, ind[0]], valV0T, markV0TbdrD, dV)
            alphaOrd = np.minimum(alphaOrd, alphaTmp)
        alpha = np.maximum(alpha, alphaOrd)
        lim[:, indDOF] = alpha[:, None] * dataTaylor[:, indDOF]
    return lim, minMax


def applySlopeLimiterTaylorStrict(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad):
    """core/applySlopeLimiterTaylorStrict.m:93-127: for ord = p..1 the
    reconstructions of degree p - ord + 1 about the degree-(ord-1)
    coefficients (from the data limited so far) give alphaOrd, which scales
    every coefficient of degree >= ord."""
    K, N = dataTaylor.shape
    p = p_of_N(N)
    lim = dataTaylor.copy()
    minMax = None
    for o in range(p, 0, -1):
        alphaOrd = np.ones(K)
        indDOF = np.arange(o * (o + 1) // 2, N)
        for i in range(1, o + 1):
            pR = p - o + 1
            ind = mult2ind(np.array([o - i, i - 1]) + multiindex(pR)) - 1
            valV0T = computeFuncDiscAtPoints(lim[:, ind], basesOnQuad['phiTaylorV0T'][:, :, 0:len(ind)])
            dV = valV0T if o > 1 else dataV0T
            alphaTmp, minMax = computeVertexBasedLimiter(g, lim[:, ind[0]], valV0T, markV0TbdrD, dV)
            alphaOrd = np.minimum(alphaOrd, alphaTmp)
        lim[:, indDOF] = alphaOrd[:, None] * lim[:, indDOF]
    return lim, minMax


def applySlopeLimiterTaylor(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad, type_='linear'):
    """core/applySlopeLimiterTaylor.m:69-85."""
    if type_ == 'linear':
        return applySlopeLimiterTaylorLinear(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad)
    if type_ == 'hierarch_vert':
        return applySlopeLimiterTaylorHierarchicalVertex(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad)
    if type_ == 'strict':
        return applySlopeLimiterTaylorStrict(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad)
    raise ValueError('Unknown limiter type')


def applySlopeLimiterDisc(g, dataDisc, markV0TbdrD, dataV0T, globM, globMDiscTaylor, basesOnQuad, type_='linear'):
    """core/applySlopeLimiterDisc.m:77-86."""
    dataTaylor = projectDataDisc2DataTaylor(dataDisc, globM, globMDiscTaylor)
    dataTaylor, minMax = applySlopeLimiterTaylor(g, dataTaylor, markV0TbdrD, dataV0T, basesOnQuad, type_)
    return projectDataTaylor2DataDisc(dataTaylor, globM, globMDiscTaylor), minMax


def projectDataDisc2DataLagr(dataDisc, p=None):
    """DG data -> values at the Lagrange nodes of degree min(p, 2) on every
    element, products summed in basis order; core/projectDataDisc2DataLagr.m:43-57."""
    K, N = dataDisc.shape
    if p is None:
        p = int(round((np.sqrt(8 * N + 1) - 3) / 2))
    if p == 0:
        L1, L2 = [1 / 3], [1 / 3]
    elif p == 1:
        L1, L2 = [0, 1, 0], [0, 0, 1]
    else:
        L1, L2 = [0, 1, 0, 1 / 2, 0, 1 / 2], [0, 0, 1, 1 / 2, 1 / 2, 0]
    out = np.zeros((K, len(L1)))
    for i in range(N):
        out = out + dataDisc[:, i:i + 1] * phi(i + 1, L1, L2)[None, :]
    return out


def applySlopeLimiterDiscLinear(g, dataDisc, markV0TbdrD, dataV0T):
    """Vertex-based limiter of the DG basis' linear part;
    core/applySlopeLimiterDiscLinear.m:109-123."""
    assert dataDisc.shape[1] >= 3, 'needs p >= 1'
    valV0T = projectDataDisc2DataLagr(dataDisc[:, 0:3])
    alphaE, minMax = computeVertexBasedLimiter(g, dataDisc[:, 0] * phi(1, 1 / 3, 1 / 3), valV0T, markV0TbdrD, dataV0T)
    lim = np.column_stack([dataDisc[:, 0], alphaE[:, None] * dataDisc[:, 1:3],
                           (alphaE == 1)[:, None] * dataDisc[:, 3:]])
    return lim, minMax, alphaE

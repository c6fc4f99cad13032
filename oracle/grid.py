"""Mesh generation and connectivity (oracle).  Restates core/domainSquare.m,
core/generateGridData.m, advection/preprocessProblem.m:72-76 and
advection/computeDerivedGridData.m.

Matlab semantics kept: ``sparse`` sums duplicates, ``find`` returns entries in
column-major order, and repeated indexed assignment is last-write-wins (here
written as "first occurrence in forward triangle order", which is the same
thing for the reverse-order writes of generateGridData.m:142-148).

The dense K x K ``markV0TV0T`` temporaries of generateGridData.m:187 are only
built when ``dense_patches=True`` (small meshes, used by the tests as the
brute-force check); otherwise the vertex patches ``markV0TT0V`` are built as
sparse incidence products with the same pattern.

Pinned by: closed forms (numV = (dim+1)^2, numE = 3 dim^2 + 2 dim,
K = 2 dim^2, Euler's formula, sum of areas = 1), outward unit normals,
consistency of E0T/T0E/E0E/markE0TE0T, a brute-force neighbour search on
small meshes, the dense markV0TV0T path against the sparse one, and the
dim = 3 numbering drawn in the reference's doxygen/images/domainSquare.png
(transcribed into the tests as data).
"""
import numpy as np
import scipy.sparse as sp


class Grid(dict):
    """Attribute-access dict mirroring Matlab's struct ``g``."""
    __getattr__ = dict.__getitem__

    def __setattr__(self, k, v):
        self[k] = v


def matlab_colon(a, h, b):
    """Values of Matlab's ``a:h:b`` for a positive step.

    Matlab documents that colon ranges are computed symmetrically
    (a + k h for the first half, b - (n - k) h for the second half).  For the
    dyadic steps of every configuration here both halves are exact; for
    non-dyadic h the last-ulp values are parity unpinned."""
    n = int(np.floor((b - a) / h + 1e-10))
    k = np.arange(n + 1, dtype=float)
    x = a + k * h
    half = (n + 1) // 2
    x[half:] = b - (n - k[half:]) * h
    return x


def domainSquare(h, a=0.0, b=1.0):
    """Friedrichs-Keller triangulation of [a,b]^2; core/domainSquare.m:67-105.

    Returns coordV (numV x 2) and V0T (K x 3, 0-based), then the full grid
    via generateGridData plus idE / idE0T boundary ids (1 S, 2 E, 3 N, 4 W)."""
    dim = int(np.ceil((b - a) / h))
    h = (b - a) / dim
    xs = matlab_colon(a, h, b)
    X, Y = np.meshgrid(xs, xs)
    coordV = np.column_stack([X.flatten(order='F'), Y.flatten(order='F')])
    n = dim * (dim + 1)
    off = np.arange(n)[:, None]
    pat1 = np.array([1, dim + 2, 2]) - 1
    pat2 = np.array([dim + 2, dim + 3, 2]) - 1
    keep = np.ones(n, dtype=bool)
    keep[np.arange(dim, n, dim + 1)] = False          # rows (dim+1):(dim+1):dim(dim+1), 1-based
    V0T1 = (pat1[None, :] + off)[keep]
    V0T2 = (pat2[None, :] + off)[keep]
    g = generateGridData(coordV, np.vstack([V0T1, V0T2]))
    g.dim = dim
    idE = np.zeros(g.numE, dtype=np.int64)
    tol = 10 * np.finfo(float).eps
    idE[np.abs(g.baryE[:, 1] - a) < tol] = 1
    idE[np.abs(g.baryE[:, 0] - b) < tol] = 2
    idE[np.abs(g.baryE[:, 1] - b) < tol] = 3
    idE[np.abs(g.baryE[:, 0] - a) < tol] = 4
    g.idE = idE
    g.idE0T = idE[g.E0T]
    return g


def generateGridData(coordV, V0T, dense_patches=False):
    """Connectivity and geometry lists; core/generateGridData.m:130-198."""
    coordV = np.asarray(coordV, dtype=float)
    V0T = np.asarray(V0T, dtype=np.int64)
    g = Grid()
    g.coordV = coordV
    g.V0T = V0T
    K = V0T.shape[0]; numV = coordV.shape[0]
    g.numT = K; g.numV = numV
    tid = np.arange(1, K + 1, dtype=float)                 # 1-based ids: 0 means "absent" in sparse
    # V2T(a, b) = k for directed edge a->b of triangle k (generateGridData.m:136-137)
    rows = V0T[:, [0, 1, 2]].ravel(); cols = V0T[:, [1, 2, 0]].ravel()
    V2T = sp.csr_matrix((np.repeat(tid, 3), (rows, cols)), shape=(numV, numV))
    V2T.sum_duplicates()
    # edge numbering: find(triu(V2T + V2T')) in column-major order (:139-141)
    S = (V2T + V2T.T).tocoo()
    m = S.row <= S.col
    r = S.row[m]; c = S.col[m]
    order = np.lexsort((r, c))
    r = r[order]; c = c[order]
    numE = r.size
    V2E = sp.csr_matrix((np.arange(1, numE + 1, dtype=float), (r, c)), shape=(numV, numV))
    V2E = (V2E + V2E.T).tocsr()

    def lookup(Msp, a, b):
        return np.asarray(Msp[a.ravel(), b.ravel()]).ravel().reshape(a.shape)

    # V0E / T0E via reverse-order last-write-wins (:142-148) == first occurrence
    # in forward order
    A = V0T[:, [0, 1, 2]]; B = V0T[:, [1, 2, 0]]
    idx = lookup(V2E, A, B).astype(np.int64) - 1           # K x 3 edge ids
    flat = idx.ravel()
    _, first = np.unique(flat, return_index=True)
    e = flat[first]
    V0E = np.zeros((numE, 2), dtype=np.int64)
    V0E[e, 0] = A.ravel()[first]
    V0E[e, 1] = B.ravel()[first]
    T0E = np.zeros((numE, 2), dtype=np.int64)
    T0E[e, 0] = lookup(V2T, A.ravel()[first], B.ravel()[first]).astype(np.int64) - 1
    T0E[e, 1] = lookup(V2T, B.ravel()[first], A.ravel()[first]).astype(np.int64) - 1
    g.V0E = V0E; g.T0E = T0E; g.numE = numE
    vecE = coordV[V0E[:, 1], :] - coordV[V0E[:, 0], :]
    g.areaE = np.sqrt(vecE[:, 0] ** 2 + vecE[:, 1] ** 2)
    g.nuE = np.column_stack([vecE[:, 1], -vecE[:, 0]]) / g.areaE[:, None]
    x = coordV[:, 0]; y = coordV[:, 1]
    v1, v2, v3 = V0T[:, 0], V0T[:, 1], V0T[:, 2]
    g.areaT = (x[v1] * y[v2] + x[v2] * y[v3] + x[v3] * y[v1]
               - x[v1] * y[v3] - x[v2] * y[v1] - x[v3] * y[v2]) / 2
    g.baryT = (coordV[v1, :] + coordV[v2, :] + coordV[v3, :]) / 3
    # E0T(k, n) = edge opposite vertex n (:158)
    g.E0T = lookup(V2E, V0T[:, [1, 2, 0]], V0T[:, [2, 0, 1]]).astype(np.int64) - 1
    g.areaE0T = g.areaE[g.E0T]
    sigE0T = 1 - 2 * (T0E[g.E0T, 1] == np.arange(K)[:, None]).astype(float)
    g.sigE0T = sigE0T
    g.baryE = 0.5 * (coordV[V0E[:, 0], :] + coordV[V0E[:, 1], :])
    g.coordV0T = np.stack([coordV[V0T[:, n], :] for n in range(3)], axis=1)       # K x 3 x 2
    g.baryE0T = np.stack([g.baryE[g.E0T[:, n], :] for n in range(3)], axis=1)
    g.nuE0T = np.stack([g.nuE[g.E0T[:, n], :] * sigE0T[:, n][:, None] for n in range(3)], axis=1)
    E0E = np.full((numE, 2), -1, dtype=np.int64)            # local edge ids, -1 = none
    for n in range(3):
        Tn = sigE0T[:, n] == 1; Tp = ~Tn
        E0E[g.E0T[Tn, n], 0] = n
        E0E[g.E0T[Tp, n], 1] = n
    g.E0E = E0E
    Bm = np.zeros((K, 2, 2))
    for m in range(2):
        Bm[:, m, 0] = g.coordV0T[:, 1, m] - g.coordV0T[:, 0, m]
        Bm[:, m, 1] = g.coordV0T[:, 2, m] - g.coordV0T[:, 0, m]
    g.B = Bm
    # markE0TE0T{nn, np} as sparse K x K (:175-185)
    markEint = E0E[:, 1] != -1
    mk = {}
    for nn in range(3):
        for np_ in range(3):
            i1 = (E0E[:, 0] == nn) & (E0E[:, 1] == np_) & markEint
            i2 = (E0E[:, 1] == nn) & (E0E[:, 0] == np_) & markEint
            rr = np.concatenate([T0E[i1, 0], T0E[i2, 1]])
            cc = np.concatenate([T0E[i1, 1], T0E[i2, 0]])
            mk[(nn, np_)] = sp.csr_matrix((np.ones(rr.size), (rr, cc)), shape=(K, K))
    g.markE0TE0T = mk
    if dense_patches:
        g.markV0TV0T = {(nn, np_): sp.csr_matrix((V0T[:, nn][:, None] == V0T[:, np_][None, :]).astype(float))
                        for nn in range(3) for np_ in range(3)}
    g.detJ0T = 2 * g.areaT
    return g


def mapRef2Phy(g, i, X1, X2):
    """g.mapRef2Phy(i, X1, X2): K x R physical coordinate i (0-based) of the
    reference points (X1, X2); generateGridData.m:195-196."""
    X1 = np.atleast_1d(np.asarray(X1, dtype=float)); X2 = np.atleast_1d(np.asarray(X2, dtype=float))
    B = g.B
    return (np.outer(B[:, i, 0], X1) + np.outer(B[:, i, 1], X2)
            + np.outer(g.coordV0T[:, 0, i], np.ones_like(X1)))


def markBoundaries(g, generateMarkE0TbdrN=None):
    """Boundary marks of advection/configureProblem.m:148-150 and
    advection/preprocessProblem.m:72-76; generateMarkE0TbdrN(g) gives the
    Neumann marks (the reference's default: none)."""
    g.markE0Tint = g.idE0T == 0
    g.markE0TbdrN = (np.zeros((g.numT, 3), dtype=bool) if generateMarkE0TbdrN is None
                     else np.asarray(generateMarkE0TbdrN(g), dtype=bool))
    g.markE0TbdrD = ~(g.markE0Tint | g.markE0TbdrN)
    bdrV = np.unique(g.V0E[g.E0T[g.markE0TbdrD], :].ravel())
    g.markV0TbdrD = np.isin(g.V0T, bdrV)
    return g


def computeDerivedGridData(g):
    """Vertex patches and boundary edge lengths;
    advection/computeDerivedGridData.m:66-81.

    markV0TT0V[n](k, k') = 1 iff vertex V0T(k, n) is a vertex of k'.  Built
    from the dense markV0TV0T when present, else from the vertex/element
    incidence (identical pattern)."""
    K = g.numT
    g.markV0TT0V = []
    if 'markV0TV0T' in g:
        for n in range(3):
            M = g.markV0TV0T[(n, 0)] + g.markV0TV0T[(n, 1)] + g.markV0TV0T[(n, 2)]
            M.data[:] = 1.0
            g.markV0TT0V.append(M.tocsr())
    else:
        inc = sp.csr_matrix((np.ones(3 * K), (g.V0T.ravel(), np.repeat(np.arange(K), 3))),
                            shape=(g.numV, K))
        inc.data[:] = 1.0
        for n in range(3):
            Sel = sp.csr_matrix((np.ones(K), (np.arange(K), g.V0T[:, n])), shape=(K, g.numV))
            M = (Sel @ inc).tocsr()
            M.data[:] = 1.0
            g.markV0TT0V.append(M)
    g.areaE0TbdrD = [g.areaE0T[:, n] * g.markE0TbdrD[:, n] for n in range(3)]
    g.areaE0TbdrN = g.markE0TbdrN * g.areaE0T
    g.areaE0TbdrNotN = [g.areaE0T[:, n] * ~g.markE0TbdrN[:, n] for n in range(3)]
    return g

#!/usr/bin/env python3
"""Write tests/golden/sliver_p4_exact.json: the dc/dt limiter's result for one
badly shaped element, evaluated in 50-digit arithmetic.

The mesh is the p = 4 random Delaunay mesh of
tests/test_gpu_parity.py::test_unstructured_mesh_steps (400 points, seed 11,
every boundary edge Dirichlet, 'rotating_humps', RK4, 'linear' limiter).
Element 319 has cond(B) = 105 and is limited with alpha = 0 in the first
stage.  The limiter of advection/solveSubStep.m:69-97 takes the stage's dc/dt
x (the oracle's, stored here) to
    t = M_T^-1 M_DT x,   l = (t_0, alpha t_1, alpha t_2, 0, ...),
    l += diag(M_T)^-1 M_T (t - l),   y = M^-1 M_DT^T l
(core/projectDataDisc2DataTaylor.m, core/applySlopeLimiterTaylorLinear.m,
the mass lumping of solveSubStep.m:75-78, core/projectDataTaylor2DataDisc.m),
with M, M_DT, M_T the element's DG, mixed and Taylor mass matrices.  Here
every step runs in mpmath at 50 digits: the element's vertices are the
mesh's doubles taken exactly, the integrals use the 28-point degree-11 rule
(festung_amd/csrc/gen_tri_rule11.py, solved at 60 digits; exact for these
degree-8 integrands), and the basis is the oracle's expressions evaluated on
mpmath numbers with 50-digit square roots.  alpha is the oracle's (0 here:
the element's centroid value is its patch's extremum, which no rounding
moves).  The fixture holds x, alpha, the exact y (rounded to double), the
oracle's double y and cond(M_T).

    python tests/golden/make_sliver_exact.py      (needs mpmath and scipy; a few minutes)
"""
import inspect
import json
import os
import sys

import mpmath as mp
import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)
sys.path.insert(0, os.path.join(ROOT, 'tests'))
sys.path.insert(0, os.path.join(ROOT, 'festung_amd', 'csrc'))

import gen_tri_rule11 as GR  # noqa: E402
from festung_amd.advection import testcases as TC  # noqa: E402
from oracle import advection as OA, basis as OB, grid as OG, limiter as OL  # noqa: E402
from test_gpu_parity import _delaunay  # noqa: E402

mp.mp.dps = 50
P, ELEM = 4, 319
EXPS = [(0, 0), (1, 0), (0, 1), (2, 0), (1, 1), (0, 2), (3, 0), (2, 1), (1, 2), (0, 3),
        (4, 0), (3, 1), (2, 2), (1, 3), (0, 4)]


def oracle_stage1():
    """The oracle's first-stage dc/dt, alpha and limited dc/dt on the mesh."""
    pts, tri = _delaunay(400, 7 + P)
    og = OG.generateGridData(pts, tri)
    idE = (og.E0E[:, 1] < 0).astype(np.int32)
    og.idE = idE
    og.idE0T = idE[og.E0T]
    tc = TC.rotating_humps()
    kw = dict(p=P, hmax=0.05, ordRK=4, isSlopeLim=True, numSteps=4, tEnd=4 * 0.1 * 0.02 / (2 * P + 1))
    cfg = OA.Struct(testcase='custom', **kw)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        cfg[k] = tc[k]
    cfg = OA.configureProblem(cfg)
    cfg.generateGridData = lambda h: og
    o = OA.initializeProblem(OA.preprocessProblem(cfg))
    o.t, o.omega = OA.rungeKuttaExplicit(o.ordRK, o.tau, 0.0)
    o.cDiscRK = [o.cDisc.reshape(-1).copy()]
    o = OA.preprocessSubStep(o, 1, 1)
    cdot = OA.computeDot(o, o.cDiscRK[0])
    K, N = o.K, o.N
    cT = OL.projectDataDisc2DataTaylor(cdot.reshape(K, N), o.globM, o.globMDiscTaylor)
    cTL, _ = OL.applySlopeLimiterTaylor(o.g, cT, o.g.markV0TbdrD, np.full((K, 3), np.nan), o.basesOnQuad,
                                        o.typeSlopeLim)
    k = ELEM
    alpha = cTL[k, 1] / cT[k, 1] if cT[k, 1] != 0 else cTL[k, 2] / cT[k, 2]
    y_or = OA.limitDot(o, cdot).reshape(K, N)[k]
    return pts, tri, cdot.reshape(K, N)[k], float(alpha), y_or, o.tau


def mp_basis():
    """The oracle's phi(i, X1, X2) expressions on mpmath numbers."""
    src = inspect.getsource(OB.phi)
    body = src.split('X1 = np.asarray(X1, dtype=float); X2 = np.asarray(X2, dtype=float)\n')[1]
    body = body.replace('np.ones_like(X1)', '1')
    ns = {'S%d' % n: mp.sqrt(n) for n in (2, 3, 5, 6, 10, 14, 30, 70)}
    exec('def mphi(i, X1, X2):\n' + body, ns)
    return ns['mphi']


def exact_limited(pts, tri, x, alpha):
    N = len(x)
    mphi = mp_basis()
    rule = GR.points(GR.solve())
    V = [[mp.mpf(float(c)) for c in pts[v]] for v in tri[ELEM]]
    B = [[V[1][0] - V[0][0], V[2][0] - V[0][0]], [V[1][1] - V[0][1], V[2][1] - V[0][1]]]
    det = B[0][0] * B[1][1] - B[0][1] * B[1][0]
    area = abs(det) / 2
    xc, yc = sum(v[0] for v in V) / 3, sum(v[1] for v in V) / 3
    dX = 2 / mp.sqrt(2 * area)
    Q = [(V[0][0] + B[0][0] * a + B[0][1] * b, V[0][1] + B[1][0] * a + B[1][1] * b, w * abs(det), a, b)
         for a, b, w in rule]
    mean = {e: sum(wq * (xq - xc) ** e[0] * (yq - yc) ** e[1] for xq, yq, wq, _, _ in Q) / area for e in EXPS}

    def taylor(i, xq, yq):  # core/phiTaylorPhy.m: centred, mean-free above degree 1, scaled
        a, b = EXPS[i]
        if i == 0:
            return mp.mpf(1)
        v = (xq - xc) ** a * (yq - yc) ** b
        if a + b >= 2:
            v -= mean[(a, b)]
        return v / (mp.factorial(a) * mp.factorial(b)) * dX ** (a + b)

    PHI = [[mphi(j + 1, aq, bq) for j in range(N)] for _, _, _, aq, bq in Q]
    TT = [[taylor(i, xq, yq) for i in range(N)] for xq, yq, _, _, _ in Q]
    MT, MTD, MD = mp.matrix(N, N), mp.matrix(N, N), mp.matrix(N, N)
    for q, (_, _, wq, _, _) in enumerate(Q):
        for i in range(N):
            for j in range(N):
                MT[i, j] += wq * TT[q][i] * TT[q][j]
                MTD[i, j] += wq * TT[q][i] * PHI[q][j]
                MD[i, j] += wq * PHI[q][i] * PHI[q][j]
    xd = mp.matrix([mp.mpf(float(v)) for v in x])
    t = mp.lu_solve(MT, MTD * xd)
    al = mp.mpf(alpha)
    lim = mp.matrix([t[0], al * t[1], al * t[2]] + [0] * (N - 3))
    corr = MT * (t - lim)
    for i in range(N):
        lim[i] += corr[i] / MT[i, i]
    y = mp.lu_solve(MD, MTD.T * lim)
    condMT = mp.norm(MT, 1) * mp.norm(mp.inverse(MT), 1)
    sv = np.linalg.svd(np.array([[float(B[i][j]) for j in range(2)] for i in range(2)]), compute_uv=False)
    return [float(v) for v in y], float(condMT), float(sv[0] / sv[1])


def main():
    pts, tri, x, alpha, y_or, tau = oracle_stage1()
    y, condMT, condB = exact_limited(pts, tri, x, alpha)
    rel = float(np.linalg.norm(np.asarray(y_or) - y) / np.linalg.norm(y))
    out = {'source': 'tests/golden/make_sliver_exact.py', 'p': P, 'element': ELEM, 'seed': 7 + P, 'points': 400,
           'tau': tau, 'condB': condB, 'condMT': condMT, 'alpha': alpha, 'dcdt': [float(v) for v in x],
           'limited_exact': y, 'limited_oracle': [float(v) for v in y_or], 'oracle_rel_error': rel}
    with open(os.path.join(HERE, 'sliver_p4_exact.json'), 'w') as f:
        json.dump(out, f, indent=1)
    print('element %d cond(B) %.1f cond(M_T) %.2e alpha %r oracle rel error %.3e' % (ELEM, condB, condMT, alpha, rel))


if __name__ == '__main__':
    main()

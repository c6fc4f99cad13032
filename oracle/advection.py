"""The advection problem module and its driver (oracle).  Restates
main.m:112-178, core/getStepLists.m, core/iterateSubSteps.m,
core/rungeKuttaExplicit.m and advection/{configureProblem, preprocessProblem,
computeDerivedGridData, initializeProblem, solveStep, preprocessSubStep,
solveSubStep, postprocessSubStep, postprocessStep, postprocessProblem}.m for
triangles (isQuadri = false).  Visualisation output is not restated.

Per stage the operators are re-assembled exactly as the reference does it
(advection/preprocessSubStep.m:69-98), then ``globM \\ (sysV - sysA c)``
(advection/solveSubStep.m:62-66), optional dc/dt and solution limiting, and
the SSP combination.  This is also the "reference-structured" CPU baseline.

Two build-defined extensions are marked as such: the 'custom' testcase
(closures supplied by the caller; the reference overwrites the closures of
its own testcases, advection/configureProblem.m:105-136) and ordRK = 4
(classic Runge-Kutta, absent from core/rungeKuttaExplicit.m): parity
unpinned for RK4, it is defined identically in the product.
"""
import time

import numpy as np
import scipy.sparse.linalg as spla

from . import assembly as A
from . import basis as Bm
from . import grid as G
from . import limiter as L


class Struct(dict):
    __getattr__ = dict.__getitem__

    def __setattr__(self, k, v):
        self[k] = v


def rungeKuttaExplicit(ord_, tau, t0):
    """Stage times and SSP weights omega; core/rungeKuttaExplicit.m:64-78."""
    if ord_ == 1:
        return np.array([t0]), np.array([0.0])
    if ord_ == 2:
        return t0 + np.array([0.0, 1.0]) * tau, np.array([0.0, 0.5])
    if ord_ == 3:
        return t0 + np.array([0.0, 1.0, 0.5]) * tau, np.array([0.0, 3 / 4, 1 / 3])
    if ord_ == 4:   # build-defined classic RK4 (not in the reference)
        return t0 + np.array([0.0, 0.5, 0.5, 1.0]) * tau, np.full(4, np.nan)
    raise ValueError('Order %d not implemented' % ord_)


# classic RK4 in accumulator form (build-defined): u_{s+1} = base + tau*a_s*cdot,
# acc <- (u0 if s == 0 else acc) + tau*b_s*cdot, base = u0 for s < 3, acc for s = 3
RK4_A = (0.5, 0.5, 1.0, 1 / 6)
RK4_B = (1 / 6, 1 / 3, 1 / 3, 0.0)


def configureProblem(pd):
    """advection/configureProblem.m:61-151 (triangles only)."""
    pd.setdefault('testcase', 'solid_body')
    pd.setdefault('hmax', 2 ** -6)
    pd.setdefault('isQuadri', False)
    assert not pd.isQuadri, 'the oracle restates the triangle path only'
    pd.setdefault('p', 0)
    pd.setdefault('qOrd', 2 * pd.p + 1)
    pd.setdefault('qOrdMax', pd.qOrd)
    pd.setdefault('ordRK', min(pd.p + 1, 3))
    pd.setdefault('numSteps', 100)
    pd.setdefault('tEnd', (pd.numSteps / (3142 * 1.5)) * 2 * np.pi)
    pd.setdefault('isSlopeLim', False)
    pd.setdefault('typeSlopeLim', 'linear')
    assert 0 <= pd.p <= 4
    assert 1 <= pd.ordRK <= 4
    assert pd.hmax > 0 and pd.numSteps > 0
    assert (not pd.isSlopeLim) or pd.p > 0
    tc = pd.testcase
    if tc == 'solid_body':
        pd.isAnalytical = False
        Gf = lambda x1, x2, a, b: (1 / 0.15) * np.sqrt((x1 - a) ** 2 + (x2 - b) ** 2)
        pd.c0Cont = lambda x1, x2: (((x1 - 0.5) ** 2 + (x2 - 0.75) ** 2 <= 0.0225)
                                    & ((x1 <= 0.475) | (x1 >= 0.525) | (x2 >= 0.85))).astype(float) \
            + (1 - Gf(x1, x2, 0.5, 0.25)) * ((x1 - 0.5) ** 2 + (x2 - 0.25) ** 2 <= 0.0225) \
            + 0.25 * (1 + np.cos(np.pi * Gf(x1, x2, 0.25, 0.5))) * ((x1 - 0.25) ** 2 + (x2 - 0.5) ** 2 <= 0.0225)
        pd.fCont = lambda t, x1, x2: np.zeros_like(x1)
        pd.u1Cont = lambda t, x1, x2: 0.5 - x2
        pd.u2Cont = lambda t, x1, x2: x1 - 0.5
        pd.cDCont = lambda t, x1, x2: np.zeros_like(x1)
        pd.gNCont = lambda t, x1, x2: np.zeros_like(x1)
    elif tc == 'transient':
        pd.isAnalytical = True
        pd.cCont = lambda t, x1, x2: np.cos(7 * x1) * np.cos(7 * x2) + np.exp(-t)
        pd.u1Cont = lambda t, x1, x2: np.exp(0.5 * (x1 + x2))
        pd.u2Cont = lambda t, x1, x2: np.exp(0.5 * (x1 - x2))
        pd.fCont = lambda t, x1, x2: (-np.exp(-t)
                                      - 7 * pd.u1Cont(t, x1, x2) * np.sin(7 * x1) * np.cos(7 * x2)
                                      - 7 * pd.u2Cont(t, x1, x2) * np.cos(7 * x1) * np.sin(7 * x2)
                                      + 0.5 * (pd.u1Cont(t, x1, x2) - pd.u2Cont(t, x1, x2)) * pd.cCont(t, x1, x2))
        pd.c0Cont = lambda x1, x2: pd.cCont(0, x1, x2)
        pd.cDCont = lambda t, x1, x2: pd.cCont(t, x1, x2)
        pd.gNCont = lambda t, x1, x2: np.zeros_like(x1)
    elif tc == 'custom':   # build-defined: caller supplies every closure
        for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
            assert k in pd, 'custom testcase needs %s' % k
        pd.setdefault('isAnalytical', 'cCont' in pd)
    else:
        raise ValueError('Invalid testcase "%s".' % tc)
    pd.generateGridData = G.domainSquare
    return pd


def preprocessProblem(pd):
    """advection/preprocessProblem.m:48-112 (triangle branch)."""
    pd.g = pd.generateGridData(pd.hmax)
    pd.K = pd.g.numT
    pd.N = (pd.p + 1) * (pd.p + 2) // 2
    pd.tau = pd.tEnd / pd.numSteps
    G.markBoundaries(pd.g, pd.get('generateMarkE0TbdrN'))
    G.computeDerivedGridData(pd.g)
    pd.basesOnQuad = Bm.computeBasesOnQuad(pd.N)
    if pd.isSlopeLim:
        L.computeTaylorBasesV0T(pd.g, pd.N, pd.basesOnQuad)
    pd.hatM = Bm.integrateRefElemPhiPhi(pd.N, pd.basesOnQuad)
    pd.hatG = Bm.integrateRefElemDphiPhiPhi(pd.N, pd.basesOnQuad)
    pd.hatRdiagOnQuad = Bm.integrateRefEdgePhiIntPhiIntPerQuad(pd.N, pd.basesOnQuad)
    pd.hatRoffdiagOnQuad = Bm.integrateRefEdgePhiIntPhiExtPerQuad(pd.N, pd.basesOnQuad)
    pd.globM = A.assembleMatElemPhiPhi(pd.g, pd.hatM)
    if pd.isSlopeLim:
        MT = L.assembleMatElemPhiTaylorPhiTaylor(pd.g, pd.N)
        pd.globMTaylor = MT
        pd.globMDiscTaylor = L.assembleMatElemPhiDiscPhiTaylor(pd.g, pd.N, pd.basesOnQuad)
        import scipy.sparse as sp
        pd.globMCorr = sp.diags(1.0 / MT.diagonal()) @ MT
    return pd


def initializeProblem(pd):
    """advection/initializeProblem.m:46-86 (no visualisation)."""
    pd.isFinished = False
    pd.cDisc = A.projectFuncCont2DataDisc(pd.g, pd.c0Cont, pd.qOrd, pd.hatM, pd.basesOnQuad)
    if pd.isSlopeLim:
        cDV0T = A.computeFuncContV0T(pd.g, lambda x1, x2: pd.cDCont(0, x1, x2))
        pd.cDisc, _ = L.applySlopeLimiterDisc(pd.g, pd.cDisc, pd.g.markV0TbdrD, cDV0T, pd.globM,
                                              pd.globMDiscTaylor, pd.basesOnQuad, pd.typeSlopeLim)
    pd.initialError = A.computeL2Error(pd.g, pd.cDisc, pd.c0Cont, 2 * pd.p, pd.basesOnQuad)
    return pd


def preprocessStep(pd, nStep):
    """advection/preprocessStep.m:58-60 (no-op)."""
    return pd


def solveStep(pd, nStep):
    """advection/solveStep.m:57-73."""
    pd.t, pd.omega = rungeKuttaExplicit(pd.ordRK, pd.tau, (nStep - 1) * pd.tau)
    pd.cDiscRK = [None] * (len(pd.t) + 1)
    pd.cDiscRK[0] = pd.cDisc.reshape(-1).copy()
    pd.isSubSteppingFinished = False
    return iterateSubSteps(pd, nStep)


def iterateSubSteps(pd, nStep):
    """core/iterateSubSteps.m:67-88 with the sub-step list of core/getStepLists.m:48."""
    nSubStep = 0
    while not pd.isSubSteppingFinished:
        nSubStep += 1
        pd = preprocessSubStep(pd, nStep, nSubStep)
        pd = solveSubStep(pd, nStep, nSubStep)
        pd = postprocessSubStep(pd, nStep, nSubStep)
    return pd


def preprocessSubStep(pd, nStep, nSubStep):
    """advection/preprocessSubStep.m:56-99 (triangle branch)."""
    K, N, p = pd.K, pd.N, pd.p
    ts = pd.t[nSubStep - 1]
    fDisc = A.projectFuncCont2DataDisc(pd.g, lambda x1, x2: pd.fCont(ts, x1, x2), 2 * p, pd.hatM, pd.basesOnQuad)
    u1Disc = A.projectFuncCont2DataDisc(pd.g, lambda x1, x2: pd.u1Cont(ts, x1, x2), 2 * p, pd.hatM, pd.basesOnQuad)
    u2Disc = A.projectFuncCont2DataDisc(pd.g, lambda x1, x2: pd.u2Cont(ts, x1, x2), 2 * p, pd.hatM, pd.basesOnQuad)
    vNormalOnQuadEdge = A.computeFuncContNuOnQuadEdge(pd.g, lambda x1, x2: pd.u1Cont(ts, x1, x2),
                                                      lambda x1, x2: pd.u2Cont(ts, x1, x2), 2 * p + 1)
    pd.u1Disc, pd.u2Disc, pd.fDisc, pd.vNormalOnQuadEdge = u1Disc, u2Disc, fDisc, vNormalOnQuadEdge
    pd.globG = A.assembleMatElemDphiPhiFuncDisc(pd.g, pd.hatG, [u1Disc, u2Disc])
    pd.globR = A.assembleMatEdgePhiPhiValUpwind(pd.g, ~pd.g.markE0TbdrN, pd.hatRdiagOnQuad,
                                                pd.hatRoffdiagOnQuad, vNormalOnQuadEdge)
    pd.globKD = A.assembleVecEdgePhiIntFuncContVal(pd.g, pd.g.markE0TbdrD, lambda x1, x2: pd.cDCont(ts, x1, x2),
                                                   vNormalOnQuadEdge, N, pd.basesOnQuad, pd.qOrd)
    gN = lambda x1, x2: (pd.gNCont(ts, x1, x2) <= 0) * pd.gNCont(ts, x1, x2)
    pd.globKN = A.assembleVecEdgePhiIntFuncCont(pd.g, pd.g.markE0TbdrN, gN, N, pd.basesOnQuad)
    pd.globL = pd.globM @ fDisc.reshape(-1)
    return pd


def limitDot(pd, cDiscDot):
    """dc/dt limiting with selective mass lumping; advection/solveSubStep.m:69-79."""
    K, N = pd.K, pd.N
    cT = L.projectDataDisc2DataTaylor(cDiscDot.reshape(K, N), pd.globM, pd.globMDiscTaylor)
    cTLim, _ = L.applySlopeLimiterTaylor(pd.g, cT, pd.g.markV0TbdrD, np.full((K, 3), np.nan),
                                         pd.basesOnQuad, pd.typeSlopeLim)
    cT = cTLim.reshape(-1) + pd.globMCorr @ (cT - cTLim).reshape(-1)
    return L.projectDataTaylor2DataDisc(cT.reshape(K, N), pd.globM, pd.globMDiscTaylor).reshape(-1)


def computeDot(pd, c):
    """cDiscDot = globM \\ (sysV - sysA c); advection/solveSubStep.m:62-66."""
    sysA = -pd.globG[0] - pd.globG[1] + pd.globR
    sysV = pd.globL - pd.globKD - pd.globKN
    return spla.spsolve(pd.globM.tocsc(), sysV - sysA @ c)


def solveSubStep(pd, nStep, nSubStep):
    """advection/solveSubStep.m:57-99 (triangle branch); ordRK = 4 is the
    build-defined accumulator form of classic RK4."""
    K, N = pd.K, pd.N
    s = nSubStep - 1
    cDiscDot = computeDot(pd, pd.cDiscRK[s])
    if pd.p > 1 and pd.isSlopeLim:
        cDiscDot = limitDot(pd, cDiscDot)
    if pd.ordRK == 4:
        base = pd.cDiscRK[0] if s < 3 else pd.rk4acc
        nxt = base + pd.tau * RK4_A[s] * cDiscDot
        if s < 3:
            pd.rk4acc = (pd.cDiscRK[0] if s == 0 else pd.rk4acc) + pd.tau * RK4_B[s] * cDiscDot
        pd.cDiscRK[s + 1] = nxt
    else:
        w = pd.omega[s]
        pd.cDiscRK[s + 1] = w * pd.cDiscRK[0] + (1 - w) * (pd.cDiscRK[s] + pd.tau * cDiscDot)
    if pd.isSlopeLim:
        tBC = pd.t[nSubStep] if nSubStep < len(pd.t) else pd.t[0] + pd.tau
        # kept for checks of the limiter's decisions (alpha of this application)
        pd.lastLimiterInput = (pd.cDiscRK[s + 1].reshape(K, N).copy(), tBC)
        cDV0T = A.computeFuncContV0T(pd.g, lambda x1, x2: pd.cDCont(tBC, x1, x2))
        lim, _ = L.applySlopeLimiterDisc(pd.g, pd.cDiscRK[s + 1].reshape(K, N), pd.g.markV0TbdrD, cDV0T,
                                         pd.globM, pd.globMDiscTaylor, pd.basesOnQuad, pd.typeSlopeLim)
        pd.cDiscRK[s + 1] = lim.reshape(-1)
    return pd


def postprocessSubStep(pd, nStep, nSubStep):
    """advection/postprocessSubStep.m:55-57."""
    pd.isSubSteppingFinished = nSubStep >= len(pd.omega)
    return pd


def postprocessStep(pd, nStep):
    """advection/postprocessStep.m:54-61."""
    pd.cDisc = pd.cDiscRK[-1].reshape(pd.K, pd.N)
    pd.isFinished = nStep >= pd.numSteps
    return pd


def postprocessProblem(pd):
    """advection/postprocessProblem.m:57-81 (triangle branch, no visualisation)."""
    if pd.isAnalytical:
        pd.error = A.computeL2Error(pd.g, pd.cDisc, lambda x1, x2: pd.cCont(pd.tEnd, x1, x2), pd.qOrd, pd.basesOnQuad)
    else:
        pd.error = A.computeL2Error(pd.g, pd.cDisc, pd.c0Cont, pd.qOrd, pd.basesOnQuad)
    return pd


def main(problemName='advection', pd=None, maxSteps=None, **kw):
    """main.m:112-178 for problemName 'advection'.  ``maxSteps`` stops the
    main loop early (used to time a bounded sample)."""
    assert problemName == 'advection'
    pd = Struct(pd or {})
    pd.update(kw)
    pd.problemName = problemName
    t0 = time.perf_counter()
    for f in (configureProblem, preprocessProblem, initializeProblem):
        pd = f(pd)
    pd.timePreprocess = time.perf_counter() - t0
    t0 = time.perf_counter()
    nStep = 0
    while not pd.isFinished:
        nStep += 1
        pd = preprocessStep(pd, nStep)
        pd = solveStep(pd, nStep)
        pd = postprocessStep(pd, nStep)
        if maxSteps is not None and nStep >= maxSteps:
            break
    pd.timeLoop = time.perf_counter() - t0
    pd.nStepsRun = nStep
    if pd.isFinished:
        pd = postprocessProblem(pd)
    return pd

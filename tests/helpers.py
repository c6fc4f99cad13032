"""Shared set-up for the parity tests: the same problem built once by the
oracle (numpy/scipy restatement of the reference) and once on the device."""
import numpy as np

from oracle import advection as OA
from oracle import assembly as OAs
from oracle import limiter as OL


def rel_l2(a, b):
    a = np.asarray(a); b = np.asarray(b)
    n = np.linalg.norm(b)
    return np.linalg.norm(a - b) / (n if n > 0 else 1.0)


def oracle_problem(tc, p, hmax, ordRK, lim, numSteps=10, tEnd=None, typeSlopeLim='linear'):
    """configure + preprocess + initialize in the oracle with the closures of
    the product testcase ``tc`` (a festung_amd.advection.testcases.Testcase)."""
    pd = OA.Struct(testcase='custom', p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=lim, numSteps=numSteps,
                   typeSlopeLim=typeSlopeLim)
    if tEnd is not None:
        pd.tEnd = tEnd
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        pd[k] = tc[k]
    if tc.get('cCont') is not None:
        pd.cCont = tc.cCont
        pd.isAnalytical = tc.isAnalytical
    pd = OA.configureProblem(pd)
    pd = OA.preprocessProblem(pd)
    pd = OA.initializeProblem(pd)
    return pd


def oracle_stage_operators(pd, ts):
    """The oracle's per-stage operators at time ts (preprocessSubStep)."""
    pd.t = np.array([ts])
    return OA.preprocessSubStep(pd, 1, 1)


def oracle_step(pd, c, nStep):
    """One oracle time step from element-major state c (K x N)."""
    pd.cDisc = c.copy()
    pd = OA.solveStep(pd, nStep)
    return pd.cDiscRK[-1].reshape(pd.K, pd.N)


def oracle_limit(pd, c, t):
    cDV0T = OAs.computeFuncContV0T(pd.g, lambda x1, x2: pd.cDCont(t, x1, x2))
    return OL.applySlopeLimiterDisc(pd.g, c, pd.g.markV0TbdrD, cDV0T, pd.globM, pd.globMDiscTaylor,
                                    pd.basesOnQuad, pd.typeSlopeLim)


def oracle_limit_disc_linear(pd, c, t):
    """core/applySlopeLimiterDiscLinear.m with the problem's Dirichlet vertex data."""
    cDV0T = OAs.computeFuncContV0T(pd.g, lambda x1, x2: pd.cDCont(t, x1, x2))
    return OL.applySlopeLimiterDiscLinear(pd.g, c, pd.g.markV0TbdrD, cDV0T)


def oracle_alpha(pd, c_pre, t):
    """alpha (per element) of core/computeVertexBasedLimiter.m for the
    solution limiter applied to c_pre at time t with the problem's Dirichlet
    vertex data: the decisions of core/applySlopeLimiterTaylorLinear.m."""
    tt = OL.projectDataDisc2DataTaylor(c_pre, pd.globM, pd.globMDiscTaylor)
    valV0T = OAs.computeFuncDiscAtPoints(tt[:, :3], pd.basesOnQuad['phiTaylorV0T'][:, :, :3])
    cD = OAs.computeFuncContV0T(pd.g, lambda x1, x2: pd.cDCont(t, x1, x2))
    a, _ = OL.computeVertexBasedLimiter(pd.g, tt[:, 0], valV0T, pd.g.markV0TbdrD, cD)
    return a

"""Step functions of the advection problem (triangles) over the device
``Problem``.  Each function keeps the name, signature and role of the
reference file it replaces; the numerics run on the GPU through the C-ABI,
the host only evaluates the user's closures at the points the library
returns (as the reference does inside projectFuncCont2DataDisc and
computeFuncContNuOnQuadEdge) and sequences the calls.

Visualisation (initializeProblem, outputStep, postprocessProblem) writes the
reference's VTK / Tecplot files when ``isVisSol`` is set; its default is
false here (the reference's is true), so that runs write no files unless
asked to.
"""
import os

import numpy as np

from .. import output as OUT
from ..core import Grid, Problem, quadRule1D, quadRule2D
from . import testcases as TC


class Struct(dict):
    """Attribute-access dict mirroring Matlab's struct ``problemData``."""
    __getattr__ = dict.__getitem__

    def __setattr__(self, k, v):
        self[k] = v


def rungeKuttaExplicit(ord_, tau, t0):
    """Stage times and SSP weights; core/rungeKuttaExplicit.m:64-78.
    ord 4 is classic Runge-Kutta (an extension; omega is not used for it)."""
    if ord_ == 1:
        return np.array([t0]), np.array([0.0])
    if ord_ == 2:
        return t0 + np.array([0.0, 1.0]) * tau, np.array([0.0, 0.5])
    if ord_ == 3:
        return t0 + np.array([0.0, 1.0, 0.5]) * tau, np.array([0.0, 3 / 4, 1 / 3])
    if ord_ == 4:
        return t0 + np.array([0.0, 0.5, 0.5, 1.0]) * tau, np.full(4, np.nan)
    raise ValueError('Order %d not implemented' % ord_)


def gammaMap(n, S):
    """core/gammaMap.m:50-57 (n 1-based)."""
    S = np.asarray(S, dtype=float).ravel()
    if n == 1:
        return 1 - S, S
    if n == 2:
        return np.zeros_like(S), 1 - S
    return S, np.zeros_like(S)


def configureProblem(pd):
    """advection/configureProblem.m:61-151 (triangles).  The reference's
    testcases overwrite the closures; 'custom' keeps caller-supplied ones."""
    pd.setdefault('testcase', 'solid_body')
    pd.setdefault('hmax', 2 ** -6)
    pd.setdefault('isQuadri', False)
    if pd.isQuadri:
        raise NotImplementedError('quadrilateral meshes are not part of this implementation')
    pd.setdefault('p', 0)
    pd.setdefault('qOrd', 2 * pd.p + 1)
    pd.setdefault('qOrdMax', pd.qOrd)
    pd.setdefault('ordRK', min(pd.p + 1, 3))
    pd.setdefault('numSteps', 100)
    pd.setdefault('tEnd', (pd.numSteps / (3142 * 1.5)) * 2 * np.pi)
    pd.setdefault('isSlopeLim', False)
    pd.setdefault('typeSlopeLim', 'linear')
    pd.setdefault('useGraph', True)
    # visualisation, advection/configureProblem.m:91-96 (isVisSol defaults to false here)
    pd.setdefault('isVisGrid', False)
    pd.setdefault('isVisSol', False)
    pd.setdefault('outputFrequency', 100)
    pd.setdefault('outputBasename', os.path.join('output', 'advection_' + pd.get('typeSlopeLim', 'linear')))
    pd.setdefault('outputTypes', ['vtk'])
    assert 0 <= pd.p <= 4, 'Polynomial order must be zero to four.'
    assert 1 <= pd.ordRK <= 4, 'Order of Runge Kutta must be one to four.'
    assert pd.hmax > 0, 'Maximum edge length must be positive.'
    assert pd.numSteps > 0, 'Number of time steps must be positive.'
    assert (not pd.isSlopeLim) or pd.p > 0, 'Slope limiting only available for p > 0.'
    if pd.isSlopeLim and pd.typeSlopeLim not in ('linear', 'hierarch_vert', 'strict'):
        raise ValueError('Unknown limiter type')
    if pd.testcase == 'custom':
        for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
            assert k in pd, 'custom testcase needs %s' % k
        pd.setdefault('isAnalytical', 'cCont' in pd)
        for k in ('fZero', 'cDZero', 'isStationary'):
            pd.setdefault(k, False)
    elif pd.testcase in TC.TESTCASES:
        pd.update(TC.TESTCASES[pd.testcase]())
    else:
        raise ValueError('Invalid testcase "%s".' % pd.testcase)
    pd.setdefault('isVelocityStationary', pd.isStationary)
    pd.generateGridData = Grid.domainSquare
    # Neumann edges, advection/configureProblem.m:149 (the reference's
    # default: none); a callable g -> K x 3 bool marks when given
    pd.setdefault('generateMarkE0TbdrN', None)
    return pd


def preprocessProblem(pd):
    """advection/preprocessProblem.m:48-112: grid, sizes, device problem
    (reference blocks and mass matrices are built inside the library)."""
    pd.g = pd.generateGridData(pd.hmax)
    if pd.get('generateMarkE0TbdrN') is not None:
        # advection/preprocessProblem.m:72-76: markE0TbdrD = ~(int | N), markV0TbdrD
        pd.g.set_boundary(pd.g.idE, np.asarray(pd.generateMarkE0TbdrN(pd.g), dtype=bool))
    pd.K = pd.g.numT
    pd.N = (pd.p + 1) * (pd.p + 2) // 2
    pd.tau = pd.tEnd / pd.numSteps
    pd.problem = Problem(pd.g, pd.p, pd.ordRK, limiter=pd.isSlopeLim, graph=pd.useGraph,
                         limiter_type=pd.typeSlopeLim if pd.isSlopeLim else 'linear')
    prob = pd.problem
    # physical evaluation points (the closures are evaluated on the host)
    Q1, Q2, _ = quadRule2D(prob.ordE)
    pd.XE = pd.g.mapRef2Phy(Q1, Q2)
    Q1, Q2, _ = quadRule2D(prob.ordF)
    pd.XF = pd.g.mapRef2Phy(Q1, Q2)
    Q, _ = quadRule1D(2 * pd.p + 1)
    X1 = np.empty((pd.K, 3, len(Q)), order='F'); X2 = np.empty_like(X1)
    for n in range(3):
        A1, A2 = gammaMap(n + 1, Q)
        X1[:, n, :], X2[:, n, :] = pd.g.mapRef2Phy(A1, A2)
    pd.XEdge = (X1, X2)
    pd.dirichletVertices = np.unique(pd.g.V0T[pd.g.markV0TbdrD])
    # K_N is formed only when the grid has Neumann edges (markE0TbdrN, set
    # through Grid.set_boundary); it is identically zero otherwise
    pd.hasNeumann = bool(np.any(pd.g.markE0TbdrN))
    pd.velocitySet = False
    pd.deviceEval = _device_points(pd) if pd.get('deviceEval', True) else None
    if pd.isVisGrid:
        # advection/preprocessProblem.m:54-59 (an image next to the solution files)
        OUT.visualizeGrid(pd.g, pd.outputBasename + '_grid.png')
    return pd


def _device_points(pd):
    """Evaluation points on the GPU (torch, [R][K] layouts of the C-ABI) when
    the closures that are evaluated every stage accept torch tensors; then
    time-dependent data are computed and handed over on the device.  None
    when torch has no GPU or a closure needs numpy arrays (the host evaluates
    them then, as before)."""
    try:
        import torch
        if not torch.cuda.is_available():
            return None
    except ImportError:
        return None
    dev = torch.device('cuda')
    t = lambda a: torch.from_numpy(np.ascontiguousarray(np.asarray(a).T)).to(dev)
    X1, X2 = pd.XEdge
    xy = pd.g.coordV[pd.dirichletVertices]
    D = dict(XE=(t(pd.XE[0]), t(pd.XE[1])), XEdge=(t(X1), t(X2)),
             dirIdx=torch.from_numpy(pd.dirichletVertices.astype(np.int64)).to(dev),
             dirXY=(torch.from_numpy(xy[:, 0].copy()).to(dev), torch.from_numpy(xy[:, 1].copy()).to(dev)))
    probe = (torch.zeros(2, dtype=torch.float64, device=dev), torch.zeros(2, dtype=torch.float64, device=dev))
    need = []
    if not pd.isVelocityStationary:
        need += [pd.u1Cont, pd.u2Cont]
    if not pd.fZero:
        need.append(pd.fCont)
    if not pd.cDZero:
        need.append(pd.cDCont)
    if pd.hasNeumann:
        need.append(pd.gNCont)
    try:
        for f in need:
            _as_device(f(0.0, *probe), probe[0])
    except Exception:
        return None
    return D


def _as_device(r, like):
    """A closure's result as a contiguous float64 tensor shaped like `like`."""
    import torch
    r = torch.as_tensor(r, dtype=torch.float64, device=like.device)
    return torch.broadcast_to(r, like.shape).contiguous()


def vertexDirichlet(pd, t):
    """c_D at the Dirichlet vertices (NaN elsewhere); the per-vertex form of
    computeFuncContV0T restricted by markV0TbdrD."""
    v = np.full(pd.g.numV, np.nan)
    idx = pd.dirichletVertices
    xy = pd.g.coordV[idx]
    v[idx] = pd.cDCont(t, xy[:, 0], xy[:, 1])
    return v


def initializeProblem(pd):
    """advection/initializeProblem.m:46-86: project c0 with qOrd = 2p+1,
    optional initial limiting, L2 error of the projection (order 2p)."""
    pd.isFinished = False
    prob = pd.problem
    prob.project(prob.ordF, pd.c0Cont(*pd.XF), set_state=True)
    if pd.isSlopeLim:
        prob.set_vertex_dirichlet(vertexDirichlet(pd, 0.0))
        pd.vertexDTime = 0.0
        prob.limit()
    pd.initialError = prob.l2_error(prob.ordE, pd.c0Cont(*pd.XE))
    if pd.isVisSol:
        _visualize(pd, prob.get_state(), 0)
    return pd


def _visualize(pd, cDisc, tLvl):
    """advection/initializeProblem.m:76-83, outputStep.m:55-64, postprocessProblem.m:45-54."""
    OUT.visualizeDataLagr(pd.g, OUT.projectDataDisc2DataLagr(cDisc), 'u_h', pd.outputBasename, tLvl,
                          pd.outputTypes)


def preprocessStep(pd, nStep):
    """advection/preprocessStep.m:58-60 (no-op)."""
    return pd


def solveStep(pd, nStep):
    """advection/solveStep.m:57-73: RK tables, then the sub-steps.  When all
    data are time-independent and already on the device the whole step runs
    as one device call (a captured hipGraph)."""
    pd.t, pd.omega = rungeKuttaExplicit(pd.ordRK, pd.tau, (nStep - 1) * pd.tau)
    pd.isSubSteppingFinished = False
    if pd.isStationary and pd.velocitySet and (not pd.isSlopeLim or pd.get('vertexDTime') is not None):
        pd.problem.step(1, pd.tau)
        pd.isSubSteppingFinished = True
        return pd
    pd.problem.begin_step()
    return iterateSubSteps(pd, nStep)


def iterateSubSteps(pd, nStep):
    """core/iterateSubSteps.m:67-88 over core/getStepLists.m:48."""
    nSubStep = 0
    while not pd.isSubSteppingFinished:
        nSubStep += 1
        for f in (preprocessSubStep, solveSubStep, postprocessSubStep):
            pd = f(pd, nStep, nSubStep)
    return pd


def preprocessSubStep(pd, nStep, nSubStep):
    """advection/preprocessSubStep.m:56-99: velocity (projected with order 2p
    and as u.nu on the edge points), globL - globKD - globKN.
    Time-independent or identically-zero data are uploaded once / skipped."""
    prob = pd.problem
    ts = pd.t[nSubStep - 1]
    D = pd.get('deviceEval')
    if not (pd.velocitySet and pd.isVelocityStationary):
        if D is not None and not pd.isVelocityStationary:   # evaluated on the device
            E, Q = D['XEdge'], D['XE']
            prob.set_velocity_points_dev(_as_device(pd.u1Cont(ts, *Q), Q[0]), _as_device(pd.u2Cont(ts, *Q), Q[0]),
                                         _as_device(pd.u1Cont(ts, *E), E[0]), _as_device(pd.u2Cont(ts, *E), E[0]))
        else:
            X1, X2 = pd.XEdge
            prob.set_velocity_points(pd.u1Cont(ts, *pd.XE), pd.u2Cont(ts, *pd.XE),
                                     pd.u1Cont(ts, X1, X2), pd.u2Cont(ts, X1, X2))
        pd.velocitySet = True
    if not (pd.fZero and pd.cDZero and not pd.hasNeumann):
        if D is not None:
            E, Q = D['XEdge'], D['XE']
            fq = None if pd.fZero else _as_device(pd.fCont(ts, *Q), Q[0])
            cDe = None if pd.cDZero else _as_device(pd.cDCont(ts, *E), E[0])
            gNe = _as_device(pd.gNCont(ts, *E), E[0]) if pd.hasNeumann else None
            prob.set_source_points_dev(fq, cDe, gNe)
        else:
            fq = None if pd.fZero else pd.fCont(ts, *pd.XE)
            cDe = None if pd.cDZero else pd.cDCont(ts, *pd.XEdge)
            gNe = pd.gNCont(ts, *pd.XEdge) if pd.hasNeumann else None
            prob.set_source_points(fq, cDe, gNe)
    return pd


def solveSubStep(pd, nStep, nSubStep):
    """advection/solveSubStep.m:57-99 as one device stage (rhs, dc/dt
    limiting, RK combination, solution limiting with c_D at t_{s+1})."""
    if pd.isSlopeLim:
        tBC = pd.t[nSubStep] if nSubStep < len(pd.t) else pd.t[0] + pd.tau
        if not (pd.cDZero and pd.get('vertexDTime') is not None):
            D = pd.get('deviceEval')
            if D is not None and not pd.cDZero:
                import torch
                v = torch.full((pd.g.numV,), float('nan'), dtype=torch.float64, device=D['dirIdx'].device)
                v[D['dirIdx']] = _as_device(pd.cDCont(tBC, *D['dirXY']), D['dirXY'][0])
                pd.problem.set_vertex_dirichlet_dev(v)
            else:
                pd.problem.set_vertex_dirichlet(vertexDirichlet(pd, tBC))
            pd.vertexDTime = tBC
    pd.problem.stage(nSubStep, pd.tau)
    return pd


def postprocessSubStep(pd, nStep, nSubStep):
    """advection/postprocessSubStep.m:55-57."""
    pd.isSubSteppingFinished = nSubStep >= len(pd.omega)
    return pd


def postprocessStep(pd, nStep):
    """advection/postprocessStep.m:54-61 (the state stays on the device;
    ``cDisc`` is fetched by postprocessProblem)."""
    pd.isFinished = nStep >= pd.numSteps
    return pd


def outputStep(pd, nStep):
    """advection/outputStep.m:53-66: Lagrange-point output every
    outputFrequency steps."""
    if pd.isVisSol and nStep % pd.outputFrequency == 0:
        _visualize(pd, pd.problem.get_state(), nStep)
    return pd


def postprocessProblem(pd):
    """advection/postprocessProblem.m:43-82: final state and L2 error with
    qOrd = 2p+1 against c(tEnd) (analytical) or c0."""
    prob = pd.problem
    pd.cDisc = prob.get_state()
    if pd.isVisSol:
        _visualize(pd, pd.cDisc, pd.numSteps)
    if pd.isAnalytical:
        cq = pd.cCont(pd.tEnd, *pd.XF)
    else:
        cq = pd.c0Cont(*pd.XF)
    pd.error = prob.l2_error(prob.ordF, cq)
    return pd

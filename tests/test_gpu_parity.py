"""Parity of the device path (through the C-ABI) with the oracle on the same
seeded inputs: projection, the DG right-hand side, single RK steps with and
without the slope limiter, multi-step runs, the analytical 'transient' case
with time-dependent source and Dirichlet data, limiter outputs, and
run-to-run determinism.  Tolerance: 1e-12 relative L2 per step (FP64)."""
import os

import numpy as np
import pytest

from festung_amd.advection import steps as S
from festung_amd.advection import testcases as TC
from helpers import (oracle_problem, oracle_stage_operators, oracle_step, oracle_limit,
                     oracle_limit_disc_linear, rel_l2)
from oracle import advection as OA

pytestmark = pytest.mark.gpu

TOL = 1e-12


def product_problem(tc, p, hmax, ordRK, lim, numSteps=10, tEnd=None, graph=False, typeSlopeLim='linear'):
    pd = S.Struct(testcase='custom', p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=lim, numSteps=numSteps,
                  useGraph=graph, typeSlopeLim=typeSlopeLim)
    if tEnd is not None:
        pd.tEnd = tEnd
    pd.update(tc)
    pd = S.configureProblem(pd)
    pd = S.preprocessProblem(pd)
    pd = S.initializeProblem(pd)
    return pd


def product_step(pd, nStep):
    pd.t, pd.omega = S.rungeKuttaExplicit(pd.ordRK, pd.tau, (nStep - 1) * pd.tau)
    pd.isSubSteppingFinished = False
    pd.problem.begin_step()
    pd = S.iterateSubSteps(pd, nStep)
    return pd.problem.get_state()


CASES = [(1, 2, False), (2, 3, False), (1, 2, True), (2, 3, True), (3, 3, False), (3, 3, True),
         (0, 1, False), (4, 3, False), (4, 3, True), (2, 4, False), (2, 4, True)]


@pytest.mark.parametrize('p,ordRK,lim', CASES)
def test_initial_projection(gpu, p, ordRK, lim):
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, ordRK, lim)
    d = product_problem(tc, p, 1 / 8, ordRK, lim)
    c = d.problem.get_state()
    assert rel_l2(c, o.cDisc) < TOL
    assert abs(d.initialError - o.initialError) <= 1e-12 * max(o.initialError, 1e-300)


@pytest.mark.parametrize('p', [0, 1, 2, 3, 4])
def test_rhs(gpu, p):
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, min(p + 1, 3), False)
    d = product_problem(tc, p, 1 / 8, min(p + 1, 3), False)
    rng = np.random.default_rng(7 + p)
    c = rng.standard_normal((o.K, o.N))
    o = oracle_stage_operators(o, 0.0)
    ref = OA.computeDot(o, c.reshape(-1)).reshape(o.K, o.N)
    d.t = np.array([0.0])
    S.preprocessSubStep(d, 1, 1)
    d.problem.set_state(c)
    got = d.problem.rhs()
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize('p,ordRK,lim', CASES)
def test_one_step(gpu, p, ordRK, lim):
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, ordRK, lim, numSteps=20, tEnd=0.5)
    d = product_problem(tc, p, 1 / 8, ordRK, lim, numSteps=20, tEnd=0.5)
    c = o.cDisc.copy()
    d.problem.set_state(c)
    got = product_step(d, 1)
    ref = oracle_step(o, c, 1)
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize('p,ordRK,lim,graph', [(1, 2, False, True), (2, 3, True, True), (2, 3, True, False),
                                               (1, 2, True, True), (3, 3, True, True)])
def test_main_loop(gpu, p, ordRK, lim, graph):
    """20 steps through main('advection', ...) against the oracle's main, at a
    stable time step (CFL 0.4 / (2p+1) with max|u| <= 0.8)."""
    import festung_amd
    tc = TC.rotating_humps()
    h = 1 / 12
    kw = dict(p=p, hmax=h, ordRK=ordRK, isSlopeLim=lim, numSteps=20, tEnd=20 * 0.4 * h / (0.8 * (2 * p + 1)))
    d = festung_amd.main('advection', dict(testcase='custom', useGraph=graph, **tc, **kw))
    ocfg = OA.Struct(testcase='custom', **kw)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        ocfg[k] = tc[k]
    o = OA.main('advection', ocfg)
    assert rel_l2(d.cDisc, o.cDisc) < 20 * TOL
    assert abs(d.error - o.error) <= 1e-10 * o.error


@pytest.mark.parametrize('p,lim', [(1, False), (2, False), (2, True)])
def test_transient_source(gpu, p, lim):
    """Time-dependent f and c_D ('transient', advection/configureProblem.m:119-132)."""
    import festung_amd
    kw = dict(p=p, hmax=1 / 8, isSlopeLim=lim, numSteps=8, tEnd=0.08)
    d = festung_amd.main('advection', dict(testcase='transient', **kw))
    o = OA.main('advection', OA.Struct(testcase='transient', **kw))
    assert rel_l2(d.cDisc, o.cDisc) < 10 * TOL
    assert abs(d.error - o.error) <= 1e-10 * o.error


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_limiter_alone(gpu, p):
    """applySlopeLimiterDisc on a rough state: result, alpha and patch bounds."""
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, min(p + 1, 3), True)
    d = product_problem(tc, p, 1 / 8, min(p + 1, 3), True)
    rng = np.random.default_rng(11 + p)
    c = o.cDisc + 0.05 * rng.standard_normal(o.cDisc.shape)
    ref, (mn, mx) = oracle_limit(o, c, 0.0)
    d.problem.set_state(c)
    d.problem.limit()
    got = d.problem.get_state()
    alpha, gmn, gmx = d.problem.limiter_info()
    assert rel_l2(got, ref) < TOL
    assert np.max(np.abs(gmn - mn)) < 1e-14 and np.max(np.abs(gmx - mx)) < 1e-14
    assert np.all((alpha >= 0) & (alpha <= 1))
    # the limiter's decisions: alpha of every element as the oracle forms it
    # (core/computeVertexBasedLimiter.m:78-106 on the Taylor data); a flip at a
    # tol threshold would show here as a difference of order one
    from oracle import limiter as OL
    from oracle import assembly as OAs
    t = OL.projectDataDisc2DataTaylor(c, o.globM, o.globMDiscTaylor)
    valV0T = OAs.computeFuncDiscAtPoints(t[:, :3], o.basesOnQuad['phiTaylorV0T'][:, :, :3])
    cD = OAs.computeFuncContV0T(o.g, lambda x1, x2: o.cDCont(0.0, x1, x2))
    aref, _ = OL.computeVertexBasedLimiter(o.g, t[:, 0], valV0T, o.g.markV0TbdrD, cD)
    assert 0 < np.sum(aref < 1) < o.K
    assert np.max(np.abs(alpha - aref)) < 1e-12


@pytest.mark.parametrize('type_', ['hierarch_vert', 'strict'])
@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_limiter_types_alone(gpu, type_, p):
    """applySlopeLimiterDisc with the 'hierarch_vert' and 'strict' types
    (vertex-patch bounds of every derivative coefficient below degree p)."""
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, min(p + 1, 3), True, typeSlopeLim=type_)
    d = product_problem(tc, p, 1 / 8, min(p + 1, 3), True, typeSlopeLim=type_)
    rng = np.random.default_rng(41 + p)
    c = o.cDisc + 0.05 * rng.standard_normal(o.cDisc.shape)
    ref, (mn, mx) = oracle_limit(o, c, 0.0)
    d.problem.set_state(c)
    d.problem.limit()
    got = d.problem.get_state()
    alpha, gmn, gmx = d.problem.limiter_info()
    assert rel_l2(got, ref) < TOL
    assert np.max(np.abs(gmn - mn)) < 1e-14 and np.max(np.abs(gmx - mx)) < 1e-14
    assert np.all((alpha >= 0) & (alpha <= 1))


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_limiter_disc_linear_alone(gpu, p):
    """core/applySlopeLimiterDiscLinear.m on a rough state with the
    problem's Dirichlet vertex data: bit-identical result, alpha and bounds
    (the same IEEE operations in the same order as the oracle)."""
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, min(p + 1, 3), True)
    d = product_problem(tc, p, 1 / 8, min(p + 1, 3), True)
    rng = np.random.default_rng(61 + p)
    c = o.cDisc + 0.05 * rng.standard_normal(o.cDisc.shape)
    ref, (mn, mx), aref = oracle_limit_disc_linear(o, c, 0.0)
    assert 0 < np.sum(aref < 1) < o.K
    d.problem.set_state(c)
    d.problem.limit_disc_linear()
    got = d.problem.get_state()
    alpha, gmn, gmx = d.problem.limiter_info()
    assert np.array_equal(alpha, aref)
    assert np.array_equal(got, ref)
    assert np.array_equal(gmn, mn) and np.array_equal(gmx, mx)


def test_limiter_disc_linear_needs_limiter(gpu):
    from festung_amd._lib import FestungError
    d = product_problem(TC.rotating_humps(), 2, 1 / 4, 3, False)
    with pytest.raises(FestungError, match='without the limiter'):
        d.problem.limit_disc_linear()


@pytest.mark.parametrize('type_', ['hierarch_vert', 'strict'])
@pytest.mark.parametrize('p,ordRK', [(2, 3), (3, 3), (4, 4), (2, 2)])
def test_limiter_types_step(gpu, type_, p, ordRK):
    """One full step with dc/dt and solution limiting of the given type."""
    tc = TC.rotating_humps()
    o = oracle_problem(tc, p, 1 / 8, ordRK, True, numSteps=20, tEnd=0.5, typeSlopeLim=type_)
    d = product_problem(tc, p, 1 / 8, ordRK, True, numSteps=20, tEnd=0.5, typeSlopeLim=type_)
    c = o.cDisc.copy()
    d.problem.set_state(c)
    got = product_step(d, 1)
    ref = oracle_step(o, c, 1)
    assert rel_l2(got, ref) < TOL


def test_determinism(gpu):
    tc = TC.rotating_humps()
    outs = []
    for _ in range(2):
        d = product_problem(tc, 2, 1 / 16, 3, True, graph=True)
        for n in range(1, 4):
            product_step(d, n)
        d.problem.step(3, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_graph_matches_stagewise(gpu):
    tc = TC.rotating_humps()
    a = product_problem(tc, 2, 1 / 16, 3, True, graph=True)
    b = product_problem(tc, 2, 1 / 16, 3, True, graph=False)
    product_step(a, 1)
    product_step(b, 1)
    a.problem.step(4, a.tau)
    for n in range(2, 6):
        product_step(b, n)
    assert np.array_equal(a.problem.get_state(), b.problem.get_state())


def test_device_pointer_state(gpu):
    """State transfer through device pointers ([N][K] in the reference's
    element order) round-trips exactly through the device's element order."""
    import torch
    tc = TC.rotating_humps()
    d = product_problem(tc, 2, 1 / 8, 3, False)
    rng = np.random.default_rng(3)
    c = rng.standard_normal((d.K, d.N))
    t = torch.tensor(np.ascontiguousarray(c.T), device='cuda')
    d.problem.set_state_dev(t)
    assert np.array_equal(d.problem.get_state(), c)
    out = torch.empty_like(t)
    d.problem.get_state_dev(out)
    d.problem.sync()
    assert np.array_equal(out.cpu().numpy().T, c)


@pytest.mark.parametrize('p,ordRK,lim', [(1, 2, True), (2, 3, True), (3, 3, True), (4, 3, True), (2, 4, True),
                                         (0, 1, False)])
def test_elem_block_invariance(gpu, p, ordRK, lim):
    """Every element's arithmetic, and the limiter's global min / max, are
    independent of the number of elements per workgroup: the states after a
    few steps agree bit for bit (K = 1152 leaves a partial workgroup)."""
    tc = TC.rotating_humps()
    outs = []
    for bs in (64, 128, 256):
        d = product_problem(tc, p, 1 / 24, ordRK, lim, graph=True)
        d.problem.elem_block = bs
        assert d.problem.elem_block == bs
        product_step(d, 1)
        d.problem.step(3, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1]) and np.array_equal(outs[0], outs[2])


def test_vertex_dirichlet_change_recaptures_graph(gpu):
    """Dropping the Dirichlet vertex data after whole steps were captured in a
    graph: the next steps use the new limiter input (bit for bit as eager
    steps do)."""
    tc = TC.rotating_humps()
    outs = []
    for graph in (True, False):
        d = product_problem(tc, 2, 1 / 16, 3, True, graph=graph)
        product_step(d, 1)
        d.problem.step(2, d.tau)
        d.problem.set_vertex_dirichlet(None)
        d.problem.step(2, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_elem_block_rejects_invalid(gpu):
    from festung_amd._lib import FestungError
    d = product_problem(TC.rotating_humps(), 1, 1 / 8, 2, False)
    for bs in (0, 32, 96, 512):
        with pytest.raises(FestungError, match='elements per workgroup'):
            d.problem.elem_block = bs


def test_autotune_keeps_state(gpu):
    tc = TC.rotating_humps()
    d = product_problem(tc, 2, 1 / 32, 3, True, graph=True)
    product_step(d, 1)
    c0 = d.problem.get_state()
    best, ms = d.problem.autotune(d.tau, 3)
    assert best in (64, 128, 256) and d.problem.elem_block == best
    assert set(ms) == {64, 128, 256} and all(v > 0 for v in ms.values())
    assert np.array_equal(d.problem.get_state(), c0)
    d.problem.step(2, d.tau)
    e = product_problem(tc, 2, 1 / 32, 3, True, graph=True)
    product_step(e, 1)
    e.problem.step(2, e.tau)
    assert np.array_equal(d.problem.get_state(), e.problem.get_state())


def test_stagewise_then_steps(gpu):
    """Stage-by-stage steps, a state reset and then captured whole steps give
    the result of the same number of steps on a fresh problem."""
    tc = TC.rotating_humps()
    a = product_problem(tc, 2, 1 / 16, 3, True, graph=True)
    product_step(a, 1)
    c1 = a.problem.get_state()
    for _ in range(2):
        a.problem.set_state(c1)
        a.problem.begin_step()
        for n in range(1, 4):
            a.problem.stage(n, a.tau)
    a.problem.step(2, a.tau)
    b = product_problem(tc, 2, 1 / 16, 3, True, graph=True)
    product_step(b, 1)
    b.problem.step(3, b.tau)
    assert np.array_equal(a.problem.get_state(), b.problem.get_state())


def test_set_state_waits_for_running_steps(gpu):
    """A state upload right after asynchronously enqueued steps lands after
    them (the problem's stream is drained first), and so does a download."""
    tc = TC.rotating_humps()
    d = product_problem(tc, 2, 1 / 32, 3, True, graph=True)
    product_step(d, 1)
    c1 = d.problem.get_state()
    d.problem.step(20, d.tau)       # returns before the steps have run
    d.problem.set_state(c1)
    assert np.array_equal(d.problem.get_state(), c1)


def _timedep_velocity_case():
    import math
    from festung_amd.advection.testcases import Testcase, _xp, gaussian_humps
    s = lambda t: 1.0 + 0.5 * math.sin(4.0 * t)
    return Testcase(isAnalytical=False, c0Cont=gaussian_humps(61),
                    u1Cont=lambda t, x1, x2: s(t) * (0.5 - x2) + 0.1 * _xp(x1).sin(3.0 * x1) * (1 - 0 * x2),
                    u2Cont=lambda t, x1, x2: s(t) * (x1 - 0.5) + 0 * x1,
                    fCont=lambda t, x1, x2: 0.2 * math.cos(t) * _xp(x1).exp(-x1 * x2),
                    cDCont=lambda t, x1, x2: 0.1 * (1 + x1 - x2) * math.exp(-t),
                    gNCont=lambda t, x1, x2: 0 * x1,
                    fZero=False, cDZero=False, isStationary=False, isVelocityStationary=False)


@pytest.mark.parametrize('lim', [False, True])
def test_time_dependent_data_on_device(gpu, lim):
    """Velocity, source and Dirichlet data that change every stage, evaluated
    by torch on the device and handed over by device pointer: against the
    oracle, and against the same run with host evaluation."""
    import festung_amd
    tc = _timedep_velocity_case()
    kw = dict(p=2, hmax=1 / 8, isSlopeLim=lim, numSteps=6, tEnd=0.06, ordRK=3)
    d = festung_amd.main('advection', dict(testcase='custom', **tc, **kw))
    assert d.deviceEval is not None
    h = festung_amd.main('advection', dict(testcase='custom', deviceEval=False, **tc, **kw))
    assert h.deviceEval is None
    ocfg = OA.Struct(testcase='custom', **kw)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        ocfg[k] = tc[k]
    o = OA.main('advection', ocfg)
    assert rel_l2(d.cDisc, o.cDisc) < 10 * TOL
    assert rel_l2(d.cDisc, h.cDisc) < 1e-13


def test_main_writes_visualisation(gpu, tmp_path):
    """isVisSol: the initial, every outputFrequency-th and final frames, in
    both formats.  The written values are the oracle's: its state after the
    same steps, taken to the Lagrange points by its own
    projectDataDisc2DataLagr (core/projectDataDisc2DataLagr.m, pinned on the
    CPU against point evaluation), in the writers' node order
    (core/visualizeDataLagr.m:125-130, 214-219).  Values are written with ten
    significant digits, hence rtol 1e-9."""
    import festung_amd
    from oracle import limiter as OL
    base = str(tmp_path / 'vis' / 'adv')
    kw = dict(testcase='solid_body', p=2, hmax=1 / 8, numSteps=4)
    festung_amd.main('advection', dict(kw, isVisSol=True, isVisGrid=True, outputFrequency=2, outputBasename=base,
                                       outputTypes=['vtk', 'tec']))
    names = sorted(p.name for p in (tmp_path / 'vis').iterdir())
    assert names == ['adv.%d.%s' % (n, e) for n in (0, 2, 4) for e in ('plt', 'vtu')] + ['adv_grid.png']
    o = OA.main('advection', dict(kw))
    want = OL.projectDataDisc2DataLagr(o.cDisc)[:, [0, 1, 2, 5, 3, 4]].ravel()
    import xml.etree.ElementTree as ET
    piece = ET.parse(base + '.4.vtu').getroot().find('UnstructuredGrid/Piece')
    u = np.array([a for a in piece.iter('DataArray') if a.get('Name') == 'u_h'][0].text.split(), dtype=float)
    assert np.allclose(u, want, rtol=1e-9, atol=1e-13)
    pts = np.array(piece.find('Points/DataArray').text.split(), dtype=float).reshape(-1, 3)
    V = o.g.coordV[o.g.V0T]                                   # K x 3 x 2 vertices
    mid = 0.5 * (V[:, [0, 1, 2]] + V[:, [1, 2, 0]])           # edge midpoints, edges 1-2, 2-3, 3-1
    assert np.allclose(pts[0::6, :2], V[:, 0], atol=1e-12) and np.allclose(pts[1::6, :2], V[:, 1], atol=1e-12)
    assert np.allclose(pts[3::6, :2], mid[:, 0], atol=1e-12) and np.allclose(pts[5::6, :2], mid[:, 2], atol=1e-12)
    with open(base + '.4.plt') as f:
        body = f.read().split('\n\n')
    nums = np.array(' '.join(body[1:]).split(), dtype=float)
    nP = 6 * o.K
    assert np.allclose(nums[2 * nP:3 * nP], want, rtol=1e-9, atol=1e-13)


def _delaunay(n, seed):
    from scipy.spatial import Delaunay
    rng = np.random.default_rng(seed)
    pts = np.vstack([rng.uniform(0, 1, (n, 2)), [[0, 0], [1, 0], [0, 1], [1, 1]]])
    tri = Delaunay(pts).simplices.astype(np.int64)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cw = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]) < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]
    return pts, tri


@pytest.mark.parametrize('p,ordRK,lim', [(1, 2, True), (2, 3, True), (3, 3, False), (2, 3, False), (3, 3, True),
                                         (4, 4, True), (4, 3, False)])
def test_unstructured_mesh_steps(gpu, p, ordRK, lim):
    """Random Delaunay meshes (irregular vertex degrees, every orientation of
    the neighbour's local edge): four steps against the oracle.  The p = 4
    mesh holds slivers (cond(B) up to 1447), on which the dc/dt limiter's
    change to the Taylor basis and back is ill-conditioned on both sides:
    the product's closed form rounds like cond(B)^p eps in double, so for
    elements with cond(B)^p > 4096 it runs in double-double
    (limit_dot_fixup_kernel), and the oracle's mass-matrix solves (cond(M_T)
    up to 4e19) lose digits too.  On element 319 (cond(B) 105, limited with
    alpha = 0 in the first stage) the same limiter evaluated in 50-digit
    arithmetic puts the oracle's limited dc/dt 2.5e-8 (relative) from the
    exact value and the product's 1.8e-9, so that case is held to 1e-11 per
    step (measured: 1.6e-12 after one step); without the limiter, and for
    p <= 3, 1e-12."""
    tol = 1e-11 if (p == 4 and lim) else TOL
    from festung_amd.core import Grid
    from oracle import grid as OG
    pts, tri = _delaunay(400, 7 + p)
    og = OG.generateGridData(pts, tri)
    idE = (og.E0E[:, 1] < 0).astype(np.int32)       # every boundary edge Dirichlet
    og.idE = idE
    og.idE0T = idE[og.E0T]
    g = Grid.generateGridData(pts, tri)
    g.set_boundary(idE)
    tc = TC.rotating_humps()
    kw = dict(p=p, hmax=0.05, ordRK=ordRK, isSlopeLim=lim, numSteps=4, tEnd=4 * 0.1 * 0.02 / (2 * p + 1))
    ocfg = OA.Struct(testcase='custom', **kw)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        ocfg[k] = tc[k]
    ocfg = OA.configureProblem(ocfg)
    ocfg.generateGridData = lambda h: og
    o = OA.preprocessProblem(ocfg)
    o = OA.initializeProblem(o)
    d = S.Struct(testcase='custom', useGraph=True, **kw)
    d.update(tc)
    d = S.configureProblem(d)
    d.generateGridData = lambda h: g
    d = S.preprocessProblem(d)
    d = S.initializeProblem(d)
    assert rel_l2(d.problem.get_state(), o.cDisc) < TOL
    c = o.cDisc.copy()
    for n in range(1, 5):
        c = oracle_step(o, c, n)
        got = product_step(d, n)
        assert rel_l2(got, c) < tol * n, n


def test_sliver_limited_dcdt_near_exact(gpu):
    """Element 319 of the p = 4 Delaunay mesh above (cond(B) 105) is limited
    with alpha = 0 in the first stage, where the dc/dt limiter's change of
    basis and mass lumping are ill-conditioned (cond(M_T) 3.9e19).  Its
    limited dc/dt, read back from the first RK4 stage (c_1 = c_0 + tau/2 y),
    against the same limiter evaluated in 50-digit arithmetic
    (tests/golden/make_sliver_exact.py): the double-double fix-up lands at
    1.8e-9 (relative), the oracle at 2.5e-8; with F and Finv as plain doubles
    the fix-up sat at 6.2e-9.  Held to 3e-9 and to a quarter of the oracle's
    error."""
    import json
    from festung_amd.core import Grid
    from oracle import grid as OG
    fx = json.load(open(os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'sliver_p4_exact.json')))
    p, k = fx['p'], fx['element']
    pts, tri = _delaunay(fx['points'], fx['seed'])
    og = OG.generateGridData(pts, tri)
    idE = (og.E0E[:, 1] < 0).astype(np.int32)
    g = Grid.generateGridData(pts, tri)
    g.set_boundary(idE)
    tc = TC.rotating_humps()
    kw = dict(p=p, hmax=0.05, ordRK=4, isSlopeLim=True, numSteps=4, tEnd=4 * 0.1 * 0.02 / (2 * p + 1))
    d = S.Struct(testcase='custom', useGraph=False, **kw)
    d.update(tc)
    d = S.configureProblem(d)
    d.generateGridData = lambda h: g
    d = S.initializeProblem(S.preprocessProblem(d))
    assert d.tau == fx['tau']
    c0 = d.problem.get_state()
    d.t, d.omega = S.rungeKuttaExplicit(d.ordRK, d.tau, 0.0)
    d.problem.begin_step()
    S.preprocessSubStep(d, 1, 1)
    S.solveSubStep(d, 1, 1)
    y = (d.problem.get_state()[k] - c0[k]) / (0.5 * d.tau)
    ex = np.asarray(fx['limited_exact'])
    err = np.linalg.norm(y - ex) / np.linalg.norm(ex)
    print('element %d: product %.2e, oracle %.2e from the 50-digit value' % (k, err, fx['oracle_rel_error']))
    assert err < 3e-9
    assert err < fx['oracle_rel_error'] / 4


def test_mid_size_step_parity(gpu):
    """One SSP-RK3 step with the limiter at K = 32 768 (hmax 2^-7: hundreds
    of workgroups over the eight XCD chunks, Hilbert element order, patch
    bounds across workgroups) against the oracle."""
    tc = TC.rotating_humps()
    o = oracle_problem(tc, 2, 1 / 128, 3, True, numSteps=20, tEnd=0.01)
    d = product_problem(tc, 2, 1 / 128, 3, True, numSteps=20, tEnd=0.01)
    assert d.problem.K == 32768
    c = o.cDisc.copy()
    d.problem.set_state(c)
    got = product_step(d, 1)
    ref = oracle_step(o, c, 1)
    assert rel_l2(got, ref) < TOL


@pytest.mark.parametrize('p,ordRK,hexp,K', [(2, 3, 9, 524288), (4, 4, 8, 131072)])
def test_full_size_free_stream(gpu, p, ordRK, hexp, K):
    """At the benchmark sizes (p = 2: K = 524 288, hmax 2^-9, SSP-RK3; p = 4:
    K = 131 072, hmax 2^-8, RK4 with the MFMA element term; limiter on,
    captured graph) c = 1 with inflow c_D = 1 under the rotation
    u = (0.5 - y, x - 0.5) -- linear and divergence-free, so represented
    exactly -- stays 1 to rounding over three steps: every element's flux
    balance closes, whatever workgroup, XCD chunk or patch it lies in."""
    import festung_amd
    zero = lambda t, x1, x2: np.zeros_like(x1)
    tc = TC.Testcase(isAnalytical=False, c0Cont=lambda x1, x2: np.ones_like(x1),
                     u1Cont=lambda t, x1, x2: 0.5 - x2, u2Cont=lambda t, x1, x2: x1 - 0.5,
                     fCont=zero, cDCont=lambda t, x1, x2: np.ones_like(x1), gNCont=zero,
                     fZero=True, cDZero=False, isStationary=True)
    h = 2.0 ** -hexp
    d = festung_amd.main('advection', dict(testcase='custom', p=p, hmax=h, ordRK=ordRK, isSlopeLim=True, numSteps=3,
                                           tEnd=3 * 0.4 * h / (0.8 * (2 * p + 1)), useGraph=True, **tc))
    c = d.cDisc
    assert c.shape == (K, (p + 1) * (p + 2) // 2)
    assert np.abs(c[:, 1:]).max() < 1e-12
    assert np.abs(c[:, 0] / np.sqrt(0.5) - 1).max() < 1e-12

"""Parity on the benchmark workloads themselves, against the oracle.

* The p = 4 workload (32-mode unit-speed stream-function velocity carrying
  disc / square indicators, RK4, the 'linear' vertex limiter on dc/dt and
  the solution): several steps at hmax 1/8 and 1/16 with the oracle run
  live, checking the state per step and every element's limiter decision
  (alpha of core/computeVertexBasedLimiter.m:94-105 in the last solution
  limiting).
* At the benchmark sizes (K = 524 288 p = 2 SSP-RK3 and K = 131 072 p = 4
  RK4, limiter on; K = 32 768 p = 4) against fixtures the oracle wrote
  (tests/golden/make_fullsize.py): the device's initial and final states are
  reduced the same way (Frobenius norm, block sums, a seeded element sample,
  the set of elements with alpha < 1).  A state within relative L2 distance
  eps of the oracle's has sqrt(sum_b |block-sum difference|^2 / m_b) <= eps
  ||ref|| and ||sample difference|| <= eps ||ref||; those are asserted with
  eps = 1e-12 per step.
"""
import os

import numpy as np
import pytest

from festung_amd.advection import steps as S
from festung_amd.advection import testcases as TC
from helpers import oracle_alpha, oracle_problem, rel_l2
from oracle import advection as OA

pytestmark = pytest.mark.gpu

TOL = 1e-12
# alpha = dMin / (d -/+ tol) (core/computeVertexBasedLimiter.m:100-101) with
# tol = 1e-8: where the data are locally flat, dMin and d are rounding-level
# (1e-16 of the data), so a rounding difference moves alpha by up to about
# 1e-16 / 1e-8 times the data scale; the state it scales (slopes of that
# size) is checked by the state tolerance.  Decisions (alpha < 1) exactly.
ALPHA_TOL = 1e-6
GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def _product(tc, p, hmax, ordRK, numSteps, tEnd, graph):
    pd = S.Struct(testcase='custom', p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=True, typeSlopeLim='linear',
                  numSteps=numSteps, tEnd=tEnd, useGraph=graph)
    pd.update(tc)
    for f in (S.configureProblem, S.preprocessProblem, S.initializeProblem):
        pd = f(pd)
    return pd


@pytest.mark.parametrize('hmax,nsteps,graph', [(1 / 8, 6, False), (1 / 16, 4, True)])
def test_p4_discontinuous_workload_steps(gpu, hmax, nsteps, graph):
    tc = TC.random_discontinuous()
    tau = 0.1 * hmax / 9.0           # CFL 0.1 h / (max|u| (2p+1)), max|u| = 1
    o = oracle_problem(tc, 4, hmax, 4, True, numSteps=100, tEnd=100 * tau)
    d = _product(tc, 4, hmax, 4, 100, 100 * tau, graph)
    assert rel_l2(d.problem.get_state(), o.cDisc) < TOL
    flips = limited = 0
    c = o.cDisc.copy()
    for n in range(1, nsteps + 1):
        o.cDisc = c
        o = OA.solveStep(o, n)
        c = o.cDiscRK[-1].reshape(o.K, o.N)
        d = S.solveStep(d, n)
        got = d.problem.get_state()
        assert rel_l2(got, c) < TOL * n, n
        aref = oracle_alpha(o, *o.lastLimiterInput)
        alpha = d.problem.limiter_info()[0]
        flips += int(np.sum((alpha < 1) != (aref < 1)))
        limited += int(np.sum(aref < 1))
        assert np.max(np.abs(alpha - aref)) < ALPHA_TOL, n
    print('p=4 K=%d: %d limited element-steps, %d alpha flips' % (o.K, limited, flips))
    assert limited > 0 and flips == 0


def _reduce(c, f):
    edges = f['block_edges']
    blocks = np.add.reduceat(c, edges[:-1], axis=0)
    return np.linalg.norm(c), blocks, c[f['sample_idx']], np.diff(edges)


def _check(c, f, tag, eps):
    fro, blocks, sample, m = _reduce(c, f)
    ref_fro = float(f[tag + '_fro'])
    assert abs(fro - ref_fro) <= eps * ref_fro, (tag, fro, ref_fro)
    db = blocks - f[tag + '_blocks']
    agg = np.sqrt(np.sum(np.sum(db ** 2, axis=1) / m))
    assert agg <= eps * ref_fro, (tag, agg / ref_fro)
    ds = np.linalg.norm(sample - f[tag + '_sample'])
    assert ds <= eps * ref_fro, (tag, ds / ref_fro)
    return agg / ref_fro, ds / ref_fro


@pytest.mark.parametrize('name', ['c3mid', 'c3', 'c2'])
def test_full_size_against_oracle_fixture(gpu, name):
    path = os.path.join(GOLDEN, 'fullsize_%s.npz' % name)
    f = np.load(path)
    tc = TC.TESTCASES[str(f['testcase'])]()
    p, ordRK, steps = int(f['p']), int(f['ordRK']), int(f['steps'])
    d = _product(tc, p, 2.0 ** -int(f['hexp']), ordRK, int(f['numSteps']), float(f['tEnd']), True)
    assert d.problem.K == int(f['K'])
    _check(d.problem.get_state(), f, 'init', TOL)
    for n in range(1, steps + 1):
        d = S.solveStep(d, n)       # step 1 stage by stage, later steps from the captured graph
    got = d.problem.get_state()
    r = _check(got, f, 'final', TOL * steps)
    alpha = d.problem.limiter_info()[0]
    ref_bits = np.unpackbits(f['alpha_lt1'])[:int(f['K'])].astype(bool)
    flips = int(np.sum((alpha < 1) != ref_bits))
    print('%s: K=%d, %d elements with alpha < 1, %d flips, block / sample rel diff %.2e / %.2e'
          % (name, d.problem.K, int(ref_bits.sum()), flips, r[0], r[1]))
    assert ref_bits.sum() > 0 and flips == 0
    assert np.max(np.abs(alpha[f['sample_idx']] - f['alpha_sample'])) < ALPHA_TOL

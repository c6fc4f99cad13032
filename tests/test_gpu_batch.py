"""Batched form: independent problems of uneven size and degree joined into
one device problem per degree and advanced as one graph launch give exactly
(bitwise) the states of the problems run one by one, and one batched
problem matches the oracle."""
import numpy as np
import pytest

from festung_amd import batch as BT
from helpers import rel_l2
from oracle import advection as OA

pytestmark = pytest.mark.gpu


@pytest.fixture(scope='module')
def inputs():
    specs = BT.batch_specs(n=9, seed0=77, kmin=64, kmax=2048)
    assert len({s['p'] for s in specs}) >= 2
    return [BT.ProblemInputs(s) for s in specs]


def test_batched_equals_individual(gpu, inputs):
    b = BT.BatchedProblems(inputs)
    b.step(6)
    b.sync()
    for i, inp in enumerate(inputs):
        single = BT.make_single(inp)
        single.step(6, inp.tau)
        assert np.array_equal(b.state(i), single.get_state()), i


def test_batched_graph_relaunch(gpu, inputs):
    """The cached graph relaunched (3 + 3 steps) equals one launch of 6."""
    a = BT.BatchedProblems(inputs)
    a.step(3)
    a.step(3)
    a.sync()
    c = BT.BatchedProblems(inputs)
    c.step(6)
    c.sync()
    for i in range(len(inputs)):
        assert np.array_equal(a.state(i), c.state(i))


def test_batched_problem_matches_oracle(gpu, inputs):
    inp = inputs[0]
    b = BT.BatchedProblems(inputs)
    c0 = b.state(0).copy()
    b.step(1)
    b.sync()
    tc = inp.tc
    pd = OA.Struct(testcase='custom', p=inp.p, hmax=inp.spec['hmax'], ordRK=inp.spec['ordRK'],
                   numSteps=1, tEnd=inp.tau)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        pd[k] = tc[k]
    for f in (OA.configureProblem, OA.preprocessProblem):
        pd = f(pd)
    pd.cDisc = c0
    pd = OA.solveStep(pd, 1)
    assert rel_l2(b.state(0), pd.cDiscRK[-1].reshape(pd.K, pd.N)) < 1e-12

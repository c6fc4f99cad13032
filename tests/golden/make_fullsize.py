#!/usr/bin/env python3
"""Write the full-size parity fixtures: the oracle (oracle/, the CPU
restatement of the reference's advection path) run at the benchmark mesh
sizes, reduced to size-independent checks that a GPU test can compare
against without running the oracle again.

Each case runs configureProblem -> preprocessProblem -> initializeProblem
(projection of c0 and the initial limiting) and then ``steps`` time steps of
advection/solveStep.m with the limiter on.  The fixture holds, for the
initial and the final state (K x N, reference element order):

* ``fro``: the Frobenius norm;
* ``blocks``: sums of every coefficient over NB blocks of consecutive
  elements (NB x N);
* ``sample``: the coefficients of a seeded sample of elements;
* for the final step's last solution limiting (the one whose alpha the
  device keeps): ``alpha_lt1`` (packed bits: alpha < 1 per element) and the
  sample's alpha values.

A device state within relative L2 distance eps of the oracle's satisfies
|block sum difference| <= sqrt(block size) eps fro and ||sample difference||
<= eps fro, which is what the GPU test asserts.

    python tests/golden/make_fullsize.py [case ...]     (cases: c2 c3 c3mid)
"""
import os
import sys
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

sys.path.insert(0, os.path.dirname(HERE))
from festung_amd.advection import testcases as TC  # noqa: E402  (closures only)
from helpers import oracle_alpha  # noqa: E402
from oracle import advection as OA  # noqa: E402

NB = 1024        # element blocks
NSAMPLE = 2048   # sampled elements

# name -> (testcase, p, hexp, ordRK, steps); tau = 0.1 h / (max|u| (2p+1)),
# max|u| over the vertices, as the benchmark
CASES = {
    'c2': ('rotating_humps', 2, 9, 3, 2),          # K = 524 288, SSP-RK3
    'c3': ('random_discontinuous', 4, 8, 4, 2),    # K = 131 072, RK4
    'c3mid': ('random_discontinuous', 4, 7, 4, 1),  # K = 32 768, RK4
}


def bench_tau(tc, hmax):
    """0.1 h / (max|u| (2p+1)) without the (2p+1): the caller divides."""
    dim = int(np.ceil(1.0 / hmax))
    n = dim + 1
    xs = np.linspace(0.0, 1.0, n)
    X, Y = np.meshgrid(xs, xs)
    umax = float(np.max(np.hypot(tc.u1Cont(0, X, Y), tc.u2Cont(0, X, Y))))
    return 0.1 * (1.0 / dim) / umax


def reduce_state(c, idx):
    K = c.shape[0]
    edges = np.linspace(0, K, NB + 1).astype(np.int64)
    blocks = np.add.reduceat(c, edges[:-1], axis=0)
    return dict(fro=np.linalg.norm(c), blocks=blocks, sample=c[idx], block_edges=edges)


def run(name):
    tcname, p, hexp, ordRK, steps = CASES[name]
    tc = TC.TESTCASES[tcname]()
    hmax = 2.0 ** -hexp
    tau = bench_tau(tc, hmax) / (2 * p + 1)
    numSteps = 100
    tEnd = tau * numSteps
    pd = OA.Struct(testcase='custom', p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=True, typeSlopeLim='linear',
                   numSteps=numSteps, tEnd=tEnd)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        pd[k] = tc[k]
    t0 = time.time()
    for f in (OA.configureProblem, OA.preprocessProblem, OA.initializeProblem):
        pd = f(pd)
    K, N = pd.K, pd.N
    rng = np.random.Generator(np.random.PCG64(1234 + p))
    idx = np.sort(rng.choice(K, size=min(NSAMPLE, K), replace=False))
    init = reduce_state(pd.cDisc, idx)
    for n in range(1, steps + 1):
        pd = OA.solveStep(pd, n)
        pd = OA.postprocessStep(pd, n)
    final = reduce_state(pd.cDisc, idx)
    alpha = oracle_alpha(pd, *pd.lastLimiterInput)
    out = dict(name=np.array(name), testcase=np.array(tcname), p=p, hexp=hexp, ordRK=ordRK, steps=steps,
               numSteps=numSteps, tEnd=tEnd, K=K, N=N, sample_idx=idx.astype(np.int64),
               alpha_lt1=np.packbits(alpha < 1), alpha_sample=alpha[idx], n_alpha_lt1=int(np.sum(alpha < 1)),
               block_edges=init['block_edges'])
    for tag, st in (('init', init), ('final', final)):
        for k in ('fro', 'blocks', 'sample'):
            out['%s_%s' % (tag, k)] = st[k]
    path = os.path.join(HERE, 'fullsize_%s.npz' % name)
    np.savez_compressed(path, **out)
    print('%s: K=%d p=%d %d steps, %d elements with alpha < 1, %.0f s -> %s'
          % (name, K, p, steps, out['n_alpha_lt1'], time.time() - t0, path))


if __name__ == '__main__':
    for nm in (sys.argv[1:] or list(CASES)):
        run(nm)

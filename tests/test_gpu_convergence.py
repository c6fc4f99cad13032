"""Convergence orders through the device path: the reference's
advection/doConvergenceTest.m:76-104 on triangles ('transient' testcase,
advection/configureProblem.m:119-132: time-dependent source and Dirichlet
data, tEnd = 2, hmax = 2^-level / 3, numSteps = 2^p 50 2^level) run through
festung_amd.main('advection', ...).  Asserted:

* on the coarse levels the oracle could finish (tests/golden/make_eoc.py),
  the device's final L2 error (core/computeL2Error.m:97-114) equals the
  oracle's to 1e-10 relative;
* the EOC log(e_l / e_(l+1)) / log 2 between levels 2 and 3 is at least
  p + 1 - 0.3 (the asymptotic order p + 1 of the upwind DG method for
  smooth solutions, minus a stated margin for the pre-asymptotic range), and
  between levels 1 and 2 at least p + 1 - 0.6.
"""
import json
import os

import numpy as np
import pytest

pytestmark = pytest.mark.gpu

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden', 'eoc_transient.json')


def _run(p, level):
    import festung_amd
    d = festung_amd.main('advection', dict(testcase='transient', p=p, hmax=2.0 ** -level / 3,
                                           numSteps=2 ** p * 50 * 2 ** level, tEnd=2.0))
    return d.error, d.K


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_transient_eoc_through_device(gpu, p):
    ref = {(e['p'], e['level']): e for e in json.load(open(GOLDEN))['errors']}
    errs = []
    for level in (1, 2, 3):
        err, K = _run(p, level)
        errs.append(err)
        if (p, level) in ref:
            r = ref[(p, level)]
            assert K == r['K']
            assert abs(err - r['error']) <= 1e-10 * r['error'], (p, level, err, r['error'])
    eoc = np.log(np.array(errs[:-1]) / np.array(errs[1:])) / np.log(2.0)
    print('p=%d errors %s EOC %s' % (p, ['%.3e' % e for e in errs], ['%.3f' % e for e in eoc]))
    assert eoc[0] >= p + 1 - 0.6, eoc
    assert eoc[1] >= p + 1 - 0.3, eoc

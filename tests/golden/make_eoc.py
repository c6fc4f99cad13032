#!/usr/bin/env python3
"""Write tests/golden/eoc_transient.json: the oracle's final L2 errors of the
'transient' testcase (advection/configureProblem.m:119-132) on the coarse
levels of advection/doConvergenceTest.m:76-90 (hmax = 2^-level / 3,
numSteps = 2^p 50 2^level, tEnd = 2, SSP-RK min(p+1, 3), no limiter), for
the device's convergence test to compare against.  Levels the oracle cannot
finish in minutes are left out.

    python tests/golden/make_eoc.py
"""
import json
import os
import sys
import time

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.dirname(os.path.dirname(HERE)))

from oracle import advection as OA  # noqa: E402

# (p, level) pairs computed by the oracle
PAIRS = [(1, 1), (1, 2), (2, 1), (2, 2), (3, 1), (4, 1)]


def main():
    out = {'source': 'oracle/advection.py main(), testcase transient, tEnd 2, '
                     'hmax 2^-level/3, numSteps 2^p*50*2^level (advection/doConvergenceTest.m:76-90)',
           'errors': []}
    for p, lev in PAIRS:
        t0 = time.time()
        n = 2 ** p * 50 * 2 ** lev
        d = OA.main('advection', OA.Struct(testcase='transient', p=p, hmax=2.0 ** -lev / 3, numSteps=n, tEnd=2.0))
        out['errors'].append(dict(p=p, level=lev, K=int(d.K), numSteps=n, error=float(d.error)))
        print('p=%d level=%d K=%d steps=%d error=%.16e (%.0f s)' % (p, lev, d.K, n, d.error, time.time() - t0),
              flush=True)
    with open(os.path.join(HERE, 'eoc_transient.json'), 'w') as f:
        json.dump(out, f, indent=1)


if __name__ == '__main__':
    main()

"""festung_apply_operator_dev_noalloc (the entry point that allocates no
device memory) against festung_apply_operator_dev on the same inputs, bit for
bit: two calls in a row, the second on a smaller problem (the shared scratch
keeps the first call's contents), every operator; and the size limit."""
import ctypes as C

import numpy as np
import pytest

from festung_amd import _lib
from festung_amd.advection import testcases as TC

pytestmark = pytest.mark.gpu


def _problem(hmax, p):
    from test_gpu_parity import product_problem
    from festung_amd.advection import steps as S
    d = product_problem(TC.rotating_humps(), p, hmax, min(p + 1, 3), False)
    d.t = np.array([0.0])
    S.preprocessSubStep(d, 1, 1)        # sets the velocity
    return d


def _call(fn, P, op, x, y):
    import torch
    torch.cuda.synchronize()
    rc = fn(P._h, P.OPERATORS[op], C.c_void_p(x.data_ptr()), C.c_void_p(y.data_ptr()))
    _lib.check(rc)
    P.sync()


def test_noalloc_matches_allocating_entry(gpu):
    import torch
    L = _lib.lib()
    big = _problem(1 / 24, 4)          # K = 1152, N = 15
    small = _problem(1 / 8, 2)         # K = 128, N = 6
    rng = np.random.default_rng(17)
    for d in (big, small):             # the second call on a smaller input
        P = d.problem
        x = torch.tensor(rng.standard_normal((P.N, P.K)), device='cuda')
        for op in ('G1', 'G2', 'G', 'R', 'M', 'Minv'):
            y0 = torch.full_like(x, np.nan)
            y1 = torch.full_like(x, np.nan)
            _call(L.festung_apply_operator_dev, P, op, x, y0)
            _call(L.festung_apply_operator_dev_noalloc, P, op, x, y1)
            assert torch.equal(y0, y1), (P.K, op)
            assert torch.isfinite(y1).all()
        y2 = torch.full_like(x, np.nan)
        P.apply_operator_dev_noalloc('R', x, y2)
        torch.cuda.synchronize()
        assert np.array_equal(y2.cpu().numpy().T, P.apply_operator('R', x.cpu().numpy().T))


def test_noalloc_rejects_oversized_problem(gpu):
    import torch
    d = _problem(1 / 374, 4)           # K = 279752, K * N = 4196280 > 2^22
    P = d.problem
    assert P.K * P.N > 1 << 22
    x = torch.zeros((P.N, P.K), dtype=torch.float64, device='cuda')
    y = torch.full_like(x, 7.0)
    with pytest.raises(_lib.FestungError, match='exceeds'):
        P.apply_operator_dev_noalloc('G', x, y)
    torch.cuda.synchronize()
    assert bool((y == 7.0).all())

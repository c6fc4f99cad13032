"""Ordering and cache-invalidation rules of the problem handle: changes made
between captured step graphs take effect (bit for bit as eager steps do), a
new time step drops the graphs only after the launches in flight, device
pointers are checked before any launch and ordered against torch's stream,
and a batch's members' own calls run after the batch without a host sync."""
import numpy as np
import pytest

from festung_amd import batch as BT
from festung_amd.advection import testcases as TC
from test_gpu_parity import product_problem, product_step

pytestmark = pytest.mark.gpu


def test_source_change_between_graph_steps(gpu):
    """Setting and clearing a source after graph steps (the graph baked in
    whether a source exists): same bits as the eager sequence."""
    tc = TC.rotating_humps()
    outs = []
    for graph in (True, False):
        d = product_problem(tc, 2, 1 / 16, 3, True, graph=graph)
        product_step(d, 1)
        d.problem.step(2, d.tau)
        X1, X2 = d.XEdge
        fq = 0.3 * np.sin(3 * d.XE[0]) * np.cos(2 * d.XE[1])
        cDe = 0.2 + 0.1 * X1 - 0.05 * X2
        d.problem.set_source_points(fq, cDe)
        d.problem.step(2, d.tau)
        d.problem.set_source_points(0.5 * fq, None)
        d.problem.step(2, d.tau)
        d.problem.set_source_points(None, None)
        d.problem.step(2, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_tau_change_back_to_back(gpu):
    """step(8, tau1) then step(8, tau2) enqueued without a sync between: the
    first graph exec is dropped only after its launches have run."""
    tc = TC.rotating_humps()
    outs = []
    for graph in (True, False):
        d = product_problem(tc, 2, 1 / 32, 3, True, graph=graph)
        product_step(d, 1)
        d.problem.step(8, d.tau)
        d.problem.step(8, 0.5 * d.tau)
        d.problem.step(3, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_tau_field_change_between_graph_steps(gpu):
    tc = TC.rotating_humps()
    outs = []
    for graph in (True, False):
        d = product_problem(tc, 1, 1 / 16, 2, False, graph=graph)
        product_step(d, 1)
        d.problem.step(8, d.tau)
        d.problem.set_tau_field(np.linspace(0.5, 1.0, d.K) * d.tau)
        d.problem.step(8, d.tau)
        d.problem.set_tau_field(None)
        d.problem.step(2, d.tau)
        outs.append(d.problem.get_state())
    assert np.array_equal(outs[0], outs[1])


def test_state_dev_rejects_bad_tensors(gpu):
    import torch
    from festung_amd._lib import FestungError  # noqa: F401  (no launch may happen)
    d = product_problem(TC.rotating_humps(), 2, 1 / 8, 3, False)
    c0 = d.problem.get_state()
    K, N = d.K, d.N
    bad = [torch.zeros((N, K), dtype=torch.float32, device='cuda'),
           torch.zeros((N, K), dtype=torch.float64),
           torch.zeros((N, K + 1), dtype=torch.float64, device='cuda'),
           torch.zeros((K, N), dtype=torch.float64, device='cuda').t()]
    for t in bad:
        with pytest.raises(ValueError):
            d.problem.set_state_dev(t)
        with pytest.raises(ValueError):
            d.problem.get_state_dev(t)
    assert np.array_equal(d.problem.get_state(), c0)


def test_state_dev_follows_torch_stream(gpu):
    """A tensor torch has just written is read after that write; a tensor the
    problem writes is read by torch after the copy (no explicit sync)."""
    import torch
    d = product_problem(TC.rotating_humps(), 2, 1 / 32, 3, True, graph=True)
    product_step(d, 1)                     # sets the velocity
    K, N = d.K, d.N
    rng = np.random.default_rng(5)
    c = rng.standard_normal((K, N))
    src = torch.tensor(np.ascontiguousarray(c.T), device='cuda')
    for _ in range(3):
        t = torch.empty_like(src)
        t.copy_(src * 2.0)
        t.mul_(0.5)
        d.problem.set_state_dev(t)
        del t
        d.problem.step(4, d.tau)
        out = torch.empty_like(src)
        d.problem.get_state_dev(out)
        got = (out * 1.0).cpu().numpy().T
        ref = d.problem.get_state()
        assert np.array_equal(got, ref)


def test_batch_member_calls_follow_the_batch(gpu):
    """A member's get_state right after batch.step (no batch.sync) sees the
    batch's result."""
    specs = BT.batch_specs(n=6, seed0=177, kmin=64, kmax=2048)
    inputs = [BT.ProblemInputs(s) for s in specs]
    a = BT.BatchedProblems(inputs)
    assert len(a.problems) >= 2
    a.step(5)
    got = [a.state(i) for i in range(len(inputs))]
    b = BT.BatchedProblems(inputs)
    b.step(5)
    b.sync()
    for i in range(len(inputs)):
        assert np.array_equal(got[i], b.state(i)), i


def test_tuning_table_applies_to_new_problems(gpu, tmp_path):
    """The autotuner's choice is recorded under the problem's key; a table
    entry loaded from a file sets the workgroup size of problems created
    later under that key, and only under it."""
    from festung_amd import core
    core.tune_clear()
    d = product_problem(TC.rotating_humps(), 1, 1 / 16, 2, False)
    product_step(d, 1)
    assert d.problem.elem_block == 64                       # default below 1024 * 128 elements
    best, ms = d.problem.autotune(d.tau, 2)
    assert best in (64, 128, 256) and core.tune_count() == 1
    path = tmp_path / 'tune.txt'
    core.tune_save(str(path))
    rows = [l.split() for l in path.read_text().splitlines() if l and not l.startswith('#')]
    assert len(rows) == 1 and rows[0][1:5] == ['1', '2', '-1', str(int(np.ceil(np.log2(d.K))))]
    other = {64: 128, 128: 256, 256: 64}[best]
    rows[0][5] = str(other)
    path.write_text(' '.join(rows[0]) + '\n')
    core.tune_clear()
    core.tune_load(str(path))
    e = product_problem(TC.rotating_humps(), 1, 1 / 16, 2, False)
    assert e.problem.elem_block == other
    f = product_problem(TC.rotating_humps(), 2, 1 / 16, 3, False)   # another key: the default
    assert f.problem.elem_block == 64
    core.tune_clear()


def test_batch_follows_members_own_work(gpu):
    """Steps enqueued on a member's own stream right before batch.step (no
    host sync between) are done before the batch reads that member."""
    specs = BT.batch_specs(n=6, seed0=177, kmin=64, kmax=2048)
    inputs = [BT.ProblemInputs(s) for s in specs]
    out = []
    for sync_between in (False, True):
        b = BT.BatchedProblems(inputs)
        assert len(b.problems) >= 2
        m = b.problems[-1]
        m.step(3, 1e-3)
        if sync_between:
            m.sync()
        b.step(2)
        b.sync()
        out.append([b.state(i) for i in range(len(inputs))])
    for x, y in zip(*out):
        assert np.array_equal(x, y)

"""Discrete adjoint on the device: dot-product test of the transposed
operator, gradients against the oracle's transposed sparse operators,
against central finite differences, invariance to the checkpoint interval,
and the torch.autograd wrapper."""
import numpy as np
import pytest

from festung_amd import adjoint as AD
from festung_amd.advection import testcases as TC
from oracle import adjoint as OAd
from oracle import advection as OA
from oracle import assembly as OAs

pytestmark = pytest.mark.gpu


def setup(p, hmax, nmodes, seed=51):
    modes = AD.stream_modes()[:nmodes]
    m = AD.StreamModel(hmax, p, modes=modes, c0=TC.gaussian_humps(seed))
    rng = np.random.default_rng(52)
    a_true = 0.05 * rng.standard_normal(len(modes))
    a = a_true + 0.005 * np.random.default_rng(53).standard_normal(len(modes))
    tau = 0.1 / (int(np.ceil(1 / hmax)) * 0.8 * (2 * p + 1))
    return m, a_true, a, tau


def test_transposed_operator_dot_product(gpu):
    m, _, a, _ = setup(2, 1 / 8, 4)
    m.set_velocity(a)
    rng = np.random.default_rng(1)
    x = rng.standard_normal((m.prob.K, m.prob.N))
    y = rng.standard_normal((m.prob.K, m.prob.N))
    m.prob.set_state(x)
    Lx = m.prob.rhs()                       # M^-1 A x
    LTy = m.prob.adj_rhs(y)                  # M^-1 A^T y
    det = 2 * m.g.areaT[:, None]
    lhs = np.sum(y * det * Lx)               # y^T A x   (hatM = I to rounding)
    rhs = np.sum(x * det * LTy)              # x^T A^T y
    assert abs(lhs - rhs) <= 1e-12 * (abs(lhs) + abs(rhs))


def oracle_fields(m, pd, a):
    """Velocity fields of u(a) and of every mode, computed by the oracle."""
    def vel(a_):
        def u1(t, x1, x2):
            r = 0.5 - x2
            for am, (k, l) in zip(a_, m.modes):
                r = r + am * AD.mode_velocity(k, l, x1, x2)[0]
            return r

        def u2(t, x1, x2):
            r = x1 - 0.5
            for am, (k, l) in zip(a_, m.modes):
                r = r + am * AD.mode_velocity(k, l, x1, x2)[1]
            return r
        return u1, u2

    def fields(u1, u2):
        d1 = OAs.projectFuncCont2DataDisc(pd.g, lambda x1, x2: u1(0, x1, x2), 2 * pd.p, pd.hatM, pd.basesOnQuad)
        d2 = OAs.projectFuncCont2DataDisc(pd.g, lambda x1, x2: u2(0, x1, x2), 2 * pd.p, pd.hatM, pd.basesOnQuad)
        vn = OAs.computeFuncContNuOnQuadEdge(pd.g, lambda x1, x2: u1(0, x1, x2), lambda x1, x2: u2(0, x1, x2),
                                             2 * pd.p + 1)
        return d1, d2, vn
    base = fields(*vel(a))
    modes = []
    for (k, l) in m.modes:
        f1 = lambda t, x1, x2, k=k, l=l: AD.mode_velocity(k, l, x1, x2)[0]
        f2 = lambda t, x1, x2, k=k, l=l: AD.mode_velocity(k, l, x1, x2)[1]
        modes.append(fields(f1, f2))
    return base, modes


@pytest.mark.parametrize('p', [1, 2])
def test_adjoint_matches_oracle(gpu, p):
    m, a_true, a, tau = setup(p, 1 / 8, 4)
    nsteps = 8
    cObs = m.run(m.c0, a_true, nsteps, tau)
    J, g0, ga, _ = m.gradient(m.c0, a, nsteps, tau, cObs, ckpt=3)
    z = lambda t, x1, x2: np.zeros_like(x1)
    pd = OA.Struct(testcase='custom', p=p, hmax=1 / 8, ordRK=3, numSteps=nsteps, c0Cont=z, u1Cont=z, u2Cont=z,
                   fCont=z, cDCont=z, gNCont=z)
    pd = OA.preprocessProblem(OA.configureProblem(pd))
    (u1d, u2d, vn), modes = oracle_fields(m, pd, a)
    Jo, g0o, gao = OAd.gradient(pd, u1d, u2d, vn, modes, m.c0, nsteps, tau, cObs)
    assert abs(J - Jo) <= 1e-11 * Jo
    assert np.linalg.norm(g0 - g0o) <= 1e-10 * np.linalg.norm(g0o)
    assert np.linalg.norm(ga - gao) <= 1e-9 * np.linalg.norm(gao)


def test_adjoint_checkpoint_interval_invariance(gpu):
    m, a_true, a, tau = setup(2, 1 / 16, 4)
    cObs = m.run(m.c0, a_true, 12, tau)
    r0 = m.gradient(m.c0, a, 12, tau, cObs, ckpt=0)
    r1 = m.gradient(m.c0, a, 12, tau, cObs, ckpt=5)
    assert r0[0] == r1[0]
    assert np.array_equal(r0[1], r1[1]) and np.array_equal(r0[2], r1[2])


def test_adjoint_vs_central_differences(gpu):
    m, a_true, a, tau = setup(2, 1 / 32, 8)
    nsteps = 40
    cObs = m.run(m.c0, a_true, nsteps, tau)
    J, g0, ga, _ = m.gradient(m.c0, a, nsteps, tau, cObs, ckpt=10)
    rng = np.random.default_rng(7)
    h = 1e-6
    for _ in range(4):
        d0 = rng.standard_normal(g0.shape) * np.abs(m.c0).max()
        da = rng.standard_normal(ga.shape) * 0.05
        Jp = m.misfit(m.c0 + h * d0, a + h * da, nsteps, tau, cObs)
        Jm = m.misfit(m.c0 - h * d0, a - h * da, nsteps, tau, cObs)
        fd = (Jp - Jm) / (2 * h)
        an = np.sum(g0 * d0) + np.sum(ga * da)
        assert abs(fd - an) <= 1e-6 * abs(an), (fd, an)


def test_autograd_wrapper(gpu):
    import torch
    m, a_true, a, tau = setup(1, 1 / 8, 4)
    cObs = m.run(m.c0, a_true, 5, tau)
    F = AD.misfit_function()
    c0 = torch.tensor(np.ascontiguousarray(m.c0), requires_grad=True)
    at = torch.tensor(a, requires_grad=True)
    J = F.apply(c0, at, m, 5, tau, cObs, 0)
    J.backward()
    _, g0, ga, _ = m.gradient(m.c0, a, 5, tau, cObs)
    assert np.allclose(c0.grad.numpy(), g0, rtol=0, atol=0)
    assert np.allclose(at.grad.numpy(), ga, rtol=0, atol=0)


def test_autograd_wrapper_on_device(gpu):
    """c0, c_obs and the gradients as CUDA tensors ([N][K]) through
    festung_adjoint_dev: the same bits as the host entry point (the velocity
    is formed on the device in the host's order of operations); with no
    gradient requested the forward run alone gives J."""
    import torch
    m, a_true, a, tau = setup(1, 1 / 8, 4)
    cObs = m.run(m.c0, a_true, 5, tau)
    F = AD.misfit_function()
    c0 = torch.tensor(np.ascontiguousarray(m.c0.T), device='cuda', requires_grad=True)
    at = torch.tensor(a, device='cuda', requires_grad=True)
    cO = torch.tensor(np.ascontiguousarray(cObs.T), device='cuda')
    J = F.apply(c0, at, m, 5, tau, cO, 0)
    J.backward()
    Jh, g0, ga, _ = m.gradient(m.c0, a, 5, tau, cObs)
    assert J.is_cuda and c0.grad.is_cuda and at.grad.is_cuda
    assert J.item() == Jh
    assert np.array_equal(c0.grad.cpu().numpy().T, g0)
    assert np.allclose(at.grad.cpu().numpy(), ga, rtol=1e-13, atol=0)
    J2 = F.apply(c0.detach(), at.detach(), m, 5, tau, cO, 0)
    assert J2.is_cuda and abs(J2.item() - Jh) <= 1e-12 * abs(Jh)


def test_adjoint_elem_block_invariance(gpu):
    """The transposed stage kernels, with their fused velocity sensitivities,
    give the same gradient bit for bit for every workgroup size."""
    m, a_true, a, tau = setup(2, 1 / 24, 4)
    cObs = m.run(m.c0, a_true, 6, tau)
    res = []
    for bs in (64, 128, 256):
        m.prob.elem_block = bs
        res.append(m.gradient(m.c0, a, 6, tau, cObs, ckpt=2))
    for r in res[1:]:
        assert r[0] == res[0][0]
        assert np.array_equal(r[1], res[0][1]) and np.array_equal(r[2], res[0][2])


@pytest.mark.parametrize('nsteps,ckpt', [(12, 2), (12, 5), (7, 0), (9, 4), (10, 3)])
def test_adjoint_variants_agree(gpu, nsteps, ckpt):
    """Every way of running the sweep gives the same results: bit for bit
    with the velocity sensitivities fused into the transposed stage kernel or
    on more threads per element, with checkpointed segments recomputed on the
    problem's stream or on a second stream overlapping the sweep.  Deferred
    to one pass per checkpoint segment, the sensitivities are the same
    per-stage terms summed per time step first (another fixed order): equal
    to 1e-13 relative, and bit for bit between its serial and overlapped
    sweeps."""
    res = []
    kinds = ((True, 'fused'), (False, 'fused'), (False, 'split'), (True, 'split'),
             (False, 'deferred'), (True, 'deferred'))
    for serial, kern in kinds:
        modes = AD.stream_modes()[:4]
        m = AD.StreamModel(1 / 16, 2, modes=modes, c0=TC.gaussian_humps(51), adj_serial=serial, adj_kernel=kern)
        a = 0.05 * np.random.default_rng(52).standard_normal(len(modes))
        tau = 0.1 / (16 * 0.8 * 5)
        cObs = m.run(m.c0, 1.1 * a, nsteps, tau)
        res.append(m.gradient(m.c0, a, nsteps, tau, cObs, ckpt=ckpt))
    for (serial, kern), r in zip(kinds[1:], res[1:]):
        assert r[0] == res[0][0]
        assert np.array_equal(r[1], res[0][1])
        ref = res[4] if kern == 'deferred' else res[0]
        assert np.array_equal(r[2], ref[2])
        for key in ('gradU1', 'gradU2', 'gradVn'):
            assert np.array_equal(r[3][key], ref[3][key])
    d, f = res[4], res[0]
    assert np.allclose(d[2], f[2], rtol=0, atol=1e-13 * np.abs(f[2]).max())
    for key in ('gradU1', 'gradU2', 'gradVn'):
        assert np.allclose(d[3][key], f[3][key], rtol=0, atol=1e-13 * np.abs(f[3][key]).max())


@pytest.mark.parametrize('kern', ['fused', 'split'])
def test_adjoint_binomial_checkpointing(gpu, kern):
    """Binomial checkpointing (state snapshots only, every step recomputed
    with its stage states before it is reversed) runs the same kernels in the
    same order as the uniform checkpoints: the same bits for every number of
    snapshots, from one (walk from the start) to one per step."""
    nsteps = 13
    modes = AD.stream_modes()[:4]
    a = 0.05 * np.random.default_rng(52).standard_normal(len(modes))
    tau = 0.1 / (16 * 0.8 * 5)

    def run(binomial, ckpt):
        m = AD.StreamModel(1 / 16, 2, modes=modes, c0=TC.gaussian_humps(51), adj_kernel=kern,
                           adj_binomial=binomial)
        cObs = m.run(m.c0, 1.1 * a, nsteps, tau)
        return m.gradient(m.c0, a, nsteps, tau, cObs, ckpt=ckpt)
    ref = run(False, 5)
    for snaps in (1, 2, 3, 6, nsteps):
        r = run(True, snaps)
        assert r[0] == ref[0]
        assert np.array_equal(r[1], ref[1]) and np.array_equal(r[2], ref[2])
        for key in ('gradU1', 'gradU2', 'gradVn'):
            assert np.array_equal(r[3][key], ref[3][key])

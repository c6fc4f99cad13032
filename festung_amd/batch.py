"""Batched form: many independent advection problems of uneven mesh size and
polynomial degree advanced together.

Problems of the same degree are joined into one device problem on the
disjoint union of their meshes (``Grid.concat``), each keeping its own time
step through a per-element tau field; the degree groups are the parallel
branches of one captured hipGraph (``Batch``), so a whole batch of time steps
is a single launch.  Without the slope limiter the arithmetic of every
element is exactly that of the problem run alone.

The reference solves one problem per ``main`` run; this is an extension for
throughput (no reference counterpart).
"""
import numpy as np

from .core import Batch, Grid, Problem, quadRule1D, quadRule2D
from .advection import testcases as TC


def gammaMap(n, S):
    S = np.asarray(S, dtype=float).ravel()
    if n == 1:
        return 1 - S, S
    if n == 2:
        return np.zeros_like(S), 1 - S
    return S, np.zeros_like(S)


def batch_specs(n=512, seed0=4000, kmin=1024, kmax=65536, degrees=(1, 2, 3)):
    """Problem i (seed seed0 + i): K_target ~ logU[kmin, kmax] -> dim =
    round(sqrt(K_target / 2)), hmax = 1 / dim; p uniform over ``degrees``;
    ordRK = min(p + 1, 3); velocity and humps seeded from the same seed."""
    specs = []
    for i in range(n):
        rng = np.random.Generator(np.random.PCG64(seed0 + i))
        kt = np.exp(rng.uniform(np.log(kmin), np.log(kmax)))
        dim = max(2, int(round(np.sqrt(kt / 2))))
        p = int(degrees[rng.integers(0, len(degrees))])
        specs.append(dict(index=i, dim=dim, hmax=1.0 / dim, p=p, ordRK=min(p + 1, 3),
                          vel_seed=seed0 + i, hump_seed=10 ** 6 + seed0 + i))
    return specs


def _evaluate(fn, X, device):
    """fn on host coordinate arrays, optionally through torch on the GPU."""
    if device is None:
        return fn(*X)
    import torch
    T = [torch.as_tensor(np.asarray(x), device=device) for x in X]
    return fn(*T).cpu().numpy()


class ProblemInputs:
    """Grid, closures, evaluation points and time step of one problem."""

    def __init__(self, spec, cfl=0.1, device=None):
        self.spec = spec
        self.p = spec['p']
        self.g = Grid.domainSquare(spec['hmax'])
        self.tc = TC.rotating_humps(vel_seed=spec['vel_seed'], hump_seed=spec['hump_seed'])
        p = self.p
        Q1, Q2, _ = quadRule2D(max(2 * p, 1))
        XE = self.g.mapRef2Phy(Q1, Q2)
        Q1, Q2, _ = quadRule2D(2 * p + 1)
        XF = self.g.mapRef2Phy(Q1, Q2)
        Q, _ = quadRule1D(2 * p + 1)
        K = self.g.numT
        X1 = np.empty((K, 3, len(Q)), order='F'); X2 = np.empty_like(X1)
        for n in range(3):
            A1, A2 = gammaMap(n + 1, Q)
            X1[:, n, :], X2[:, n, :] = self.g.mapRef2Phy(A1, A2)
        u1, u2, c0 = self.tc.u1Cont, self.tc.u2Cont, self.tc.c0Cont
        self.u1q = _evaluate(lambda a, b: u1(0, a, b), XE, device)
        self.u2q = _evaluate(lambda a, b: u2(0, a, b), XE, device)
        self.u1e = _evaluate(lambda a, b: u1(0, a, b), (X1, X2), device)
        self.u2e = _evaluate(lambda a, b: u2(0, a, b), (X1, X2), device)
        self.c0q = _evaluate(c0, XF, device)
        xy = self.g.coordV
        umax = float(np.max(np.hypot(u1(0, xy[:, 0], xy[:, 1]), u2(0, xy[:, 0], xy[:, 1]))))
        self.tau = cfl * spec['hmax'] / (umax * (2 * p + 1))
        self.K = K


def make_single(inp, graph=True):
    """The problem alone on the device (the unbatched reference path)."""
    prob = Problem(inp.g, inp.p, inp.spec['ordRK'], graph=graph)
    prob.set_velocity_points(inp.u1q, inp.u2q, inp.u1e, inp.u2e)
    prob.project(prob.ordF, inp.c0q, set_state=True)
    return prob


class BatchedProblems:
    """Union problems per degree, advanced as one graph launch."""

    def __init__(self, inputs):
        self.inputs = list(inputs)
        self.groups = {}
        for idx, inp in enumerate(self.inputs):
            self.groups.setdefault((inp.p, inp.spec['ordRK']), []).append(idx)
        self.problems = []
        self.where = {}
        for (p, ordRK), members in sorted(self.groups.items()):
            grids = [self.inputs[i].g for i in members]
            union = Grid.concat(grids)
            prob = Problem(union, p, ordRK)
            cat = lambda name: np.concatenate([getattr(self.inputs[i], name) for i in members], axis=0)
            prob.set_velocity_points(cat('u1q'), cat('u2q'), cat('u1e'), cat('u2e'))
            prob.project(prob.ordF, cat('c0q'), set_state=True)
            prob.set_tau_field(np.concatenate([np.full(self.inputs[i].K, self.inputs[i].tau) for i in members]))
            off = 0
            for i in members:
                self.where[i] = (len(self.problems), off, self.inputs[i].K)
                off += self.inputs[i].K
            self.problems.append(prob)
        self.batch = Batch(self.problems)

    def step(self, nsteps):
        """nsteps time steps of every problem: one graph launch."""
        self.batch.step(nsteps, np.zeros(len(self.problems)))

    def sync(self):
        self.batch.sync()

    def state(self, i):
        """cDisc (K_i x N_i) of problem i."""
        g, off, K = self.where[i]
        return self.problems[g].get_state()[off:off + K]

    @property
    def dofs(self):
        return sum(inp.K * (inp.p + 1) * (inp.p + 2) // 2 for inp in self.inputs)

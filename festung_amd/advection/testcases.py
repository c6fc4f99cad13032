"""Testcases of the advection problem: the reference's own
('solid_body', 'transient'; advection/configureProblem.m:105-136) and the
synthetic benchmark workloads (seeded stream-function velocities, Gaussian
humps, discontinuous indicators).

Every closure follows the reference's calling convention: c0(x1, x2),
u1/u2/f/cD/gN(t, x1, x2), elementwise on numpy arrays.  A testcase also
states which closures are identically zero and whether the data depend on
time, so the host can skip uploading zero or unchanged data.
"""
import math

import numpy as np


class Testcase(dict):
    __getattr__ = dict.__getitem__


def _xp(x):
    """Array namespace of x: the closures accept numpy arrays and, for fast
    set-up of large batches, torch tensors (evaluated where they live)."""
    if type(x).__module__.startswith('torch'):
        import torch
        return torch
    return np


def _zeros(t, x1, x2):
    return _xp(x1).zeros_like(x1)


def solid_body():
    """LeVeque's solid body rotation; advection/configureProblem.m:106-117."""
    G = lambda x1, x2, a, b: (1 / 0.15) * np.sqrt((x1 - a) ** 2 + (x2 - b) ** 2)
    c0 = lambda x1, x2: ((((x1 - 0.5) ** 2 + (x2 - 0.75) ** 2 <= 0.0225)
                          & ((x1 <= 0.475) | (x1 >= 0.525) | (x2 >= 0.85))).astype(float)
                         + (1 - G(x1, x2, 0.5, 0.25)) * ((x1 - 0.5) ** 2 + (x2 - 0.25) ** 2 <= 0.0225)
                         + 0.25 * (1 + np.cos(np.pi * G(x1, x2, 0.25, 0.5)))
                         * ((x1 - 0.25) ** 2 + (x2 - 0.5) ** 2 <= 0.0225))
    zero = lambda t, x1, x2: np.zeros_like(x1)
    return Testcase(isAnalytical=False, c0Cont=c0, fCont=zero,
                    u1Cont=lambda t, x1, x2: 0.5 - x2, u2Cont=lambda t, x1, x2: x1 - 0.5,
                    cDCont=zero, gNCont=zero, fZero=True, cDZero=True, isStationary=True)


def transient():
    """Analytical transient case; advection/configureProblem.m:119-132."""
    def cCont(t, x1, x2):
        xp = _xp(x1)
        return xp.cos(7 * x1) * xp.cos(7 * x2) + math.exp(-t)

    def u1(t, x1, x2):
        return _xp(x1).exp(0.5 * (x1 + x2))

    def u2(t, x1, x2):
        return _xp(x1).exp(0.5 * (x1 - x2))

    def f(t, x1, x2):
        xp = _xp(x1)
        return (-math.exp(-t) - 7 * u1(t, x1, x2) * xp.sin(7 * x1) * xp.cos(7 * x2)
                - 7 * u2(t, x1, x2) * xp.cos(7 * x1) * xp.sin(7 * x2)
                + 0.5 * (u1(t, x1, x2) - u2(t, x1, x2)) * cCont(t, x1, x2))
    return Testcase(isAnalytical=True, cCont=cCont, c0Cont=lambda x1, x2: cCont(0, x1, x2),
                    u1Cont=u1, u2Cont=u2, fCont=f, cDCont=cCont, gNCont=_zeros,
                    fZero=False, cDZero=False, isStationary=False, isVelocityStationary=True)


def stream_velocity(seed, kmax, lmax, amplitude, rotation=True, unit_max_speed=False):
    """Divergence-free velocity u = (d psi/dy, -d psi/dx) from the stream
    function psi = psi_0 + sum_kl a_kl sin(k pi x) sin(l pi y) / (pi sqrt(k^2+l^2)),
    psi_0 = -((x-1/2)^2 + (y-1/2)^2)/2 (the rotation u_0 = (1/2-y, x-1/2)).
    a_kl ~ N(0, 1) from numpy's PCG64(seed), rescaled so that the
    perturbation's maximum speed on a 257^2 lattice equals ``amplitude`` (or,
    with ``unit_max_speed``, the whole field's maximum speed is 1)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    ks = [(k, l) for k in range(1, kmax + 1) for l in range(1, lmax + 1)]
    a = rng.standard_normal(len(ks))

    kx = sorted({k for k, _ in ks}); ly = sorted({l for _, l in ks})

    def pert(x1, x2):
        xp = _xp(x1)
        sx = {k: xp.sin(k * np.pi * x1) for k in kx}
        cx = {k: xp.cos(k * np.pi * x1) for k in kx}
        sy = {l: xp.sin(l * np.pi * x2) for l in ly}
        cy = {l: xp.cos(l * np.pi * x2) for l in ly}
        p1 = xp.zeros_like(x1)
        p2 = xp.zeros_like(x1)
        for (k, l), ak in zip(ks, a):
            s = float(ak / (np.pi * np.sqrt(k * k + l * l)))
            p1 = p1 + s * sx[k] * (l * np.pi) * cy[l]       # d psi / dy
            p2 = p2 - s * (k * np.pi) * cx[k] * sy[l]       # -d psi / dx
        return p1, p2

    xs = np.linspace(0, 1, 257)
    X, Y = np.meshgrid(xs, xs)
    q1, q2 = pert(X, Y)
    if unit_max_speed:
        r1 = (0.5 - Y if rotation else 0) + q1
        r2 = (X - 0.5 if rotation else 0) + q2
        scale_p = 1.0
        scale_all = 1.0 / np.max(np.hypot(r1, r2))
    else:
        scale_p = amplitude / np.max(np.hypot(q1, q2))
        scale_all = 1.0

    scale_all = float(scale_all); scale_p = float(scale_p)

    def u1(t, x1, x2):
        return scale_all * ((0.5 - x2 if rotation else 0.0) + scale_p * pert(x1, x2)[0])

    def u2(t, x1, x2):
        return scale_all * ((x1 - 0.5 if rotation else 0.0) + scale_p * pert(x1, x2)[1])
    return u1, u2, a


def gaussian_humps(seed, n=8):
    """c0 = sum_j A_j exp(-|x - x_j|^2 / s_j^2), x_j ~ U[0.2, 0.8]^2,
    s_j ~ U[0.03, 0.1], A_j ~ U[0.5, 1] from PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    xc = rng.uniform(0.2, 0.8, size=(n, 2))
    sg = rng.uniform(0.03, 0.1, size=n)
    A = rng.uniform(0.5, 1.0, size=n)

    def c0(x1, x2):
        xp = _xp(x1)
        r = xp.zeros_like(x1)
        for j in range(n):
            r = r + float(A[j]) * xp.exp(-((x1 - float(xc[j, 0])) ** 2 + (x2 - float(xc[j, 1])) ** 2)
                                         / float(sg[j]) ** 2)
        return r
    return c0


def indicators(seed, ndisc=6, nsq=6):
    """Sum of seeded disc (r ~ U[0.05, 0.15]) and square (side ~ U[0.1, 0.25])
    indicator functions, centres ~ U[0.2, 0.8]^2."""
    rng = np.random.Generator(np.random.PCG64(seed))
    dc = rng.uniform(0.2, 0.8, size=(ndisc, 2)); dr = rng.uniform(0.05, 0.15, size=ndisc)
    sc = rng.uniform(0.2, 0.8, size=(nsq, 2)); ss = rng.uniform(0.1, 0.25, size=nsq)

    def c0(x1, x2):
        r = np.zeros_like(x1, dtype=float)
        for j in range(ndisc):
            r += ((x1 - dc[j, 0]) ** 2 + (x2 - dc[j, 1]) ** 2 <= dr[j] ** 2)
        for j in range(nsq):
            r += (np.abs(x1 - sc[j, 0]) <= ss[j] / 2) & (np.abs(x2 - sc[j, 1]) <= ss[j] / 2)
        return r
    return c0


def rotating_humps(vel_seed=21, hump_seed=22, amplitude=0.05, kmax=4, lmax=4):
    """Rotation + 16-mode perturbation (amplitude 0.05) carrying 8 Gaussian
    humps, zero inflow and source: the p = 2 drop-in and batched workloads."""
    u1, u2, _ = stream_velocity(vel_seed, kmax, lmax, amplitude)
    zero = _zeros
    return Testcase(isAnalytical=False, c0Cont=gaussian_humps(hump_seed), u1Cont=u1, u2Cont=u2,
                    fCont=zero, cDCont=zero, gNCont=zero, fZero=True, cDZero=True, isStationary=True)


def random_discontinuous(vel_seed=31, c_seed=32):
    """32-mode stream-function velocity scaled to unit maximum speed carrying
    discontinuous indicator data: the p = 4 workload."""
    u1, u2, _ = stream_velocity(vel_seed, 8, 4, 1.0, rotation=False, unit_max_speed=True)
    zero = _zeros
    return Testcase(isAnalytical=False, c0Cont=indicators(c_seed), u1Cont=u1, u2Cont=u2,
                    fCont=zero, cDCont=zero, gNCont=zero, fZero=True, cDZero=True, isStationary=True)


TESTCASES = {'solid_body': solid_body, 'transient': transient, 'rotating_humps': rotating_humps,
             'random_discontinuous': random_discontinuous}

"""Gradient of a final-time misfit through the device's discrete adjoint.

The velocity is parametrised by stream-function modes,

    u(a) = u0 + sum_m a_m u_m,   u_m = (d psi_m/dy, -d psi_m/dx),
    psi_m = sin(k pi x) sin(l pi y) / (pi sqrt(k^2 + l^2)),

so u1Disc, u2Disc (DG projections, order 2p) and u.nu (edge points) are
affine in a.  ``festung_adjoint`` returns dJ/dc0 and dJ with respect to those
three fields; the parameter gradient is their contraction with each mode's
fields.  ``AdvectionMisfit`` wraps one evaluation as a torch.autograd.Function;
with CUDA tensors (c0 and c_obs laid out [N][K]) the state, the observation
and the gradients stay on the GPU (``festung_adjoint_dev``), and the adjoint
runs only when a gradient is needed.

The reference has no adjoint; this is an extension of the forward path.
"""
import numpy as np

from .core import Grid, Problem, quadRule1D, quadRule2D, ref_table


def gammaMap(n, S):
    S = np.asarray(S, dtype=float).ravel()
    if n == 1:
        return 1 - S, S
    if n == 2:
        return np.zeros_like(S), 1 - S
    return S, np.zeros_like(S)


def stream_modes(kmax=8, lmax=4):
    """(k, l) index pairs of the modes, k fastest within l."""
    return [(k, l) for l in range(1, lmax + 1) for k in range(1, kmax + 1)]


def mode_velocity(k, l, x1, x2):
    s = 1.0 / (np.pi * np.sqrt(k * k + l * l))
    u1 = s * np.sin(k * np.pi * x1) * (l * np.pi) * np.cos(l * np.pi * x2)
    u2 = -s * (k * np.pi) * np.cos(k * np.pi * x1) * np.sin(l * np.pi * x2)
    return u1, u2


def rotation(x1, x2):
    return 0.5 - x2, x1 - 0.5


class StreamModel:
    """Device problem (limiter off) whose velocity is u0 + sum a_m u_m."""

    def __init__(self, hmax, p, ordRK=3, modes=None, c0=None, adj_serial=False, adj_kernel=None,
                 adj_binomial=False, graph=False):
        self.g = Grid.domainSquare(hmax)
        self.p = p
        self.prob = Problem(self.g, p, ordRK, adj_serial=adj_serial, adj_kernel=adj_kernel,
                            adj_binomial=adj_binomial, graph=graph)
        self.modes = modes if modes is not None else stream_modes()
        pr = self.prob
        Q1, Q2, _ = quadRule2D(pr.ordE)
        self.XE = self.g.mapRef2Phy(Q1, Q2)
        Q1, Q2, _ = quadRule2D(pr.ordF)
        self.XF = self.g.mapRef2Phy(Q1, Q2)
        Q, _ = quadRule1D(2 * p + 1)
        K = self.g.numT
        X1 = np.empty((K, 3, len(Q)), order='F'); X2 = np.empty_like(X1)
        for n in range(3):
            A1, A2 = gammaMap(n + 1, Q)
            X1[:, n, :], X2[:, n, :] = self.g.mapRef2Phy(A1, A2)
        self.XEdge = (X1, X2)
        nu = self.g.nuE0T                     # K x 3 x 2
        self.n1 = nu[:, :, 0][:, :, None]; self.n2 = nu[:, :, 1][:, :, None]
        # point values of the base field and of every mode
        self.base = self._points(rotation)
        self.mode_pts = [self._points(lambda a, b, k=k, l=l: mode_velocity(k, l, a, b)) for (k, l) in self.modes]
        # each mode's DG projections and normal velocity (the derivatives of the fields in a_m)
        self.mode_fields = []
        for (u1q, u2q, u1e, u2e) in self.mode_pts:
            d1 = pr.project(pr.ordE, u1q)
            d2 = pr.project(pr.ordE, u2q)
            vn = self.n1 * u1e + self.n2 * u2e
            self.mode_fields.append((d1, d2, vn))
        if c0 is not None:
            self.set_c0(c0)

    def _points(self, f):
        u1q, u2q = f(*self.XE)
        u1e, u2e = f(*self.XEdge)
        return [np.asfortranarray(x) for x in (u1q, u2q, u1e, u2e)]

    def set_c0(self, c0fun):
        self.c0 = self.prob.project(self.prob.ordF, c0fun(*self.XF))

    def set_velocity(self, a):
        a = np.asarray(a, dtype=float)
        pts = [b.copy() for b in self.base]
        for am, mp in zip(a, self.mode_pts):
            for i in range(4):
                pts[i] += am * mp[i]
        self.prob.set_velocity_points(*pts)

    def run(self, c0, a, nsteps, tau):
        """Final state after nsteps (forward only)."""
        self.set_velocity(a)
        self.prob.set_state(c0)
        self.prob.step(nsteps, tau)
        return self.prob.get_state()

    def misfit(self, c0, a, nsteps, tau, cObs):
        """J = (c(T) - cObs)^T M (c(T) - cObs) by the forward path."""
        cT = self.run(c0, a, nsteps, tau)
        d = cT - cObs
        N = self.prob.N
        hatM = ref_table(self.p, 'hatM').reshape(N, N, order='F')
        B = self.g.B
        det = B[:, 0, 0] * B[:, 1, 1] - B[:, 0, 1] * B[:, 1, 0]
        return float(np.sum(det * np.einsum('ki,ij,kj->k', d, hatM, d)))

    # ---- device-resident variants (torch CUDA tensors, [cols][K] layouts) ----
    def _dev(self):
        """Point values and mode fields as CUDA tensors (built once)."""
        if getattr(self, '_devf', None) is None:
            import torch

            def q(x):      # K x R (Fortran) -> [R][K]
                return torch.from_numpy(np.ascontiguousarray(np.asarray(x).T)).cuda()

            def e(x):      # K x 3 x Re (Fortran) -> [Re][3][K]
                return torch.from_numpy(np.ascontiguousarray(np.asarray(x).transpose(2, 1, 0))).cuda()
            conv = (q, q, e, e)
            base = [conv[i](self.base[i]) for i in range(4)]
            modes = [torch.stack([conv[i](mp[i]) for mp in self.mode_pts]) for i in range(4)]
            d1 = torch.stack([q(f[0]) for f in self.mode_fields])
            d2 = torch.stack([q(f[1]) for f in self.mode_fields])
            vn = torch.stack([e(f[2]) for f in self.mode_fields])
            N = self.prob.N
            hatM = torch.from_numpy(ref_table(self.p, 'hatM').reshape(N, N, order='F').copy()).cuda()
            B = self.g.B
            det = torch.from_numpy(np.ascontiguousarray(B[:, 0, 0] * B[:, 1, 1] - B[:, 0, 1] * B[:, 1, 0])).cuda()
            self._devf = dict(base=base, modes=modes, d1=d1, d2=d2, vn=vn, hatM=hatM, det=det)
        return self._devf

    def set_velocity_dev(self, a):
        """set_velocity with the point values formed on the device from a
        (a CUDA or CPU tensor of mode coefficients), in the same order of
        operations as the host (the same bits)."""
        import torch
        D = self._dev()
        a = torch.as_tensor(a, dtype=torch.float64).to(D['base'][0].device)
        pts = [b.clone() for b in D['base']]
        for m in range(a.shape[0]):
            for i in range(4):
                pts[i] += a[m] * D['modes'][i][m]
        self.prob.set_velocity_points_dev(*pts)

    def misfit_dev(self, c0, a, nsteps, tau, cObs):
        """J by the forward path with c0, cObs CUDA tensors [N][K]; a 0-d tensor."""
        import torch
        D = self._dev()
        self.set_velocity_dev(a)
        self.prob.set_state_dev(c0)
        self.prob.step(nsteps, tau)
        cT = torch.empty_like(c0)
        self.prob.get_state_dev(cT)
        d = cT - cObs
        return torch.sum(D['det'] * torch.einsum('ik,ij,jk->k', d, D['hatM'], d))

    def gradient_dev(self, c0, a, nsteps, tau, cObs, ckpt=0):
        """J (float), dJ/dc0 ([N][K] CUDA tensor), dJ/da (CUDA tensor) and the
        device times, without host copies of the fields."""
        import torch
        D = self._dev()
        self.set_velocity_dev(a)
        self.prob.set_state_dev(c0)
        g0 = torch.empty_like(c0)
        g1 = torch.empty_like(c0)
        g2 = torch.empty_like(c0)
        gv = torch.empty((self.prob.Re, 3, self.prob.K), dtype=torch.float64, device=c0.device)
        r = self.prob.adjoint_dev(nsteps, tau, cObs, ckpt, g0, g1, g2, gv)
        ga = (torch.einsum('mik,ik->m', D['d1'], g1) + torch.einsum('mik,ik->m', D['d2'], g2)
              + torch.einsum('mrnk,rnk->m', D['vn'], gv))
        return r['J'], g0, ga, r

    def gradient(self, c0, a, nsteps, tau, cObs, ckpt=0):
        """J, dJ/dc0 (K x N), dJ/da (len(modes)) and the device times."""
        self.set_velocity(a)
        self.prob.set_state(c0)
        r = self.prob.adjoint(nsteps, tau, cObs, ckpt)
        ga = np.array([np.sum(r['gradU1'] * d1) + np.sum(r['gradU2'] * d2) + np.sum(r['gradVn'] * vn)
                       for (d1, d2, vn) in self.mode_fields])
        return r['J'], r['gradC0'], ga, r


def misfit_function():
    """torch.autograd.Function computing J(c0, a) with the device adjoint
    as its backward.  With CUDA tensors c0, cObs ([N][K]) everything stays on
    the device; with CPU tensors (c0 K x N) the host entry points are used.
    The adjoint (which needs the forward pass it checkpoints) runs in forward
    only when a gradient is requested; otherwise forward is the plain forward
    run."""
    import torch

    class AdvectionMisfit(torch.autograd.Function):
        @staticmethod
        def forward(ctx, c0, a, model, nsteps, tau, cObs, ckpt):
            need = ctx.needs_input_grad[0] or ctx.needs_input_grad[1]
            if c0.is_cuda:
                if need:
                    J, g0, ga, _ = model.gradient_dev(c0.detach(), a.detach(), nsteps, tau, cObs, ckpt)
                    ga = ga.to(a.device)
                else:
                    return model.misfit_dev(c0.detach(), a.detach(), nsteps, tau, cObs).detach()
            else:
                co = cObs.detach().cpu().numpy() if torch.is_tensor(cObs) else cObs
                if not need:
                    return torch.tensor(model.misfit(c0.detach().numpy(), a.detach().cpu().numpy(), nsteps, tau, co),
                                        dtype=torch.float64)
                J, g0, ga, _ = model.gradient(c0.detach().numpy(), a.detach().cpu().numpy(), nsteps, tau, co, ckpt)
                g0 = torch.from_numpy(np.ascontiguousarray(g0)).to(c0)
                ga = torch.from_numpy(ga).to(a)
            ctx.save_for_backward(g0, ga)
            return torch.tensor(J, dtype=torch.float64, device=c0.device)

        @staticmethod
        def backward(ctx, gJ):
            g0, ga = ctx.saved_tensors
            return gJ * g0, (gJ.to(ga.device) * ga), None, None, None, None, None

    return AdvectionMisfit

"""Python handles over the C-ABI: ``Grid`` (host-side mesh, the struct ``g``
of core/generateGridData.m) and ``Problem`` (device-resident advection
state and operators).

Arrays use the reference's Matlab shapes (e.g. cDisc is K x N) and are
exchanged in column-major (Fortran) order, which is the memory layout the
C-ABI expects, so no transposes are made.  Ids are 0-based.
"""
import ctypes as C
import os

import numpy as np

from . import _lib as L
from ._lib import check, lib, ptr

# shapes of the grid fields as functions of (numT, numV, numE)
_FIELD_SHAPES = {
    'V0T': lambda T, V, E: (T, 3), 'V0E': lambda T, V, E: (E, 2), 'T0E': lambda T, V, E: (E, 2),
    'E0E': lambda T, V, E: (E, 2), 'E0T': lambda T, V, E: (T, 3), 'idE': lambda T, V, E: (E,),
    'idE0T': lambda T, V, E: (T, 3), 'nbr': lambda T, V, E: (T, 3), 'nbrEdge': lambda T, V, E: (T, 3),
    'v2tPtr': lambda T, V, E: (V + 1,), 'v2tIdx': lambda T, V, E: (3 * T,),
    'markE0Tint': lambda T, V, E: (T, 3), 'markE0TbdrD': lambda T, V, E: (T, 3),
    'markE0TbdrN': lambda T, V, E: (T, 3), 'markV0TbdrD': lambda T, V, E: (T, 3),
    'coordV': lambda T, V, E: (V, 2), 'areaE': lambda T, V, E: (E,), 'nuE': lambda T, V, E: (E, 2),
    'baryE': lambda T, V, E: (E, 2), 'areaT': lambda T, V, E: (T,), 'baryT': lambda T, V, E: (T, 2),
    'B': lambda T, V, E: (T, 2, 2), 'areaE0T': lambda T, V, E: (T, 3), 'sigE0T': lambda T, V, E: (T, 3),
    'nuE0T': lambda T, V, E: (T, 3, 2), 'coordV0T': lambda T, V, E: (T, 3, 2),
}
_BOOL_FIELDS = {'markE0Tint', 'markE0TbdrD', 'markE0TbdrN', 'markV0TbdrD'}


def fortran(a, shape, dtype=np.float64):
    a = np.asfortranarray(a, dtype=dtype)
    if a.shape != tuple(shape):
        raise ValueError('expected shape %s, got %s' % (tuple(shape), a.shape))
    return a


class Grid:
    """Mesh with the reference's numbering (core/generateGridData.m:130-198).
    Fields are read lazily as attributes: ``g.E0T``, ``g.nuE0T``, ..."""

    def __init__(self, handle):
        self._h = C.c_void_p(handle)
        T, V, E = C.c_int32(), C.c_int32(), C.c_int32()
        check(lib().festung_grid_sizes(self._h, C.byref(T), C.byref(V), C.byref(E)))
        self.numT, self.numV, self.numE = T.value, V.value, E.value
        self._cache = {}

    @classmethod
    def domainSquare(cls, hmax, a=0.0, b=1.0):
        """core/domainSquare.m:67-105 (Friedrichs-Keller triangulation)."""
        h = C.c_void_p()
        check(lib().festung_domain_square(float(hmax), float(a), float(b), C.byref(h)))
        return cls(h.value)

    @classmethod
    def generateGridData(cls, coordV, V0T):
        """core/generateGridData.m:130-198 for a CCW mesh (V0T 0-based)."""
        coordV = fortran(coordV, (np.shape(coordV)[0], 2))
        V0T = fortran(V0T, (np.shape(V0T)[0], 3), np.int32)
        h = C.c_void_p()
        check(lib().festung_grid_create(ptr(coordV), coordV.shape[0], ptr(V0T), V0T.shape[0], C.byref(h)))
        return cls(h.value)

    @classmethod
    def concat(cls, grids):
        """Disjoint union of grids (ids offset in order)."""
        arr = (C.c_void_p * len(grids))(*[g._h.value for g in grids])
        h = C.c_void_p()
        check(lib().festung_grid_concat(arr, len(grids), C.byref(h)))
        return cls(h.value)

    def set_boundary(self, idE, markE0TbdrN=None):
        idE = fortran(idE, (self.numE,), np.int32)
        mN = None if markE0TbdrN is None else fortran(markE0TbdrN, (self.numT, 3), np.uint8)
        check(lib().festung_grid_set_boundary(self._h, ptr(idE), ptr(mN)))
        self._cache.clear()

    def field(self, name):
        if name in self._cache:
            return self._cache[name]
        cnt = C.c_int64()
        isd = C.c_int()
        check(lib().festung_grid_field_size(self._h, name.encode(), C.byref(cnt), C.byref(isd)))
        shape = _FIELD_SHAPES[name](self.numT, self.numV, self.numE)
        if isd.value:
            out = np.empty(shape, dtype=np.float64, order='F')
            check(lib().festung_grid_get_f64(self._h, name.encode(), ptr(out)))
        else:
            out = np.empty(shape, dtype=np.int32, order='F')
            check(lib().festung_grid_get_i32(self._h, name.encode(), ptr(out)))
            if name in _BOOL_FIELDS:
                out = out.astype(bool)
        assert out.size == cnt.value
        self._cache[name] = out
        return out

    def __getattr__(self, name):
        if name in _FIELD_SHAPES:
            return self.field(name)
        raise AttributeError(name)

    def mapRef2Phy(self, Q1, Q2):
        """Physical coordinates of reference points (g.mapRef2Phy):
        returns X1, X2 of shape K x len(Q1)."""
        Q1 = np.ascontiguousarray(np.atleast_1d(Q1), dtype=np.float64)
        Q2 = np.ascontiguousarray(np.atleast_1d(Q2), dtype=np.float64)
        nq = Q1.size
        X1 = np.empty((self.numT, nq), order='F')
        X2 = np.empty((self.numT, nq), order='F')
        check(lib().festung_grid_map_ref_to_phy(self._h, nq, ptr(Q1), ptr(Q2), ptr(X1), ptr(X2)))
        return X1, X2

    def __del__(self):
        h = getattr(self, '_h', None)
        if h is not None and h.value and L._lib is not None:
            L._lib.festung_grid_destroy(h)
            self._h = None


def ref_table(p, name, ord_=0):
    """Reference tables computed by the library (quadrature rules, basis
    tables, reference blocks), column-major with the reference's shapes."""
    cnt = C.c_int64()
    check(lib().festung_ref_size(int(p), name.encode(), int(ord_), C.byref(cnt)))
    out = np.empty(cnt.value, dtype=np.float64)
    check(lib().festung_ref_get(int(p), name.encode(), int(ord_), ptr(out)))
    return out


def quadRule1D(qOrd):
    v = ref_table(0, 'quad1D', qOrd)
    R = v.size // 2
    return v[:R], v[R:]


def quadRule2D(qOrd):
    v = ref_table(0, 'quad2D', qOrd)
    R = v.size // 3
    return v[:R], v[R:2 * R], v[2 * R:]


class Problem:
    """Device-resident advection problem (festung_problem_* of the C-ABI)."""

    def __init__(self, grid, p, ordRK, limiter=False, graph=False, adj_serial=False, adj_kernel=None,
                 limiter_type='linear', adj_binomial=False):
        """limiter_type: 'linear', 'hierarch_vert' or 'strict'
        (core/applySlopeLimiterTaylor.m:69-85); adj_serial: recompute
        checkpointed segments on the problem's own stream; adj_kernel: None (by
        mesh size), 'fused', 'split' or 'deferred' -- where the adjoint forms
        the velocity sensitivities (in the transposed stage kernel's thread, on
        two more threads per element, or once per checkpoint segment after its
        sweep); adj_binomial: binomial checkpointing in adjoint() (its ckpt
        argument is then the number of state snapshots)."""
        if adj_kernel not in (None, 'fused', 'split', 'deferred'):
            raise ValueError('adj_kernel must be None, "fused", "split" or "deferred"')
        if adj_binomial and adj_kernel == 'deferred':
            raise ValueError('binomial checkpointing forms the sensitivities fused or split')
        if limiter_type not in L.LIMITER_TYPES:
            raise ValueError('Unknown limiter type %r' % (limiter_type,))
        flags = ((L.FLAG_LIMITER if limiter else 0) | (L.FLAG_GRAPH if graph else 0)
                 | (L.FLAG_ADJ_SERIAL if adj_serial else 0)
                 | {None: 0, 'fused': L.FLAG_ADJ_FUSED, 'split': L.FLAG_ADJ_SPLIT,
                    'deferred': L.FLAG_ADJ_DEFER}[adj_kernel]
                 | L.LIMITER_TYPES[limiter_type] | (L.FLAG_ADJ_BINOMIAL if adj_binomial else 0))
        h = C.c_void_p()
        check(lib().festung_problem_create(grid._h, int(p), int(ordRK), flags, C.byref(h)))
        self._h = h
        self.grid = grid
        self.p, self.ordRK, self.limiter = p, ordRK, limiter
        K, N, RE, RF, Re = (C.c_int32() for _ in range(5))
        check(lib().festung_problem_sizes(h, C.byref(K), C.byref(N), C.byref(RE), C.byref(RF), C.byref(Re)))
        self.K, self.N, self.RE, self.RF, self.Re = K.value, N.value, RE.value, RF.value, Re.value
        self.ordE = max(2 * p, 1)
        self.ordF = 2 * p + 1

    def close(self):
        if self._h is not None and self._h.value:
            lib().festung_problem_destroy(self._h)
        self._h = None

    def __del__(self):
        if L._lib is not None and getattr(self, '_h', None) is not None:
            self.close()

    @property
    def stream(self):
        s = C.c_void_p()
        check(lib().festung_problem_stream(self._h, C.byref(s)))
        return s.value

    def set_velocity_disc(self, u1Disc, u2Disc, vNormalOnQuadEdge):
        """u1Disc, u2Disc (K x N), vNormalOnQuadEdge (K x 3 x Re); None keeps
        that field."""
        K, N = self.K, self.N
        a = None if u1Disc is None else fortran(u1Disc, (K, N))
        b = None if u2Disc is None else fortran(u2Disc, (K, N))
        v = None if vNormalOnQuadEdge is None else fortran(vNormalOnQuadEdge, (K, 3, self.Re))
        check(lib().festung_set_velocity_disc(self._h, ptr(a), ptr(b), ptr(v)))

    def set_velocity_points(self, u1q, u2q, u1e, u2e):
        K = self.K
        a = fortran(u1q, (K, self.RE)); b = fortran(u2q, (K, self.RE))
        c = fortran(u1e, (K, 3, self.Re)); d = fortran(u2e, (K, 3, self.Re))
        check(lib().festung_set_velocity_points(self._h, ptr(a), ptr(b), ptr(c), ptr(d)))

    def set_source_points(self, fq=None, cDe=None, gNe=None):
        """globL - globKD - globKN from f at the element points (K x RE), c_D
        and g_N at the edge points (K x 3 x Re); None = zero data."""
        f = None if fq is None else fortran(fq, (self.K, self.RE))
        c = None if cDe is None else fortran(cDe, (self.K, 3, self.Re))
        n = None if gNe is None else fortran(gNe, (self.K, 3, self.Re))
        check(lib().festung_set_source_points(self._h, ptr(f), ptr(c), ptr(n)))

    def _dev_call(self, fn, tensors, shapes, pre=(), post=()):
        """Call a _dev entry point fn(h, *pre, pointers..., *post) on float64
        CUDA tensors in [cols][K] layout (None passes NULL).  Every tensor is
        checked before anything is enqueued; the problem's stream waits for
        torch's current stream before the call, and torch's current stream
        waits for the problem's after it, so torch neither frees nor reads
        the tensors before the library is done with them (no allocator stream
        bookkeeping that could outlive the problem's stream)."""
        import torch
        for t, shape in zip(tensors, shapes):
            if t is None:
                continue
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == torch.float64 and t.is_contiguous()
                    and t.numel() == int(np.prod(shape))):
                raise ValueError('expected a contiguous float64 CUDA tensor with %d values' % int(np.prod(shape)))
        ours = torch.cuda.ExternalStream(self.stream)
        cur = torch.cuda.current_stream()
        ours.wait_stream(cur)
        ptrs = [None if t is None else C.c_void_p(t.data_ptr()) for t in tensors]
        try:
            check(fn(self._h, *pre, *ptrs, *post))
        finally:
            cur.wait_stream(ours)

    def set_velocity_points_dev(self, u1q, u2q, u1e, u2e):
        """set_velocity_points from CUDA tensors laid out [RE][K] and [Re][3][K]
        (reference element order)."""
        K = self.K
        self._dev_call(lib().festung_set_velocity_points_dev, [u1q, u2q, u1e, u2e],
                       [(self.RE, K)] * 2 + [(self.Re, 3, K)] * 2)

    def set_source_points_dev(self, fq=None, cDe=None, gNe=None):
        self._dev_call(lib().festung_set_source_points_dev, [fq, cDe, gNe],
                       [(self.RE, self.K), (self.Re, 3, self.K), (self.Re, 3, self.K)])

    def set_vertex_dirichlet_dev(self, vertexD):
        self._dev_call(lib().festung_set_vertex_dirichlet_dev, [vertexD], [(self.grid.numV,)])

    def set_tau_field(self, tauK):
        t = None if tauK is None else fortran(tauK, (self.K,))
        check(lib().festung_set_tau_field(self._h, ptr(t)))

    @property
    def elem_block(self):
        """Elements per workgroup of the element kernels (64, 128 or 256)."""
        v = C.c_int32()
        check(lib().festung_get_elem_block(self._h, C.byref(v)))
        return v.value

    @elem_block.setter
    def elem_block(self, bs):
        check(lib().festung_set_elem_block(self._h, int(bs)))

    def autotune(self, tau, nsteps=10):
        """Time nsteps steps with 64, 128 and 256 elements per workgroup,
        keep the fastest; the state is left as it was.  The choice is
        recorded in the process's tuning table (problems created later with
        the same device, p, ordRK, limiter and ceil(log2 K) start with it;
        tune_save / tune_load or FESTUNG_TUNE_FILE carry it to other
        processes).  Returns (best, {elements per workgroup: ms per step})."""
        best = C.c_int32()
        ms = np.zeros(3)
        check(lib().festung_autotune(self._h, float(tau), int(nsteps), C.byref(best), ptr(ms)))
        return best.value, dict(zip((64, 128, 256), ms.tolist()))

    def set_vertex_dirichlet(self, vertexD):
        v = None if vertexD is None else fortran(vertexD, (self.grid.numV,))
        check(lib().festung_set_vertex_dirichlet(self._h, ptr(v)))

    def project(self, ord_, fq, set_state=False):
        R = self.RE if ord_ == self.ordE else self.RF
        f = fortran(fq, (self.K, R))
        out = np.empty((self.K, self.N), order='F')
        check(lib().festung_project(self._h, int(ord_), ptr(f), ptr(out), int(bool(set_state))))
        return out

    def set_state(self, cDisc):
        c = fortran(cDisc, (self.K, self.N))
        check(lib().festung_set_state(self._h, ptr(c)))

    def get_state(self):
        out = np.empty((self.K, self.N), order='F')
        check(lib().festung_get_state(self._h, ptr(out)))
        return out

    def set_state_dev(self, t):
        """State from a contiguous float64 CUDA tensor laid out [N][K]
        (reference element order); ordered after torch's pending work on it."""
        self._dev_call(lib().festung_set_state_dev, [t], [(self.N, self.K)])

    def get_state_dev(self, t):
        """Copy the state into a contiguous float64 CUDA tensor laid out
        [N][K]; torch's current stream waits for the copy."""
        self._dev_call(lib().festung_get_state_dev, [t], [(self.N, self.K)])

    def begin_step(self):
        check(lib().festung_begin_step(self._h))

    def stage(self, nSubStep, tau):
        check(lib().festung_stage(self._h, int(nSubStep), float(tau)))

    def step(self, nsteps, tau):
        check(lib().festung_step(self._h, int(nsteps), float(tau)))

    def limit(self):
        check(lib().festung_limit(self._h))

    def limit_disc_linear(self):
        """Limit the current state with the DG-basis linear limiter
        (core/applySlopeLimiterDiscLinear.m:109-123); limiter_info() then
        reports its alpha and bounds."""
        check(lib().festung_limit_disc_linear(self._h))

    def rhs(self):
        out = np.empty((self.K, self.N), order='F')
        check(lib().festung_rhs(self._h, ptr(out)))
        return out

    def limiter_info(self):
        K = self.K
        a = np.empty(K); mn = np.empty((K, 3), order='F'); mx = np.empty((K, 3), order='F')
        check(lib().festung_limiter_info(self._h, ptr(a), ptr(mn), ptr(mx)))
        return a, mn, mx

    def l2_error(self, ord_, cq):
        R = self.RE if ord_ == self.ordE else self.RF
        q = fortran(cq, (self.K, R))
        e = C.c_double()
        check(lib().festung_l2_error(self._h, int(ord_), ptr(q), C.byref(e)))
        return e.value

    def adjoint(self, nsteps, tau, cObs, ckpt=0):
        """Forward nsteps from the current state, loss and reverse sweep;
        ckpt: steps per checkpoint segment (0: all stored), or the number of
        state snapshots with binomial checkpointing.  Returns dict(J, gradC0
        (K x N), gradU1, gradU2 (K x N), gradVn (K x 3 x Re), forward_ms,
        backward_ms)."""
        K, N, Re = self.K, self.N, self.Re
        o = fortran(cObs, (K, N))
        J = np.zeros(1)
        g0 = np.empty((K, N), order='F'); g1 = np.empty((K, N), order='F'); g2 = np.empty((K, N), order='F')
        gv = np.empty((K, 3, Re), order='F')
        tm = np.zeros(2)
        check(lib().festung_adjoint(self._h, int(nsteps), float(tau), ptr(o), int(ckpt), ptr(J), ptr(g0), ptr(g1),
                                    ptr(g2), ptr(gv), ptr(tm)))
        return dict(J=float(J[0]), gradC0=g0, gradU1=g1, gradU2=g2, gradVn=gv, forward_ms=tm[0], backward_ms=tm[1])

    def adj_rhs(self, mu):
        m = fortran(mu, (self.K, self.N))
        out = np.empty((self.K, self.N), order='F')
        check(lib().festung_adj_rhs(self._h, ptr(m), ptr(out)))
        return out

    def adjoint_dev(self, nsteps, tau, cObs, ckpt=0, gradC0=None, gradU1=None, gradU2=None, gradVn=None):
        """adjoint() with cObs and the gradients as float64 CUDA tensors
        ([N][K], [N][K], [N][K], [Re][3][K]; any output may be None).
        Returns dict(J, forward_ms, backward_ms); torch's current stream
        waits for the outputs."""
        K, N, Re = self.K, self.N, self.Re
        J = np.zeros(1)
        tm = np.zeros(2)
        L_ = lib()

        def call(h, cO, g0, g1, g2, gv):
            return L_.festung_adjoint_dev(h, int(nsteps), float(tau), cO, int(ckpt), ptr(J), g0, g1, g2, gv, ptr(tm))
        self._dev_call(call, [cObs, gradC0, gradU1, gradU2, gradVn], [(N, K)] * 4 + [(Re, 3, K)])
        return dict(J=float(J[0]), forward_ms=tm[0], backward_ms=tm[1])

    OPERATORS = {'G': L.OP_G, 'G1': L.OP_G1, 'G2': L.OP_G2, 'R': L.OP_R, 'M': L.OP_M, 'Minv': L.OP_MINV}

    def apply_operator(self, op, x):
        """y = A x (K x N) for A = 'G1', 'G2', 'G' (globG{1}, globG{2}, their
        sum), 'R' (globR), 'M' (globM) or 'Minv' (globM \\ x), at the current
        velocity."""
        xx = fortran(x, (self.K, self.N))
        y = np.empty((self.K, self.N), order='F')
        check(lib().festung_apply_operator(self._h, self.OPERATORS[op], ptr(xx), ptr(y)))
        return y

    def apply_operator_dev(self, op, x, y):
        """apply_operator on float64 CUDA tensors laid out [N][K]."""
        self._dev_call(lib().festung_apply_operator_dev, [x, y], [(self.N, self.K)] * 2, pre=(self.OPERATORS[op],))

    def apply_operator_dev_noalloc(self, op, x, y):
        """apply_operator_dev through the entry point that allocates nothing
        (library-owned scratch; K*N <= 2^22)."""
        self._dev_call(lib().festung_apply_operator_dev_noalloc, [x, y], [(self.N, self.K)] * 2,
                       pre=(self.OPERATORS[op],))

    def assemble_vectors(self, fq=None, cDe=None, gNe=None):
        """(globL, globKD, globKN) as K x N arrays from point values (None
        where the input is None)."""
        K, N = self.K, self.N
        f = None if fq is None else fortran(fq, (K, self.RE))
        c = None if cDe is None else fortran(cDe, (K, 3, self.Re))
        n = None if gNe is None else fortran(gNe, (K, 3, self.Re))
        outs = [None if v is None else np.empty((K, N), order='F') for v in (f, c, n)]
        check(lib().festung_assemble_vectors(self._h, ptr(f), ptr(c), ptr(n), *[ptr(o) for o in outs]))
        return tuple(outs)

    def vertex_based_limiter(self, valCentroid, valV0T, markV0TbdrD=None, dataV0T=None):
        """core/computeVertexBasedLimiter.m: (alphaE (K), [minV0T, maxV0T] (K x 3))."""
        K = self.K
        vc = fortran(valCentroid, (K,))
        vv = fortran(valV0T, (K, 3))
        mk = None if markV0TbdrD is None else fortran(markV0TbdrD, (K, 3), np.uint8)
        dv = None if dataV0T is None else fortran(dataV0T, (K, 3))
        a = np.empty(K); mn = np.empty((K, 3), order='F'); mx = np.empty((K, 3), order='F')
        check(lib().festung_vertex_based_limiter(self._h, ptr(vc), ptr(vv), ptr(mk), ptr(dv), ptr(a), ptr(mn), ptr(mx)))
        return a, [mn, mx]

    def apply_slope_limiter_disc(self, dataDisc, markV0TbdrD=None, dataV0T=None):
        """core/applySlopeLimiterDisc.m: (dataDisc limited, [minV0T, maxV0T])."""
        K, N = self.K, self.N
        c = fortran(dataDisc, (K, N))
        mk = None if markV0TbdrD is None else fortran(markV0TbdrD, (K, 3), np.uint8)
        dv = None if dataV0T is None else fortran(dataV0T, (K, 3))
        out = np.empty((K, N), order='F'); mn = np.empty((K, 3), order='F'); mx = np.empty((K, 3), order='F')
        check(lib().festung_apply_slope_limiter_disc(self._h, ptr(c), ptr(mk), ptr(dv), ptr(out), ptr(mn), ptr(mx)))
        return out, [mn, mx]

    KERNEL_KINDS = ('stage', 'vertex', 'limit_dot', 'limit', 'copy')

    def profile(self, enable):
        check(lib().festung_profile(self._h, int(bool(enable))))

    def profile_read(self):
        ms = np.zeros(5)
        cnt = np.zeros(5, dtype=np.int64)
        check(lib().festung_profile_read(self._h, ptr(ms), cnt.ctypes.data_as(C.POINTER(C.c_int64))))
        return {k: (ms[i], int(cnt[i])) for i, k in enumerate(self.KERNEL_KINDS)}

    def sync(self):
        check(lib().festung_sync(self._h))


class Batch:
    """Several problems advanced together as one hipGraph launch
    (festung_batch_* of the C-ABI)."""

    def __init__(self, problems):
        self.problems = list(problems)
        arr = (C.c_void_p * len(self.problems))(*[p._h.value for p in self.problems])
        h = C.c_void_p()
        check(lib().festung_batch_create(arr, len(self.problems), C.byref(h)))
        self._h = h

    def step(self, nsteps, taus):
        t = np.ascontiguousarray(taus, dtype=np.float64)
        assert t.size == len(self.problems)
        check(lib().festung_batch_step(self._h, int(nsteps), ptr(t)))

    def sync(self):
        check(lib().festung_batch_sync(self._h))

    def close(self):
        if self._h is not None and self._h.value:
            lib().festung_batch_destroy(self._h)
        self._h = None

    def __del__(self):
        if L._lib is not None and getattr(self, '_h', None) is not None:
            self.close()


def rungeKuttaExplicit(ord_, tau, t0):
    """core/rungeKuttaExplicit.m:64-78 through festung_runge_kutta_explicit:
    (stage times, omega)."""
    t = np.zeros(4); w = np.zeros(4); n = C.c_int32()
    check(lib().festung_runge_kutta_explicit(int(ord_), float(tau), float(t0), ptr(t), ptr(w), C.byref(n)))
    return t[:n.value].copy(), w[:n.value].copy()


def tune_save(path):
    """Write the autotuner's table (festung_tune_save) to a text file."""
    check(lib().festung_tune_save(os.fsencode(path)))


def tune_load(path):
    """Merge a tuning file written by tune_save (festung_tune_load)."""
    check(lib().festung_tune_load(os.fsencode(path)))


def tune_clear():
    check(lib().festung_tune_clear())


def tune_count():
    n = C.c_int32()
    check(lib().festung_tune_count(C.byref(n)))
    return n.value


def binomial_schedule(nsteps, snapshots):
    """(forward steps besides the per-step recomputations, actions) of the
    binomial checkpointing schedule for nsteps steps and the given number of
    state snapshots (festung_binomial_schedule)."""
    adv = C.c_int64(); act = C.c_int64()
    check(lib().festung_binomial_schedule(int(nsteps), int(snapshots), C.byref(adv), C.byref(act)))
    return adv.value, act.value

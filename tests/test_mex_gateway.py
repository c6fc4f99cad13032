"""The MEX gateway (festung_amd/mex/festung_mex.cpp) compiled against the test
double of the MEX API (tests/mex/mex.h, tests/mex/mex_standin.cpp) and driven
from Python: argument checking, 1-based id conversion, shapes, the
element-major KN x 1 vectors of advection/solveStep.m:68, handle bookkeeping
(mexLock while handles live), and errors raised only through
mexErrMsgIdAndTxt.  Host-only commands run on the CPU; the device commands
are compared bit for bit with the Python binding on a GPU."""
import ctypes as C
import os
import subprocess

import numpy as np
import pytest

from festung_amd import _lib
from festung_amd.core import Grid, quadRule2D, ref_table, rungeKuttaExplicit

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class Gateway:
    """festung_mex(cmd, ...) through the test double."""

    def __init__(self, path):
        L = C.CDLL(path)
        vp, dp, sz = C.c_void_p, C.POINTER(C.c_double), C.c_size_t
        for name, res, args in (('standin_double', vp, [sz, sz, dp]), ('standin_logical', vp, [sz, sz, C.c_char_p]),
                                ('standin_string', vp, [C.c_char_p]), ('standin_class', C.c_int, [vp]),
                                ('standin_rows', sz, [vp]), ('standin_cols', sz, [vp]),
                                ('standin_read', None, [vp, dp]), ('standin_read_string', None, [vp, C.c_char_p, sz]),
                                ('standin_free', None, [vp]), ('standin_lock_count', C.c_int, []),
                                ('standin_at_exit', None, []),
                                ('standin_call', C.c_int, [C.c_int, C.POINTER(vp), C.c_int, C.POINTER(vp),
                                                           C.c_char_p, C.c_char_p, sz])):
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        self.L = L
        self.handles = {}

    def _arg(self, x):
        L = self.L
        if isinstance(x, str):
            return L.standin_string(x.encode())
        if isinstance(x, tuple) and x and x[0] == 'handle':
            return self.handles[x[1]]
        a = np.asarray(x)
        if a.dtype == bool:
            m, n = (a.shape + (1,))[:2] if a.ndim else (1, 1)
            b = np.asfortranarray(a.reshape(m, n) if a.ndim <= 2 else a.reshape(a.shape[0], -1, order='F'))
            return L.standin_logical(m, n, b.astype(np.uint8).tobytes(order='F'))
        a = np.asarray(a, dtype=np.float64)
        if a.ndim == 0:
            a = a.reshape(1, 1)
        elif a.ndim == 1:
            a = a.reshape(-1, 1)
        m = a.shape[0]
        n = int(np.prod(a.shape[1:]))
        f = np.asfortranarray(a).reshape(m, n, order='F')
        f = np.asfortranarray(f)
        return L.standin_double(m, n, f.ctypes.data_as(C.POINTER(C.c_double)))

    def call(self, nlhs, *args):
        """Returns the outputs (numpy arrays; strings for char; handles kept
        as mxArrays under ('out', i))."""
        L = self.L
        ins = [self._arg(a) for a in args]
        inarr = (C.c_void_p * len(ins))(*ins)
        out = (C.c_void_p * max(nlhs, 1))()
        idb = C.create_string_buffer(256)
        msg = C.create_string_buffer(1024)
        rc = L.standin_call(nlhs, out, len(ins), inarr, idb, msg, 1024)
        for a, x in zip(ins, args):
            if not (isinstance(x, tuple) and x and x[0] == 'handle'):
                L.standin_free(a)
        if rc:
            raise MexError(idb.value.decode(), msg.value.decode())
        res = []
        for i in range(nlhs):
            a = out[i]
            if not a:
                res.append(None)
                continue
            cls = L.standin_class(a)
            if cls == 4:
                b = C.create_string_buffer(256)
                L.standin_read_string(a, b, 256)
                res.append(b.value.decode())
                L.standin_free(a)
            elif cls == 13:
                res.append(a)            # a handle: kept alive as an mxArray
            else:
                m, n = L.standin_rows(a), L.standin_cols(a)
                v = np.empty(m * n)
                if m * n:
                    L.standin_read(a, v.ctypes.data_as(C.POINTER(C.c_double)))
                v = v.reshape(m, n, order='F')
                res.append(v.astype(bool) if cls == 3 else v)
                L.standin_free(a)
        return res

    def keep(self, name, h):
        self.handles[name] = h
        return ('handle', name)


class MexError(RuntimeError):
    def __init__(self, ident, msg):
        super().__init__('%s: %s' % (ident, msg))
        self.ident = ident
        self.msg = msg


@pytest.fixture(scope='module')
def gw(tmp_path_factory):
    _lib.lib()
    out = str(tmp_path_factory.mktemp('mex') / 'libfestung_mex_standin.so')
    pkg = os.path.join(ROOT, 'festung_amd')
    cmd = ['g++', '-std=c++17', '-O1', '-Wall', '-Wextra', '-Werror', '-shared', '-fPIC',
           '-I' + os.path.join(ROOT, 'tests', 'mex'), '-I' + os.path.join(ROOT, 'include'),
           os.path.join(pkg, 'mex', 'festung_mex.cpp'), os.path.join(ROOT, 'tests', 'mex', 'mex_standin.cpp'),
           '-L' + pkg, '-lfestung_hip', '-Wl,-rpath,' + pkg, '-o', out]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return Gateway(out)


def test_gateway_compiles_and_reports_version(gw):
    (v,) = gw.call(1, 'version')
    assert v == _lib.lib().festung_version().decode()


def test_grid_commands_match_binding(gw):
    (gh,) = gw.call(1, 'grid_square', 0.25)
    g = gw.keep('sq', gh)
    assert gw.L.standin_lock_count() == 1
    ref = Grid.domainSquare(0.25)
    (V0T,) = gw.call(1, 'grid_field', g, 'V0T')
    assert np.array_equal(V0T, ref.V0T + 1)                     # 1-based
    (T0E,) = gw.call(1, 'grid_field', g, 'T0E')
    assert np.array_equal(T0E, ref.T0E + 1) and T0E.min() == 0  # 0 = none
    (mD,) = gw.call(1, 'grid_field', g, 'markE0TbdrD')
    assert mD.dtype == bool and np.array_equal(mD, ref.markE0TbdrD)
    (nu,) = gw.call(1, 'grid_field', g, 'nuE0T')
    assert nu.shape == (ref.numT, 6) and np.array_equal(nu, ref.nuE0T.reshape(ref.numT, 6, order='F'))
    (cv,) = gw.call(1, 'grid_field', g, 'coordV')
    assert np.array_equal(cv, ref.coordV)
    Q1, Q2, _ = quadRule2D(4)
    X1, X2 = gw.call(2, 'map_ref_to_phy', g, Q1, Q2)
    R1, R2 = ref.mapRef2Phy(Q1, Q2)
    assert np.array_equal(X1, R1) and np.array_equal(X2, R2)
    # an unstructured mesh through grid_create (1-based V0T, boundary ids)
    (g2h,) = gw.call(1, 'grid_create', cv, V0T, gw.call(1, 'grid_field', g, 'idE')[0])
    g2 = gw.keep('c', g2h)
    (E0T,) = gw.call(1, 'grid_field', g2, 'E0T')
    assert np.array_equal(E0T, ref.E0T + 1)
    assert np.array_equal(gw.call(1, 'grid_field', g2, 'markV0TbdrD')[0], ref.markV0TbdrD)
    gw.call(0, 'grid_destroy', g2)
    gw.call(0, 'grid_destroy', g)
    assert gw.L.standin_lock_count() == 0


def test_tables_match_binding(gw):
    for ord_, tau, t0 in ((1, 0.1, 0.0), (2, 0.25, 1.0), (3, 0.01, 2.5), (4, 0.2, 0.0)):
        t, w = gw.call(2, 'rk', ord_, tau, t0)
        tr, wr = rungeKuttaExplicit(ord_, tau, t0)
        assert np.array_equal(t.ravel(), tr) and np.array_equal(w.ravel(), wr, equal_nan=True)
    for p, name, ord_ in ((2, 'hatG', 0), (3, 'hatRoff', 0), (4, 'phi2D', 8)):
        (T,) = gw.call(1, 'ref', p, name, ord_)
        assert np.array_equal(T.ravel(), ref_table(p, name, ord_))


def test_errors_are_raised_through_the_interpreter(gw):
    with pytest.raises(MexError, match="unknown command 'nope'"):
        gw.call(0, 'nope')
    with pytest.raises(MexError) as e:
        gw.call(1, 'rk', 5, 0.1, 0.0)
    assert e.value.ident == 'festung:hip' and 'not supported' in e.value.msg
    with pytest.raises(MexError, match='usage'):
        gw.call(1, 'grid_square')
    with pytest.raises(MexError, match='festung handle'):
        gw.call(1, 'grid_field', 3.0, 'V0T')
    (gh,) = gw.call(1, 'grid_square', 0.5)
    g = gw.keep('e', gh)
    with pytest.raises(MexError, match='V0T holds a vertex id out of range'):
        gw.call(1, 'grid_create', np.zeros((3, 2)), np.array([[1.0, 2.0, 4.0]]))
    import torch
    if not torch.cuda.is_available():
        with pytest.raises(MexError) as e:
            gw.call(1, 'create', g, 2.0, 3.0, 0.0)
        assert e.value.ident == 'festung:hip'
    gw.call(0, 'grid_destroy', g)
    with pytest.raises(MexError, match='not a live grid handle'):
        gw.call(0, 'grid_destroy', g)
    assert gw.L.standin_lock_count() == 0


@pytest.mark.gpu
def test_device_commands_match_binding(gpu, gw):
    """A p = 2 problem through the gateway and through the Python binding:
    rhs, operator applies on the element-major vector, the stages of one
    step, the stand-alone limiters: the same bits."""
    from festung_amd.advection import testcases as TC
    from test_gpu_parity import product_problem, product_step
    tc = TC.rotating_humps()
    d = product_problem(tc, 2, 1 / 8, 3, True, numSteps=20, tEnd=0.5)
    P = d.problem
    (gh,) = gw.call(1, 'grid_square', 1 / 8)
    g = gw.keep('dev_g', gh)
    (hh,) = gw.call(1, 'create', g, 2.0, 3.0, float(_lib.FLAG_LIMITER))
    h = gw.keep('dev_h', hh)
    X1, X2 = d.XEdge
    gw.call(0, 'velocity_points', h, tc.u1Cont(0, *d.XE), tc.u2Cont(0, *d.XE), tc.u1Cont(0, X1, X2),
            tc.u2Cont(0, X1, X2))
    d.t = np.array([0.0])
    from festung_amd.advection import steps as S
    S.preprocessSubStep(d, 1, 1)
    rng = np.random.default_rng(8)
    c = rng.standard_normal((P.K, P.N))
    gw.call(0, 'set_state', h, c)
    P.set_state(c)
    assert np.array_equal(gw.call(1, 'rhs', h)[0], P.rhs())
    ev = c.reshape(-1, 1)                                  # element-major KN x 1
    for op in ('G1', 'G2', 'R', 'Minv'):
        (y,) = gw.call(1, 'apply', h, op, ev)
        assert y.shape == (P.K * P.N, 1)
        assert np.array_equal(y.reshape(P.K, P.N), P.apply_operator(op, c)), op
    # one step stage by stage (with the limiter)
    c0 = d.problem.get_state()
    gw.call(0, 'set_state', h, c0)
    gw.call(0, 'vertex_dirichlet', h, S.vertexDirichlet(d, 0.0))
    gw.call(0, 'begin_step', h)
    for s in (1, 2, 3):
        gw.call(0, 'stage', h, float(s), d.tau)
    ref = product_step(d, 1)
    assert np.array_equal(gw.call(1, 'get_state', h)[0], ref)
    cen = rng.standard_normal(P.K)
    val = cen[:, None] + rng.standard_normal((P.K, 3))
    a, mn, mx = gw.call(3, 'vertex_limiter', h, cen, val, np.zeros((P.K, 3), dtype=bool), np.zeros((P.K, 3)))
    ar, (mnr, mxr) = P.vertex_based_limiter(cen, val)
    assert np.array_equal(a.ravel(), ar) and np.array_equal(mn, mnr) and np.array_equal(mx, mxr)
    gw.call(0, 'destroy', h)
    gw.call(0, 'grid_destroy', g)
    assert gw.L.standin_lock_count() == 0


@pytest.mark.gpu
def test_operator_objects_run_solve_substep_unchanged(gpu, gw):
    """What the Matlab operator objects do for advection/solveSubStep.m:62-66
    (sysA = -globG{1} - globG{2} + globR; cDiscDot = globM \\ (sysV - sysA *
    cDiscRK{s})): the element and edge assembly shims hand their velocity
    fields to the device separately ('velocity' with [] for the other
    fields), every product is a gateway 'apply' on the element-major KN x 1
    vector, and sysV comes from the reference's own vector assembly (the
    oracle here).  Against the oracle's sparse path, 1e-12 relative."""
    from helpers import oracle_problem, oracle_stage_operators, rel_l2
    from festung_amd.advection import testcases as TC
    from oracle import advection as OA
    tc = dict(TC.rotating_humps())
    tc['cDCont'] = lambda t, x1, x2: 0.2 + 0.1 * x1 + 0 * t
    for p in (1, 2, 4):
        o = oracle_problem(tc, p, 1 / 8, min(p + 1, 3), False)
        o = oracle_stage_operators(o, 0.0)
        (gh,) = gw.call(1, 'grid_square', 1 / 8)
        g = gw.keep('oo_g%d' % p, gh)
        (hh,) = gw.call(1, 'create', g, float(p), float(min(p + 1, 3)), 0.0)
        h = gw.keep('oo_h%d' % p, hh)
        gw.call(0, 'velocity', h, o.u1Disc, o.u2Disc, np.empty((0, 0)))     # assembleMatElemDphiPhiFuncDisc
        gw.call(0, 'velocity', h, np.empty((0, 0)), np.empty((0, 0)), o.vNormalOnQuadEdge)  # ...EdgePhiPhiValUpwind
        c = np.random.default_rng(60 + p).standard_normal(o.K * o.N)[:, None]
        sysV = (o.globL - o.globKD - o.globKN)[:, None]
        G1c, G2c, Rc = (gw.call(1, 'apply', h, op, c)[0] for op in ('G1', 'G2', 'R'))
        sysAc = -G1c - G2c + Rc
        (cdot,) = gw.call(1, 'apply', h, 'Minv', sysV - sysAc)
        ref = OA.computeDot(o, c.ravel())
        assert rel_l2(cdot.ravel(), ref) < 1e-12, p
        gw.call(0, 'destroy', h)
        gw.call(0, 'grid_destroy', g)

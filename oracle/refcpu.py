"""ctypes wrapper of oracle/refcpu (test infrastructure): the
reference-structured C++ restatement of the step (per-stage CSR assembly of
G1, G2 and R, their sparse sum, one SpMV, block solves, the Taylor-basis
limiter with CSR vertex patches; OpenMP), fed with the numpy oracle's
preprocessed problem.  Used as the timed CPU baseline and, in the tests, as
a second oracle path against the numpy one.

Only workloads whose coefficients do not depend on time are taken (the
point values are evaluated once, at t = 0), and no Neumann marks.
"""
import ctypes as C
import os
import subprocess

import numpy as np
import scipy.sparse as sp

from . import assembly as A
from .basis import gammaMap
from .grid import mapRef2Phy
from .quadrature import quadRule1D, quadRule2D

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, 'refcpu', 'refcpu.cpp')
LIB = os.path.join(HERE, 'refcpu', 'librefcpu.so')

dp = C.POINTER(C.c_double)
ip = C.POINTER(C.c_int32)
lp = C.POINTER(C.c_int64)
bp = C.POINTER(C.c_uint8)


class _Input(C.Structure):
    _fields_ = [('K', C.c_int), ('N', C.c_int), ('R', C.c_int), ('R2', C.c_int), ('ordRK', C.c_int),
                ('limiter', C.c_int),
                ('hatM', dp), ('hatG', dp), ('hatRdiag', dp), ('hatRoff', dp), ('W2', dp), ('phi2', dp),
                ('W1', dp), ('phi1', dp),
                ('B', dp), ('detJ', dp), ('areaE0T', dp), ('nuE0T', dp),
                ('nb', ip), ('np', ip),
                ('markElemEdges', bp), ('markN', bp), ('markD', bp),
                ('fq', dp), ('u1q', dp), ('u2q', dp), ('u1e', dp), ('u2e', dp), ('cDe', dp),
                ('MT', dp), ('MDT', dp), ('phiTV0T', dp),
                ('patchPtr', lp * 3), ('patchIdx', ip * 3),
                ('markV0TbdrD', bp), ('cDV0T', dp), ('c0', dp)]


def build(force=False):
    """Compile librefcpu.so next to its source (g++ -fopenmp)."""
    if force or not os.path.exists(LIB) or os.path.getmtime(LIB) < os.path.getmtime(SRC):
        tmp = LIB + '.tmp'
        subprocess.run(['g++', '-O2', '-std=c++17', '-fopenmp', '-fPIC', '-shared', '-Wall', SRC, '-o', tmp],
                       check=True, capture_output=True, text=True)
        os.replace(tmp, LIB)
    return LIB


_lib = None


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB):
            build()
        L = C.CDLL(LIB)
        L.refcpu_create.argtypes = [C.POINTER(_Input), C.POINTER(C.c_void_p)]
        L.refcpu_run.argtypes = [C.c_void_p, C.c_int, C.c_double, C.c_int, dp]
        L.refcpu_get_state.argtypes = [C.c_void_p, dp]
        L.refcpu_set_state.argtypes = [C.c_void_p, dp]
        L.refcpu_destroy.argtypes = [C.c_void_p]
        L.refcpu_destroy.restype = None
        _lib = L
    return _lib


def _blocks(M, K, N):
    """Diagonal N x N blocks of a block-diagonal KN x KN sparse matrix -> K x N x N."""
    Mc = M.tocoo()
    out = np.zeros((K, N, N))
    k = Mc.row // N
    assert np.all(k == Mc.col // N)
    out[k, Mc.row % N, Mc.col % N] = Mc.data
    return out


class RefCpu:
    """The oracle problem pd (after preprocessProblem / initializeProblem)
    as a refcpu problem; steps(n, threads) runs n steps and returns the
    seconds of the step loop."""

    def __init__(self, pd):
        g, p, K, N = pd.g, pd.p, pd.K, pd.N
        assert not np.any(g.markE0TbdrN), 'Neumann data are not restated in refcpu'
        keep = []

        def arr(a, dtype=np.float64):
            a = np.ascontiguousarray(a, dtype=dtype)
            keep.append(a)
            return a.ctypes.data_as({np.float64: dp, np.int32: ip, np.int64: lp, np.uint8: bp}[dtype])

        bases = pd.basesOnQuad
        q2 = max(2 * p, 1)
        Q1, Q2, W2 = quadRule2D(q2)
        Q, W1 = quadRule1D(2 * p + 1)
        R = len(W1)
        XE = (mapRef2Phy(g, 0, Q1, Q2), mapRef2Phy(g, 1, Q1, Q2))
        X1e = np.zeros((K, 3, R)); X2e = np.zeros((K, 3, R))
        for n in range(3):
            A1, A2 = gammaMap(n + 1, Q)
            X1e[:, n, :] = mapRef2Phy(g, 0, A1, A2)
            X2e[:, n, :] = mapRef2Phy(g, 1, A1, A2)
        t0 = 0.0
        nb = np.full((K, 3), -1, dtype=np.int32)
        npn = np.full((K, 3), -1, dtype=np.int32)
        mee = np.zeros((K, 3), dtype=np.uint8)
        for nn in range(3):
            anyNb = None
            for q in range(3):
                M = sp.csr_matrix(g.markE0TE0T[(nn, q)])
                rows, cols = M.nonzero()
                nb[rows, nn] = cols
                npn[rows, nn] = q
                anyNb = M if anyNb is None else anyNb + M
            mee[:, nn] = (np.asarray(anyNb.sum(axis=1)).ravel() > 0) | (g.idE0T[:, nn] != 0)
        s = _Input()
        s.K, s.N, s.R, s.R2, s.ordRK, s.limiter = K, N, R, len(W2), pd.ordRK, int(bool(pd.isSlopeLim))
        s.hatM, s.hatG = arr(pd.hatM), arr(pd.hatG)
        s.hatRdiag, s.hatRoff = arr(pd.hatRdiagOnQuad), arr(pd.hatRoffdiagOnQuad)
        s.W2, s.phi2 = arr(W2), arr(bases['phi2D'][q2][:, :N])
        s.W1, s.phi1 = arr(W1), arr(bases['phi1D'][2 * p + 1])
        s.B, s.detJ, s.areaE0T, s.nuE0T = arr(g.B), arr(g.detJ0T), arr(g.areaE0T), arr(g.nuE0T)
        s.nb, s.np = arr(nb, np.int32), arr(npn, np.int32)
        s.markElemEdges = arr(mee, np.uint8)
        s.markN = arr(g.markE0TbdrN, np.uint8)
        s.markD = arr(g.markE0TbdrD, np.uint8)
        s.fq = arr(pd.fCont(t0, *XE) + 0 * XE[0])
        s.u1q = arr(pd.u1Cont(t0, *XE) + 0 * XE[0])
        s.u2q = arr(pd.u2Cont(t0, *XE) + 0 * XE[0])
        s.u1e = arr(pd.u1Cont(t0, X1e, X2e) + 0 * X1e)
        s.u2e = arr(pd.u2Cont(t0, X1e, X2e) + 0 * X1e)
        s.cDe = arr(pd.cDCont(t0, X1e, X2e) + 0 * X1e)
        if pd.isSlopeLim:
            s.MT = arr(_blocks(pd.globMTaylor, K, N))
            s.MDT = arr(_blocks(pd.globMDiscTaylor, K, N))
            s.phiTV0T = arr(bases['phiTaylorV0T'])
            for n in range(3):
                P = sp.csr_matrix(g.markV0TT0V[n])
                P.sort_indices()
                s.patchPtr[n] = arr(P.indptr, np.int64)
                s.patchIdx[n] = arr(P.indices, np.int32)
            s.markV0TbdrD = arr(g.markV0TbdrD, np.uint8)
            s.cDV0T = arr(A.computeFuncContV0T(g, lambda x1, x2: pd.cDCont(t0, x1, x2) + 0 * x1))
        s.c0 = arr(pd.cDisc.reshape(K, N))
        h = C.c_void_p()
        rc = lib().refcpu_create(C.byref(s), C.byref(h))
        if rc != 0:
            raise RuntimeError('refcpu_create failed (%d)' % rc)
        self._h, self._keep, self.K, self.N, self.tau = h, keep, K, N, pd.tau

    def steps(self, n, threads=1, tau=None):
        sec = C.c_double()
        rc = lib().refcpu_run(self._h, int(n), float(self.tau if tau is None else tau), int(threads), C.byref(sec))
        if rc != 0:
            raise RuntimeError('refcpu_run failed (%d)' % rc)
        return sec.value

    def state(self):
        c = np.zeros((self.K, self.N))
        lib().refcpu_get_state(self._h, c.ctypes.data_as(dp))
        return c

    def set_state(self, c):
        c = np.ascontiguousarray(c, dtype=np.float64).reshape(self.K, self.N)
        lib().refcpu_set_state(self._h, c.ctypes.data_as(dp))

    def __del__(self):
        if getattr(self, '_h', None) is not None and _lib is not None:
            _lib.refcpu_destroy(self._h)
            self._h = None

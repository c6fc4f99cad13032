"""Host-side checks of the C-ABI library (no GPU): it loads, exports every
symbol the header declares, builds grids bit-identical to the oracle's
(connectivity, numbering and geometry), computes the same reference tables,
and reports errors as status codes."""
import ctypes as C
import os
import re

import numpy as np
import pytest
from scipy.spatial import Delaunay

from festung_amd import _lib
from festung_amd.core import Grid, quadRule1D, quadRule2D, ref_table
from oracle import basis as OB
from oracle import grid as OG
from oracle import quadrature as OQ

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
INT_FIELDS = ('V0T', 'V0E', 'T0E', 'E0T', 'E0E', 'idE', 'idE0T')
MARK_FIELDS = ('markE0Tint', 'markE0TbdrD', 'markV0TbdrD')
F64_FIELDS = ('coordV', 'areaE', 'nuE', 'baryE', 'areaT', 'baryT', 'B', 'areaE0T', 'sigE0T', 'nuE0T',
              'coordV0T')


def header_symbols():
    txt = open(os.path.join(ROOT, 'include', 'festung_hip.h')).read()
    txt = re.sub(r'/\*.*?\*/', '', txt, flags=re.S)
    return sorted(set(re.findall(r'\b(festung_\w+)\s*\(', txt)))


def test_library_exports_header_symbols():
    L = _lib.lib()
    syms = header_symbols()
    assert len(syms) >= 30
    for s in syms:
        assert hasattr(L, s), s
        assert s in _lib.SIGNATURES, s
    assert L.festung_version().startswith(b'festung-hip')


@pytest.mark.parametrize('h', [1 / 3, 1 / 8, 1 / 10, 1 / 12, 2 ** -6])
def test_domain_square_bit_exact(h):
    g = Grid.domainSquare(h)
    o = OG.domainSquare(h)
    OG.markBoundaries(o)
    assert (g.numT, g.numV, g.numE) == (o.numT, o.numV, o.numE)
    for f in INT_FIELDS + MARK_FIELDS:
        assert np.array_equal(getattr(g, f), o[f]), f
    for f in F64_FIELDS:
        assert np.array_equal(getattr(g, f), o[f]), f
    # neighbour lists are the nonzeros of markE0TE0T{nn, np}
    nbr, nbe = g.nbr, g.nbrEdge
    for (nn, np_), M in o.markE0TE0T.items():
        r, c = M.nonzero()
        assert np.array_equal(nbr[r, nn], c) and np.all(nbe[r, nn] == np_)
    # vertex patches == the rows of markV0TT0V
    OG.computeDerivedGridData(o)
    ptr, idx = g.v2tPtr, g.v2tIdx
    for n in range(3):
        M = o.markV0TT0V[n].tocsr()
        for k in range(0, g.numT, max(1, g.numT // 50)):
            v = g.V0T[k, n]
            assert np.array_equal(np.sort(idx[ptr[v]:ptr[v + 1]]), np.sort(M.indices[M.indptr[k]:M.indptr[k + 1]]))


def _delaunay_mesh(n, seed):
    rng = np.random.default_rng(seed)
    pts = np.vstack([rng.uniform(0, 1, (n, 2)), [[0, 0], [1, 0], [0, 1], [1, 1]]])
    tri = Delaunay(pts).simplices.astype(np.int64)
    a, b, c = pts[tri[:, 0]], pts[tri[:, 1]], pts[tri[:, 2]]
    cw = (b[:, 0] - a[:, 0]) * (c[:, 1] - a[:, 1]) - (b[:, 1] - a[:, 1]) * (c[:, 0] - a[:, 0]) < 0
    tri[cw] = tri[cw][:, [0, 2, 1]]
    return pts, tri


@pytest.mark.parametrize('seed', [0, 1, 2])
def test_unstructured_grid_bit_exact(seed):
    pts, tri = _delaunay_mesh(300, seed)
    g = Grid.generateGridData(pts, tri)
    o = OG.generateGridData(pts, tri)
    for f in ('V0T', 'V0E', 'T0E', 'E0T', 'E0E'):
        assert np.array_equal(getattr(g, f), o[f]), f
    for f in F64_FIELDS:
        assert np.array_equal(getattr(g, f), o[f]), f
    # boundary ids from the caller, as in the reference (generateGridData leaves idE to the caller)
    idE = (o.E0E[:, 1] < 0).astype(np.int32)
    g.set_boundary(idE)
    o.idE = idE
    o.idE0T = idE[o.E0T]
    OG.markBoundaries(o)
    for f in MARK_FIELDS:
        assert np.array_equal(getattr(g, f), o[f]), f


def test_quadrature_and_tables_match_oracle():
    for q in range(0, 18):
        a = quadRule1D(q); b = OQ.quadRule1D(q)
        assert np.array_equal(a[0], b[0]) and np.array_equal(a[1], b[1])
    for q in range(0, 13):
        a = quadRule2D(q); b = OQ.quadRule2D(q)
        tol = 0 if q <= 6 else 2e-15       # collapsed rules: eigen-solver last bits
        for x, y in zip(a, b):
            assert np.max(np.abs(x - y)) <= tol, q
    for p in range(5):
        N = (p + 1) * (p + 2) // 2
        bq = OB.computeBasesOnQuad(N)
        refs = {'hatM': OB.integrateRefElemPhiPhi(N, bq), 'hatG': OB.integrateRefElemDphiPhiPhi(N, bq),
                'hatRdiag': OB.integrateRefEdgePhiIntPhiIntPerQuad(N, bq),
                'hatRoff': OB.integrateRefEdgePhiIntPhiExtPerQuad(N, bq)}
        for name, ref in refs.items():
            got = ref_table(p, name)
            assert np.allclose(got, ref.flatten(order='F'), rtol=0, atol=1e-12 * max(1, np.abs(ref).max())), (p, name)
        for ord_ in ([2 * p, 2 * p + 1] if p > 0 else [1]):
            assert np.allclose(ref_table(p, 'phi2D', ord_), bq['phi2D'][ord_].flatten(order='F'), atol=1e-13)
            assert np.allclose(ref_table(p, 'gradPhi2D', ord_), bq['gradPhi2D'][ord_].flatten(order='F'), atol=1e-12)
            assert np.array_equal(ref_table(p, 'phi1D', ord_), bq['phi1D'][ord_].flatten(order='F'))
            assert np.array_equal(ref_table(p, 'thetaPhi1D', ord_), bq['thetaPhi1D'][ord_].flatten(order='F'))


def test_element_term_rule():
    """The element term's rule: the reference's quadRule2D(max(2p, 1)) for
    p <= 3; at p = 4 a fully symmetric 28-point rule with interior points and
    positive weights that integrates every monomial of degree <= 11 exactly
    (closed form a! b! / (a + b + 2)!) and fails at degree 12, and with which
    the hatG block (integrands of degree 11) equals the oracle's
    quadRule2D(8) block to rounding."""
    from math import factorial
    for p in range(4):
        got = ref_table(p, 'quadG')
        ref = np.concatenate(quadRule2D(max(2 * p, 1)))
        assert np.array_equal(got, ref), p
    x, y, w = ref_table(4, 'quadG').reshape(3, -1)
    assert x.size == 28
    assert (w > 0).all() and (x > 0).all() and (y > 0).all() and (x + y < 1).all()
    bary = np.stack([1 - x - y, x, y], axis=1)
    for perm in ([1, 0, 2], [0, 2, 1], [2, 1, 0]):
        pb = bary[:, perm]
        for r in range(28):
            d = np.abs(bary - pb[r]).max(axis=1) + np.abs(w - w[r])
            assert d.min() < 1e-15
    for deg in range(13):
        err = max(abs((w * x ** a * y ** (deg - a)).sum() - factorial(a) * factorial(deg - a) / factorial(deg + 2))
                  * factorial(deg + 2) / (factorial(a) * factorial(deg - a)) for a in range(deg + 1))
        assert (err < 2e-15) if deg <= 11 else (err > 1e-6), (deg, err)
    N = 15
    phi = np.array([OB.phi(i, x, y) for i in range(1, N + 1)])
    grad = np.array([[OB.gradPhi(i, m, x, y) for m in (1, 2)] for i in range(1, N + 1)])
    G28 = np.einsum('q,imq,jq,lq->ijlm', w, grad, phi, phi)
    Gref = OB.integrateRefElemDphiPhiPhi(N, OB.computeBasesOnQuad(N))
    # both sums round at about 1e-14 of the largest entry (terms of size ~50
    # cancel); a wrong point or weight shows at 1e-4 and above
    assert np.abs(G28 - Gref).max() <= 3e-14 * np.abs(Gref).max()


@pytest.mark.parametrize('p', [1, 2, 3, 4])
def test_taylor_tables(p):
    """F maps DG coefficients to [mean; reference derivatives at the centroid]:
    row 0 is sqrt(2) e_1, rows of degree 1 are gradPhi at (1/3, 1/3), rows of
    degree d >= 2 are finite differences of gradPhi; Finv is its inverse."""
    N = (p + 1) * (p + 2) // 2
    F = ref_table(p, 'taylorF').reshape(N, N, order='F')
    Fi = ref_table(p, 'taylorFinv').reshape(N, N, order='F')
    assert np.allclose(F @ Fi, np.eye(N), atol=1e-10)
    assert np.allclose(F[0], np.eye(N)[0] * np.sqrt(2))
    # a Taylor row of degree d involves only DG functions of degree >= d
    # (exact zeros, which the kernels skip)
    deg = [d for d in range(p + 1) for _ in range(d + 1)]
    for r in range(N):
        for j in range(N):
            if deg[r] > deg[j]:
                assert F[r, j] == 0 and Fi[r, j] == 0
    c = 1 / 3
    for j in range(N):
        assert abs(F[1, j] - OB.gradPhi(j + 1, 1, c, c)) < 1e-12
        assert abs(F[2, j] - OB.gradPhi(j + 1, 2, c, c)) < 1e-12
        if p >= 2:
            h = 1e-5
            d11 = (OB.gradPhi(j + 1, 1, c + h, c) - OB.gradPhi(j + 1, 1, c - h, c)) / (2 * h)
            d12 = (OB.gradPhi(j + 1, 1, c, c + h) - OB.gradPhi(j + 1, 1, c, c - h)) / (2 * h)
            d22 = (OB.gradPhi(j + 1, 2, c, c + h) - OB.gradPhi(j + 1, 2, c, c - h)) / (2 * h)
            assert np.allclose(F[3:6, j], [d11, d12, d22], atol=1e-5 * max(1, np.abs(F[3:6, j]).max()))


def test_error_paths():
    L = _lib.lib()
    h = C.c_void_p()
    assert L.festung_domain_square(-1.0, 0.0, 1.0, C.byref(h)) == _lib.ERR_ARG
    assert b'h > 0' in L.festung_last_error()
    assert L.festung_domain_square(0.5, 0.0, 1.0, None) == _lib.ERR_ARG
    # clockwise element is rejected
    coord = np.asfortranarray([[0, 0], [1, 0], [0, 1]], dtype=np.float64)
    V0T = np.asfortranarray([[0, 2, 1]], dtype=np.int32)
    st = L.festung_grid_create(_lib.ptr(coord), 3, _lib.ptr(V0T), 1, C.byref(h))
    assert st == _lib.ERR_ARG and b'counter-clockwise' in L.festung_last_error()
    # out-of-range vertex id
    V0T = np.asfortranarray([[0, 1, 7]], dtype=np.int32)
    assert L.festung_grid_create(_lib.ptr(coord), 3, _lib.ptr(V0T), 1, C.byref(h)) == _lib.ERR_ARG
    # unsupported degree, unknown table
    cnt = C.c_int64()
    assert L.festung_ref_size(5, b'hatM', 0, C.byref(cnt)) == _lib.ERR_UNSUPPORTED
    assert L.festung_ref_size(1, b'nope', 0, C.byref(cnt)) == _lib.ERR_ARG
    assert L.festung_ref_size(1, b'quad2D', 99, C.byref(cnt)) == _lib.ERR_ARG
    # unknown grid field
    g = Grid.domainSquare(0.5)
    with pytest.raises(_lib.FestungError):
        g.field('nope') if False else _lib.check(L.festung_grid_field_size(g._h, b'nope', C.byref(cnt), None))
    # a successful call clears the message
    assert L.festung_grid_sizes(g._h, None, None, None) == _lib.OK
    assert L.festung_last_error() == b''


def test_problem_needs_device_and_fails_loudly():
    """Without a visible GPU the device path raises; there is no CPU fallback."""
    import torch
    if torch.cuda.is_available():
        pytest.skip('GPU visible')
    from festung_amd.core import Problem
    g = Grid.domainSquare(0.5)
    with pytest.raises(_lib.FestungError) as e:
        Problem(g, 1, 2)
    assert e.value.code == _lib.ERR_HIP


def test_host_mirror_defaults_follow_reference():
    """advection/configureProblem.m:64-96 defaults (triangles)."""
    from festung_amd.advection import steps as S
    pd = S.configureProblem(S.Struct())
    assert pd.testcase == 'solid_body' and pd.hmax == 2 ** -6 and pd.p == 0
    assert pd.qOrd == 1 and pd.ordRK == 1 and pd.numSteps == 100
    assert abs(pd.tEnd - (100 / (3142 * 1.5)) * 2 * np.pi) < 1e-15
    assert pd.isSlopeLim is False and pd.typeSlopeLim == 'linear'
    pd = S.configureProblem(S.Struct(p=2))
    assert pd.qOrd == 5 and pd.ordRK == 3
    with pytest.raises(AssertionError):
        S.configureProblem(S.Struct(p=0, isSlopeLim=True))
    with pytest.raises(NotImplementedError):
        S.configureProblem(S.Struct(isQuadri=True))


def test_lagrange_projection_and_vtk_tecplot_files(tmp_path):
    """Lagrange-point values of DG data (p = 1: a linear function's projection
    is exact at the vertices) and the reference's .vtu / .plt layouts, read
    back with an XML parser and by counting the Tecplot blocks."""
    import xml.etree.ElementTree as ET
    from festung_amd import output as O
    g = Grid.domainSquare(1 / 3)
    f = lambda x1, x2: 0.3 * x1 - 0.7 * x2 + 0.1
    # DG coefficients of f for p = 1 from its values at the vertices (exact interpolation)
    ph = O.phi_at_points(1, [0, 1, 0], [0, 0, 1])                     # 3 x 3
    vals = f(g.coordV0T[:, :, 0], g.coordV0T[:, :, 1])                # K x 3
    c = np.linalg.solve(ph, vals.T).T
    lag = O.projectDataDisc2DataLagr(c)
    assert np.allclose(lag, vals, atol=1e-14)
    c2 = np.zeros((g.numT, 6)); c2[:, :3] = c                          # the same function at p = 2
    lag2 = O.projectDataDisc2DataLagr(c2)
    mid = lambda a, b: f(0.5 * (g.coordV0T[:, a, 0] + g.coordV0T[:, b, 0]), 0.5 * (g.coordV0T[:, a, 1] + g.coordV0T[:, b, 1]))
    assert np.allclose(lag2[:, 3], mid(1, 2)) and np.allclose(lag2[:, 4], mid(0, 2)) and np.allclose(lag2[:, 5], mid(0, 1))
    for d, nodes, cid in ((lag, 3, 5), (lag2, 6, 22)):
        vtu, plt = O.visualizeDataLagr(g, d, 'u_h', str(tmp_path / 'out' / 'sol'), 7, ['vtk', 'tec'])
        assert vtu.endswith('sol.7.vtu') and plt.endswith('sol.7.plt')
        piece = ET.parse(vtu).getroot().find('UnstructuredGrid/Piece')
        assert int(piece.get('NumberOfPoints')) == nodes * g.numT and int(piece.get('NumberOfCells')) == g.numT
        arrays = {a.get('Name'): np.array(a.text.split(), dtype=float) for a in piece.iter('DataArray')}
        assert np.all(arrays['types'] == cid)
        pts = np.array(piece.find('Points/DataArray').text.split(), dtype=float).reshape(-1, 3)
        # every node value is f at the node (the data are exact for the linear f)
        assert np.allclose(arrays['u_h'], f(pts[:, 0], pts[:, 1]), atol=1e-8)
        txt = open(plt).read().split()
        assert 'ET=TRIANGLE,' in txt and ('E=%d,' % (g.numT if nodes == 3 else 4 * g.numT)) in txt


def test_hotstart_roundtrip(tmp_path):
    from festung_amd import output as O
    c = np.arange(12.0).reshape(4, 3)
    fn = O.outputHotstart(str(tmp_path / 'hs' / 'state'), cDisc=c, t=0.25, nStep=7)
    assert fn.endswith('state.mat')
    d = O.readHotstart(fn)
    assert np.array_equal(d['cDisc'], c) and d['t'] == 0.25 and d['nStep'] == 7


def _least_advances(n, c):
    """Least number of forward steps (besides one recomputation per reversed
    step) that reverses n steps keeping at most c states, the start state
    included, by exhaustive recursion over the position of the next
    snapshot: reversing [0, n) either walks from the start to every step
    (no free snapshot) or advances k steps, stores that state, reverses
    [k, n) with one snapshot fewer and then [0, k) with the same number."""
    from functools import lru_cache

    @lru_cache(None)
    def best(n, s):
        if n <= 1:
            return 0
        if s == 0:
            return n * (n - 1) // 2
        return min(k + best(n - k, s - 1) + best(k, s) for k in range(1, n))
    return best(n, min(c, n) - 1)


def test_binomial_schedule_is_optimal():
    """The schedule reaches the least forward-step count on every n <= 60,
    c <= 7, and Griewank's closed form r n - C(c + r, c + 1) (r the least
    integer with C(c + r, c) >= n) beyond; every step is reversed once, so
    the action count is at least n."""
    from math import comb
    from festung_amd.core import binomial_schedule
    for c in range(1, 8):
        for n in range(1, 61):
            adv, act = binomial_schedule(n, c)
            assert adv == _least_advances(n, c), (n, c)
            assert act >= n
    for n, c in ((400, 21), (400, 6), (1000, 10), (777, 2)):
        r = 0
        while comb(c + r, c) < n:
            r += 1
        assert binomial_schedule(n, c)[0] == r * n - comb(c + r, c + 1)
    assert binomial_schedule(10, 1)[0] == 45          # no snapshot: walk from the start
    assert binomial_schedule(10, 10)[0] == 9          # every state stored
    with pytest.raises(_lib.FestungError):
        binomial_schedule(0, 3)
    with pytest.raises(_lib.FestungError):
        binomial_schedule(5, 0)


def test_tuning_table_file_roundtrip(tmp_path):
    """festung_tune_load / festung_tune_save (host only): comments and blank
    lines are skipped, a malformed file is rejected as a whole with its line
    number, and a saved table loads back to the same entries."""
    from festung_amd import core
    from festung_amd._lib import FestungError
    core.tune_clear()
    assert core.tune_count() == 0
    src = tmp_path / 'in.txt'
    src.write_text('# device p ordRK limiter log2K elemsPerBlock\n\n'
                   'gfx950:sramecc+:xnack-/256cu 2 3 0 19 256   # config 2\n'
                   'gfx950:sramecc+:xnack-/256cu 4 4 -1 17 128\n')
    core.tune_load(str(src))
    assert core.tune_count() == 2
    out = tmp_path / 'out.txt'
    core.tune_save(str(out))
    rows = sorted(tuple(l.split()) for l in out.read_text().splitlines() if l and not l.startswith('#'))
    assert rows == sorted([('gfx950:sramecc+:xnack-/256cu', '2', '3', '0', '19', '256'),
                           ('gfx950:sramecc+:xnack-/256cu', '4', '4', '-1', '17', '128')])
    for bad, line in (('x 2 3 0 19 96\n', 1), ('ok 1 2 -1 10 64\nx 2 3\n', 2), ('x 7 3 0 19 64\n', 1)):
        f = tmp_path / 'bad.txt'
        f.write_text(bad)
        with pytest.raises(FestungError, match='bad.txt:%d' % line):
            core.tune_load(str(f))
        assert core.tune_count() == 2
    with pytest.raises(FestungError, match='cannot open'):
        core.tune_load(str(tmp_path / 'missing.txt'))
    core.tune_clear()
    assert core.tune_count() == 0


def test_visualize_grid(tmp_path):
    """core/visualizeGrid.m: every label at its place, checked by geometry --
    a local edge number n lies inside its triangle next to the midpoint of the
    edge opposite local vertex n, a local vertex number next to that vertex,
    global edge numbers on the edge midpoints (V0E), triangle numbers at the
    centroids, boundary ids outside the domain -- and the image is written."""
    from festung_amd import output as O
    g = Grid.domainSquare(1 / 3)
    lab = O.grid_labels(g)
    kinds = {}
    for x1, x2, t, k in lab:
        kinds.setdefault(k, []).append((x1, x2, int(t)))
    nb = int(np.sum(g.idE != 0))
    assert {k: len(v) for k, v in kinds.items()} == {'local edge': 3 * g.numT, 'vertex': g.numV,
                                                     'local vertex': 3 * g.numT, 'edge': g.numE,
                                                     'triangle': g.numT, 'boundary id': nb}
    V = g.coordV[g.V0T]                                  # numT x 3 x 2
    for n in range(3):
        le = np.array([(a, b) for a, b, t in kinds['local edge'][n * g.numT:(n + 1) * g.numT]])
        assert all(t == n + 1 for _, _, t in kinds['local edge'][n * g.numT:(n + 1) * g.numT])
        mids = (V[:, (n + 1) % 3] + V[:, (n + 2) % 3]) / 2
        others = [(V[:, (m + 1) % 3] + V[:, (m + 2) % 3]) / 2 for m in range(3) if m != n]
        d0 = np.linalg.norm(le - mids, axis=1)
        assert all(np.all(d0 < np.linalg.norm(le - o, axis=1)) for o in others)
        lv = np.array([(a, b) for a, b, _ in kinds['local vertex'][n * g.numT:(n + 1) * g.numT]])
        dv = np.linalg.norm(lv - V[:, n], axis=1)
        assert all(np.all(dv < np.linalg.norm(lv - V[:, m], axis=1)) for m in range(3) if m != n)
    ev = np.array([(a, b) for a, b, _ in kinds['edge']])
    assert np.allclose(ev, g.coordV[g.V0E].mean(axis=1))
    assert [t for _, _, t in kinds['edge']] == list(range(1, g.numE + 1))
    assert np.allclose(np.array([(a, b) for a, b, _ in kinds['triangle']]), V.mean(axis=1))
    bid = np.array([(a, b) for a, b, _ in kinds['boundary id']])
    assert np.all(np.any((bid < 0) | (bid > 1), axis=1))
    fn = O.visualizeGrid(g, str(tmp_path / 'vis' / 'grid.png'))
    assert open(fn, 'rb').read(8) == b'\x89PNG\r\n\x1a\n'

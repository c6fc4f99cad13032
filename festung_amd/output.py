"""Files the reference writes: Lagrange-point visualisation in VTK (.vtu)
and Tecplot (.plt) text formats, and hotstart dumps (.mat).

* ``projectDataDisc2DataLagr``  core/projectDataDisc2DataLagr.m:52-72
* ``visualizeDataLagr``         core/visualizeDataLagr.m:76-234 (triangles)
* ``outputHotstart`` / ``readHotstart``  core/outputHotstart.m, core/readHotstart.m
* ``visualizeGrid``             core/visualizeGrid.m:52-78 (an image file instead of a figure)

The file layouts (headers, node order of quadratic cells, number formats)
are the reference's, so its post-processing tools read these files as they
read its own.  All of this is host work on the state fetched from the device.
"""
import os

import numpy as np

from ._lib import check, lib, ptr

# Lagrange points on the reference triangle, core/projectDataDisc2DataLagr.m:59-65
_LAGR = {0: ([1 / 3], [1 / 3]),
         1: ([0, 1, 0], [0, 0, 1]),
         2: ([0, 1, 0, 1 / 2, 0, 1 / 2], [0, 0, 1, 1 / 2, 1 / 2, 0])}


def phi_at_points(p, X1, X2):
    """The orthonormal basis (core/phi.m) at reference points: npts x N."""
    X1 = np.ascontiguousarray(X1, dtype=np.float64).ravel()
    X2 = np.ascontiguousarray(X2, dtype=np.float64).ravel()
    N = (p + 1) * (p + 2) // 2
    out = np.empty((len(X1), N), order='F')
    check(lib().festung_phi_at_points(int(p), len(X1), ptr(X1), ptr(X2), ptr(out)))
    return out


def projectDataDisc2DataLagr(dataDisc, p=None):
    """Values of the DG function at the Lagrange points of degree min(p, 2)
    (K x 1, 3 or 6); core/projectDataDisc2DataLagr.m:52-72."""
    dataDisc = np.asarray(dataDisc, dtype=np.float64)
    K, N = dataDisc.shape
    if p is None:
        p = int(round((np.sqrt(8 * N + 1) - 3) / 2))
    L1, L2 = _LAGR[min(p, 2)]
    ph = phi_at_points(int(round((np.sqrt(8 * N + 1) - 3) / 2)), L1, L2)   # nL x N
    out = np.zeros((K, len(L1)))
    for i in range(N):
        out += dataDisc[:, i][:, None] * ph[:, i][None, :]
    return out


def _points(g, nL):
    """Cell nodes (K * nL x 2): vertices, then for quadratic cells the edge
    midpoints in the order edge 3, 1, 2 (core/visualizeDataLagr.m:158-167)."""
    cV = g.coordV0T                     # K x 3 x 2
    if nL in (1, 3):
        P = cV
    else:
        E0T = g.E0T
        bE = g.baryE[E0T[:, [2, 0, 1]]]  # K x 3 x 2
        P = np.concatenate([cV, bE], axis=1)
    return P.reshape(-1, 2)


def _node_values(d):
    """Per-node values in cell order; quadratic nodes permuted [1,2,3,6,4,5]."""
    K, nL = d.shape
    if nL == 1:
        return np.repeat(d[:, 0], 3)
    if nL == 3:
        return d.reshape(-1)
    return d[:, [0, 1, 2, 5, 3, 4]].reshape(-1)


def _vtk(g, dataLagr, varName, fileName, tLvl, vecName):
    K, nL = dataLagr[0].shape
    numP, cid = (3, 5) if nL in (1, 3) else (6, 22)
    P = _points(g, nL if nL != 1 else 3)
    vecNames = list(vecName)
    maskVec = np.zeros(len(varName), dtype=bool)
    for vn in vecNames:
        maskVec |= np.isin(varName, vecName[vn])
    isScalar = not np.all(maskVec)
    fn = '%s.%d.vtu' % (fileName, tLvl)
    vals = [_node_values(d) for d in dataLagr]
    with open(fn, 'w') as f:
        f.write('<?xml version="1.0"?>\n')
        f.write('<VTKFile type="UnstructuredGrid" version="0.1" byte_order="LittleEndian" '
                'compressor="vtkZLibDataCompressor">\n')
        f.write('  <UnstructuredGrid>\n')
        f.write('    <Piece NumberOfPoints="%d" NumberOfCells="%d">\n' % (K * numP, K))
        f.write('      <Points>\n')
        f.write('        <DataArray type="Float32" NumberOfComponents="3" format="ascii">\n')
        np.savetxt(f, np.column_stack([P, np.zeros(len(P))]), fmt=['          %.12e', '%.12e', '%.1e'])
        f.write('        </DataArray>\n')
        f.write('      </Points>\n')
        f.write('      <Cells>\n')
        f.write('        <DataArray type="Int32" Name="connectivity" format="ascii">\n')
        f.write('           ' + ''.join('%d ' % i for i in range(K * numP)))
        f.write('\n        </DataArray>\n')
        f.write('        <DataArray type="Int32" Name="offsets" format="ascii">\n')
        np.savetxt(f, np.arange(numP, numP * K + 1, numP), fmt='           %d')
        f.write('        </DataArray>\n')
        f.write('        <DataArray type="UInt8" Name="types" format="ascii">\n')
        np.savetxt(f, np.full(K, cid), fmt='           %d')
        f.write('        </DataArray>\n')
        f.write('      </Cells>\n')
        if vecNames and isScalar:
            f.write('      <PointData Scalars="%s" Vectors="%s">\n' % (varName[0], vecNames[0]))
        elif isScalar:
            f.write('      <PointData Scalars="%s">\n' % varName[0])
        else:
            f.write('      <PointData Vectors="%s">\n' % vecNames[0])
        for vn in vecNames:
            comp = [vals[i] for i in range(len(varName)) if varName[i] in vecName[vn]]
            arr = np.column_stack(comp + [np.zeros(len(vals[0]))] * (3 - len(comp)))
            f.write('        <DataArray type="Float32" Name="%s" NumberOfComponents="3" format="ascii">\n' % vn)
            np.savetxt(f, arr, fmt=['          %.9e', '%.9e', '%.9e'])
            f.write('        </DataArray>\n')
        for i in np.flatnonzero(~maskVec):
            f.write('        <DataArray type="Float32" Name="%s" NumberOfComponents="1" format="ascii">\n'
                    % varName[i])
            np.savetxt(f, vals[i], fmt='          %.9e')
            f.write('        </DataArray>\n')
        f.write('      </PointData>\n')
        f.write('    </Piece>\n')
        f.write('  </UnstructuredGrid>\n')
        f.write('</VTKFile>\n')
    return fn


def _rows(f, v, fmt, per):
    """fprintf of a flat vector with `per` values a line (Matlab recycles the
    format: the last line may be shorter)."""
    v = np.asarray(v).ravel()
    n = len(v) // per * per
    if n:
        np.savetxt(f, v[:n].reshape(-1, per), fmt=fmt, delimiter=' ')
    if n < len(v):
        f.write(' '.join(fmt % x for x in v[n:]) + ' ')


def _tec(g, dataLagr, varName, fileName, tLvl):
    K, nL = dataLagr[0].shape
    P = _points(g, nL if nL != 1 else 3)
    if nL in (1, 3):
        numT = K
        V0T = np.arange(1, len(P) + 1)
    else:
        numT = 4 * K
        # four sub-triangles per quadratic cell, core/visualizeDataLagr.m:214-219
        b = np.arange(0, 6 * K, 6)
        V0T = np.stack([b + 1, b + 4, b + 6, b + 2, b + 5, b + 4, b + 3, b + 6, b + 5, b + 4, b + 5, b + 6])
        V0T = V0T.ravel(order='F')
    fn = '%s.%d.plt' % (fileName, tLvl)
    with open(fn, 'w') as f:
        f.write('TITLE="FESTUNG output file"\n')
        f.write('VARIABLES=X, Y' + ''.join(', "%s"' % v for v in varName) + '\n')
        f.write('ZONE T="Time=%.3e", ' % tLvl)
        f.write('N=%d, E=%d, ' % (len(P), numT))
        f.write('ET=TRIANGLE, F=FEBLOCK, ')
        f.write('SOLUTIONTIME=%.3e\n\n' % tLvl)
        _rows(f, P[:, 0], '%.12e', 5)
        f.write('\n\n')
        _rows(f, P[:, 1], '%.12e', 5)
        f.write('\n\n')
        _rows(f, np.concatenate([_node_values(d) for d in dataLagr]), '%.9e', 5)
        f.write('\n')
        _rows(f, V0T, '%d', 3)
    return fn


def visualizeDataLagr(g, dataLagr, varName, fileName, tLvl, fileTypes='vtk', vecName=None):
    """Write Lagrange-point data (K x 1, 3 or 6 per variable) as
    <fileName>.<tLvl>.vtu and / or .plt; core/visualizeDataLagr.m:76-106.
    Returns the written file names."""
    if isinstance(dataLagr, np.ndarray):
        dataLagr = [dataLagr]
    if isinstance(varName, str):
        varName = [varName]
    if isinstance(fileTypes, str):
        fileTypes = [fileTypes]
    vecName = dict(vecName or {})
    assert len(dataLagr) == len(varName), 'Number of data sets in dataLagr does not match varName'
    assert all(d.shape[0] == g.numT for d in dataLagr), 'Wrong number of elements in dataLagr'
    assert all(d.shape[1] == dataLagr[0].shape[1] for d in dataLagr), \
        'All data sets in dataLagr must have same approximation order'
    d = os.path.dirname(fileName)
    if d:
        os.makedirs(d, exist_ok=True)
    out = []
    for t in fileTypes:
        if t == 'vtk':
            out.append(_vtk(g, dataLagr, list(varName), fileName, tLvl, vecName))
        elif t == 'tec':
            out.append(_tec(g, dataLagr, list(varName), fileName, tLvl))
        else:
            raise ValueError('Unknown file type: %s' % t)
    return out


def grid_labels(g):
    """The labels core/visualizeGrid.m:52-78 draws, as (x1, x2, text, kind)
    rows: local edge numbers 1..3 inside each triangle (vertex weights 1/12,
    11/24, 11/24 with 1/12 at the opposite vertex), global vertex numbers at
    the vertices, local vertex numbers (weights 5/6, 1/12, 1/12), global edge
    numbers at the edge midpoints, triangle numbers at the centroids, and the
    boundary ids of boundary edges, moved outward by |E| / 8 along the normal.
    Numbers are 1-based, as the reference prints them."""
    cx = np.asarray(g.coordV, dtype=float)
    V0T = np.asarray(g.V0T)
    X, Y = cx[V0T, 0], cx[V0T, 1]            # numT x 3
    out = []

    def add(x1, x2, vals, kind):
        out.extend((float(a), float(b), str(int(c)), kind) for a, b, c in zip(x1, x2, vals))
    wE = np.array([[1 / 12, 11 / 24, 11 / 24], [11 / 24, 1 / 12, 11 / 24], [11 / 24, 11 / 24, 1 / 12]])
    for kE in range(3):
        add(X @ wE[:, kE], Y @ wE[:, kE], np.full(g.numT, kE + 1), 'local edge')
    add(cx[:, 0], cx[:, 1], np.arange(1, g.numV + 1), 'vertex')
    wV = np.array([[5 / 6, 1 / 12, 1 / 12], [1 / 12, 5 / 6, 1 / 12], [1 / 12, 1 / 12, 5 / 6]])
    for kV in range(3):
        add(X @ wV[:, kV], Y @ wV[:, kV], np.full(g.numT, kV + 1), 'local vertex')
    bE, bT = np.asarray(g.baryE), np.asarray(g.baryT)
    add(bE[:, 0], bE[:, 1], np.arange(1, g.numE + 1), 'edge')
    add(bT[:, 0], bT[:, 1], np.arange(1, g.numT + 1), 'triangle')
    idE = np.asarray(g.idE)
    m = idE != 0
    nu, aE = np.asarray(g.nuE), np.asarray(g.areaE)
    add(bE[m, 0] + nu[m, 0] * aE[m] / 8, bE[m, 1] + nu[m, 1] * aE[m] / 8, idE[m], 'boundary id')
    return out


def visualizeGrid(g, fileName='grid.png'):
    """core/visualizeGrid.m:52-78: the triangles (black, white background,
    equal aspect, no axes) and the labels of ``grid_labels`` in blue, written
    to ``fileName`` (any format matplotlib writes: .png, .svg, .pdf) instead
    of opened as a figure.  Returns the file name."""
    import matplotlib
    matplotlib.use('Agg')
    import matplotlib.pyplot as plt
    d = os.path.dirname(fileName)
    if d:
        os.makedirs(d, exist_ok=True)
    cx = np.asarray(g.coordV, dtype=float)
    fig, ax = plt.subplots(figsize=(8, 8), facecolor='white')
    ax.triplot(cx[:, 0], cx[:, 1], np.asarray(g.V0T), color='black', linewidth=0.6)
    for x1, x2, t, _ in grid_labels(g):
        ax.text(x1, x2, t, color='blue', fontsize=6, ha='center', va='center')
    ax.set_aspect('equal')
    ax.axis('off')
    fig.savefig(fileName, dpi=150)
    plt.close(fig)
    return fileName


def outputHotstart(filename, **variables):
    """Save variables as the struct `data` of a .mat file;
    core/outputHotstart.m:55-75."""
    import scipy.io
    if not variables:
        raise ValueError('No variable to save specified!')
    d = os.path.dirname(filename)
    if d:
        os.makedirs(d, exist_ok=True)
    if not os.path.splitext(filename)[1]:
        filename += '.mat'
    scipy.io.savemat(filename, {'data': variables})
    return filename


def readHotstart(filename):
    """The variables saved by outputHotstart (core/readHotstart.m)."""
    import scipy.io
    m = scipy.io.loadmat(filename, squeeze_me=True, struct_as_record=False)
    data = m['data']
    return {k: getattr(data, k) for k in data._fieldnames}

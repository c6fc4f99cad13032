"""CPU oracle for the explicit DG advection step (test infrastructure only).

This package is a numpy/scipy restatement of the reference's Matlab/Octave
path for ``main('advection', ...)`` on triangles.  It is the checker: only
``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` may import it.  The product (``festung_amd``) never calls it.

Each function cites the reference file:line it follows.  The restatement keeps
the reference's data layout (K x N element-major matrices, element-major KN
vectors) and its operator structure (globally assembled sparse
``kron``/``kronVec`` matrices re-built every Runge-Kutta stage, ``M \\ rhs``
block solves), so that timing it gives a "reference-structured" CPU baseline.

Index conventions differ from Matlab only in being 0-based: vertex, element
and edge ids start at 0 and "no neighbour" is -1.  Local edge / vertex numbers
n = 0, 1, 2 correspond to Matlab's 1, 2, 3.

``oracle/refcpu`` (C++, OpenMP, ``oracle.refcpu`` wraps it) restates the same
per-stage algorithm in compiled code for the timed CPU baseline; it is
checked against the numpy restatement in ``tests/test_oracle.py``.

Parity status of each module is stated in its own docstring.
"""

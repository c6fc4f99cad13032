"""MI355X-native explicit DG advection on triangles (the ``main('advection',
...)`` path of FESTUNG), FP64, hand-written gfx950 HIP kernels behind a C-ABI.

    import festung_amd
    pd = festung_amd.main('advection', p=2, hmax=2**-5, isSlopeLim=True, ...)
"""
from .core import Grid, Problem, ref_table, quadRule1D, quadRule2D  # noqa: F401
from .framework import main, getStepLists  # noqa: F401

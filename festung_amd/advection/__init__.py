"""The advection problem module on the device: the step functions of
advection/*.m for triangles, with the same names, arguments and sequence
(core/getStepLists.m:44-49), over a device-resident ``Problem``.

    pd = festung_amd.main('advection', isQuadri=False, p=2, hmax=2**-5, ...)
"""
from .steps import (configureProblem, preprocessProblem, initializeProblem, preprocessStep, solveStep,
                    postprocessStep, outputStep, postprocessProblem, preprocessSubStep, solveSubStep,
                    postprocessSubStep, rungeKuttaExplicit)

__all__ = ['configureProblem', 'preprocessProblem', 'initializeProblem', 'preprocessStep', 'solveStep',
           'postprocessStep', 'outputStep', 'postprocessProblem', 'preprocessSubStep', 'solveSubStep',
           'postprocessSubStep', 'rungeKuttaExplicit']

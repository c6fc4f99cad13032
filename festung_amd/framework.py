"""Generic solver driver: main.m:112-178, core/getStepLists.m:44-49 and
core/getStepHandles.m.  Runs the preprocess list, the step list until
``isFinished``, then the postprocess list, and records the same timings the
reference prints ("Pre-processing time", "Loop time ... per iteration")."""
import importlib
import time


class Struct(dict):
    __getattr__ = dict.__getitem__

    def __setattr__(self, k, v):
        self[k] = v


def getStepLists():
    """core/getStepLists.m:44-49."""
    preprocessList = ['configureProblem', 'preprocessProblem', 'initializeProblem']
    stepList = ['preprocessStep', 'solveStep', 'postprocessStep', 'outputStep']
    postprocessList = ['postprocessProblem']
    subStepList = ['preprocessSubStep', 'solveSubStep', 'postprocessSubStep']
    return preprocessList, stepList, postprocessList, subStepList


def getStepHandles(problemName, names):
    """core/getStepHandles.m: the functions of the problem module by name."""
    mod = importlib.import_module('festung_amd.' + problemName)
    missing = [n for n in names if not hasattr(mod, n)]
    assert not missing, 'Not all the required functions of the problem found: %s' % missing
    return {n: getattr(mod, n) for n in names}


def main(problemName, *args, verbose=False, **kw):
    """main.m:112-178: ``main('advection', struct)`` or key/value pairs."""
    assert isinstance(problemName, str) and problemName
    if len(args) == 1:
        problemData = Struct(args[0])
    elif args:
        assert len(args) % 2 == 0, 'Optional arguments must take the form "key, value"'
        problemData = Struct(zip(args[0::2], args[1::2]))
    else:
        problemData = Struct()
    problemData.update(kw)
    problemData = Struct(problemData)
    problemData.problemName = problemName
    preprocessList, stepList, postprocessList, _ = getStepLists()
    t0 = time.perf_counter()
    h = getStepHandles(problemName, preprocessList)
    for f in preprocessList:
        problemData = Struct(h[f](problemData))
    problemData.timePreprocess = time.perf_counter() - t0
    if verbose:
        print('Pre-processing time: %g seconds.' % problemData.timePreprocess)
    assert isinstance(problemData.get('isFinished'), bool), \
        'Struct "problemData" must contain a logical variable "isFinished".'
    t0 = time.perf_counter()
    h = getStepHandles(problemName, stepList)
    nStep = 0
    while not problemData.isFinished:
        nStep += 1
        for f in stepList:
            problemData = h[f](problemData, nStep)
    problemData.problem.sync()
    problemData.timeLoop = time.perf_counter() - t0
    if verbose:
        print('Loop time: %d iterations in %g seconds (on avg. %g seconds per iteration).'
              % (nStep, problemData.timeLoop, problemData.timeLoop / max(nStep, 1)))
    t0 = time.perf_counter()
    h = getStepHandles(problemName, postprocessList)
    for f in postprocessList:
        problemData = h[f](problemData)
    problemData.timePostprocess = time.perf_counter() - t0
    return problemData

#!/usr/bin/env python3
"""Throughput of the explicit DG advection step on one MI355X.

Workload (the drop-in GPU step): domainSquare hmax = 2^-9 (K = 524 288
triangles), p = 2 (N = 6), FP64, rotation velocity plus a seeded 16-mode
divergence-free perturbation (amplitude 0.05), 8 seeded Gaussian humps,
zero inflow, SSP-RK3, Taylor-linear vertex-based limiter on (dc/dt and
solution limiting), time step tau = 0.1 h / (max|u| (2p+1)).

A "step" is one full time step (3 RK stages, all kernels), run from a
captured hipGraph with every input resident in HBM.  value = K N S steps /
wall time (DoF-stage updates per second).

Also reported: the roofline of the dominant kernel (algorithmic bytes per
launch / average launch time from HIP events on the problem's stream), a
per-kernel breakdown, and the oracle (the numpy/scipy restatement of the
reference's sparse-assembly path) timed on one step of the same mesh on the
host (one thread and all threads).  The default run adds, as extra keys,
the same step at K = 1 048 352 (hmax 1/724, about 1M triangles), the
batched workload (512 problems) and the adjoint workload (K = 32 768, 400
steps, gradient cost and finite-difference agreement).

    python bench.py [--gpus 1] [--steps 100] [--warmup 10] [--no-extras]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

METRIC = 'DoF-stage updates/s at p=2 + % HBM roofline; batched & adjoint vs plain'
HBM_PEAK_GBS = 8000.0


def kernel_bytes(kind, p, K, numV, lim, ordRK, maxDeg=6, lim_frac=1.0):
    """Algorithmic HBM bytes of all launches of one kernel kind in one time
    step (each datum read or written once; neighbour and vertex gathers count
    once per element / vertex).  lim_frac: the fraction of elements the
    solution limiter changes (at p >= 2 it limits the state in place and
    stores only those)."""
    N = (p + 1) * (p + 2) // 2
    Re = p + 1
    # velocity data the stage kernel reads per element: at p = 1 .. 3 its two
    # weighted factors at the RE points of quadRule2D(2p), at p = 4 (matrix
    # cores) the DG coefficients u1, u2
    NU = 2 * {0: 0, 1: 3, 2: 6, 3: 12}[p] if p <= 3 else 2 * N
    # per stage: (reads c0, reads the RK4 accumulator, writes it, reads the
    # stage state c in the RK update: weight a1 != 0, not in classic RK4)
    rk = {1: [(0, 0, 0, 1)], 2: [(0, 0, 0, 1), (1, 0, 0, 1)], 3: [(0, 0, 0, 1), (1, 0, 0, 1), (1, 0, 0, 1)],
          4: [(1, 0, 1, 0), (1, 1, 1, 0), (1, 1, 1, 0), (0, 1, 0, 0)]}[ordRK]
    tot = 0.0
    for (c0, ai, ao, a1) in rk:
        comb = N * (c0 + ai + ao)    # RK-combination operands besides c and the output
        if kind == 'stage':
            # edge flux factors |E| w_r u.nu: 3 Re; geometry: 1 / det at p <= 3 with a DG
            # output (the rest of it is in the velocity weights and flux factors), B (4)
            # at p = 4 and for Taylor output; nbr, flags: 16 B
            if lim and p >= 2:       # cdot in DG + centroid partials: c, velocity, flux factors, geometry, nbr, flags
                geo = 1 if p <= 3 else 4
                tot += K * (8 * (N + NU + 3 * Re + geo + N) + 4 * 4)
            else:                    # RK -> Taylor (p = 1 with the limiter) / DG
                geo = 1 if (p <= 3 and not lim) else 4
                tot += K * (8 * (N + comb + NU + 3 * Re + geo + N) + 16)
        elif kind == 'vertex' and lim:
            nv = 2 if p >= 2 else 1
            tot += nv * (4 * maxDeg * numV + 8 * K + 16 * numV)   # padded lists, centroids, bounds
        elif kind == 'limit_dot' and lim and p >= 2:   # cdot, c (a1 != 0), RK operands -> new state
            # (B, 32 B, is read only for the elements it limits: not counted, a lower bound)
            tot += K * (8 * (N + N * a1 + comb + N) + 12) + 16 * numV
        elif kind == 'limit' and lim:
            # p >= 2 ('linear' on DG coefficients): the state read, no geometry,
            # only the limited elements stored; p = 1: state and B read, all stored
            if p >= 2:
                tot += K * (8 * (N + N * lim_frac) + 12) + 16 * numV
            else:
                tot += K * (8 * (N + 4 + N) + 12) + 16 * numV
    return tot


def host_info():
    """Logical CPUs this process may run on, os.cpu_count() and the CPU model."""
    model = None
    try:
        with open('/proc/cpuinfo') as f:
            for line in f:
                if line.startswith('model name'):
                    model = line.split(':', 1)[1].strip()
                    break
    except OSError:
        pass
    return dict(affinity=len(os.sched_getaffinity(0)), nproc=os.cpu_count(), cpu_model=model)


def cpu_baseline(hmax, p, ordRK, lim, tc_factory):
    """The reference-structured CPU restatement (oracle/refcpu, C++ with
    OpenMP: every sub-step re-assembles G1, G2 and R as CSR matrices, forms
    their sparse sum, one SpMV, factors and solves every mass block, and
    limits through block solves and CSR vertex patches), one time step on
    the given mesh with the workload's closures, set-up excluded; once on
    one thread and once on all the process's CPUs.  Returns the
    cpu_baseline record (value = the all-threads rate)."""
    from oracle import advection as OA
    from oracle import refcpu
    tc = tc_factory()
    N = (p + 1) * (p + 2) // 2
    S = {1: 1, 2: 2, 3: 3, 4: 4}[ordRK]
    info = host_info()
    cfg = OA.Struct(testcase='custom', p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=lim, numSteps=100)
    for k in ('c0Cont', 'u1Cont', 'u2Cont', 'fCont', 'cDCont', 'gNCont'):
        cfg[k] = tc[k]
    for f in (OA.configureProblem, OA.preprocessProblem, OA.initializeProblem):
        cfg = f(cfg)
    r = refcpu.RefCpu(cfg)
    K = cfg.K
    del cfg
    runs = {}
    for label, nthr in (('1_thread', 1), ('all_threads', info['affinity'])):
        c0 = r.state()
        sec = r.steps(1, threads=nthr)
        r.set_state(c0)
        runs[label] = dict(value=K * N * S / sec, loop_s=sec, threads=nthr)
    del r
    return dict(value=runs['all_threads']['value'], unit='DoF-stage updates/s', cores=info['affinity'],
                kind='port', one_thread=runs['1_thread'], all_threads_loop_s=runs['all_threads']['loop_s'],
                host=info, sample_K=K, sample_hmax=hmax,
                sample='oracle/refcpu (C++/OpenMP reference-structured restatement: per-stage CSR assembly of '
                       'G1, G2, R, their sparse sum, SpMV, mass-block factorisations, limiter by block solves '
                       'and CSR vertex patches), 1 time step (%d stages) on domainSquare hmax=%g (K=%d), the '
                       'workload\'s p / limiter / closures; set-up excluded; %d threads (value) and 1 thread '
                       '(one_thread)' % (S, hmax, K, info['affinity']))


def run_batched(nprob, nsteps, seed0=4000):
    """The batched workload: nprob problems (K log-uniform in [1k, 64k],
    p uniform in {1, 2, 3}, SSP-RK min(p+1, 3), seeded velocities and humps),
    nsteps steps each, joined into one union problem per degree whose step
    sequences are the parallel branches of one graph; against one graph
    launch per problem.  Returns a dict of seconds, DoF-stage updates, setup
    seconds and the total K."""
    import torch
    from festung_amd import batch as BT
    from festung_amd.core import Batch
    t0 = time.perf_counter()
    specs = BT.batch_specs(nprob, seed0=seed0)
    inputs = [BT.ProblemInputs(s, device='cuda') for s in specs]
    bp = BT.BatchedProblems(inputs)
    singles = [BT.make_single(inp) for inp in inputs]
    sb = [Batch([s]) for s in singles]
    # warm-up: graph capture + one run of each form
    bp.step(nsteps)
    bp.sync()
    for b, inp in zip(sb, inputs):
        b.step(nsteps, [inp.tau])
        b.sync()
    setup = time.perf_counter() - t0
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    bp.step(nsteps)
    bp.sync()
    t_batched = time.perf_counter() - t0
    t0 = time.perf_counter()
    for b, inp in zip(sb, inputs):
        b.step(nsteps, [inp.tau])
        b.sync()
    t_single = time.perf_counter() - t0
    upd = sum(inp.K * (inp.p + 1) * (inp.p + 2) // 2 * inp.spec['ordRK'] for inp in inputs) * nsteps
    K_tot = sum(inp.K for inp in inputs)
    for b in sb:
        b.close()
    return dict(batched_s=t_batched, single_s=t_single, updates=upd, setup_s=setup, K_total=K_tot,
                groups=len(bp.problems))


def run_adjoint(hmax, p, nsteps, ckpt, nmodes=32, ndir=16, h=1e-6, variants=True):
    """Gradient of J = ||c(T) - c_obs||^2_M w.r.t. c0 and nmodes stream-function
    coefficients: cost against the faster plain forward run (eager or captured
    graphs), and central differences (step h) along ndir random directions."""
    import numpy as np
    import torch
    from festung_amd import adjoint as AD
    from festung_amd.advection import testcases as TC
    modes = AD.stream_modes()[:nmodes]
    m = AD.StreamModel(hmax, p, modes=modes, c0=TC.gaussian_humps(51))
    a_true = 0.05 * np.random.default_rng(52).standard_normal(nmodes)
    a = a_true + 0.005 * np.random.default_rng(53).standard_normal(nmodes)
    dim = int(np.ceil(1 / hmax))
    tau = 0.1 / (dim * 0.8 * (2 * p + 1))
    cObs = m.run(m.c0, a_true, nsteps, tau)

    # plain forward, eager and as captured graphs, timed with events on the
    # problem's stream; the faster one is the cost unit
    def time_forward(model):
        stream = torch.cuda.ExternalStream(model.prob.stream)
        model.set_velocity(a)
        model.prob.set_state(model.c0)
        model.prob.step(nsteps, tau)
        model.prob.sync()
        model.prob.set_state(model.c0)
        e0 = torch.cuda.Event(enable_timing=True); e1 = torch.cuda.Event(enable_timing=True)
        e0.record(stream)
        model.prob.step(nsteps, tau)
        e1.record(stream)
        model.prob.sync()
        return e0.elapsed_time(e1)
    eager_ms = time_forward(m)
    mg = AD.StreamModel(hmax, p, modes=modes, c0=TC.gaussian_humps(51), graph=True)
    graph_ms = time_forward(mg)
    del mg
    fwd_ms = min(eager_ms, graph_ms)
    out = {'forward_ms': fwd_ms, 'forward_eager_ms': eager_ms, 'forward_graph_ms': graph_ms}
    for c in (ckpt, 0):
        t0 = time.perf_counter()
        J, g0, ga, r = m.gradient(m.c0, a, nsteps, tau, cObs, ckpt=c)
        wall = time.perf_counter() - t0
        key = 'ckpt%d' % c if c else 'all_states'
        out[key] = {'adjoint_forward_ms': r['forward_ms'], 'adjoint_reverse_ms': r['backward_ms'],
                    'cost_vs_forward': (r['forward_ms'] + r['backward_ms']) / fwd_ms,
                    'reverse_vs_forward': r['backward_ms'] / fwd_ms, 'wall_s': wall}
    # the other sweep variants, same checkpoint interval
    for serial, kern in (((True, 'fused'), (False, 'fused'), (True, 'split'), (False, 'split'),
                          (True, 'deferred'), (False, 'deferred')) if variants else ()):
        mf = AD.StreamModel(hmax, p, modes=modes, c0=TC.gaussian_humps(51), adj_serial=serial, adj_kernel=kern)
        mf.gradient(mf.c0, a, 2, tau, cObs, ckpt=1)  # warm-up
        r = mf.gradient(mf.c0, a, nsteps, tau, cObs, ckpt=ckpt)[3]
        out['ckpt%d_%s_%s' % (ckpt, kern, 'serial' if serial else 'overlap')] = {
            'adjoint_forward_ms': r['forward_ms'], 'adjoint_reverse_ms': r['backward_ms'],
            'cost_vs_forward': (r['forward_ms'] + r['backward_ms']) / fwd_ms}
        del mf
    # binomial checkpointing with as many state snapshots as the uniform
    # schedule keeps segment checkpoints (nsteps / ckpt + 1), and with a few
    from festung_amd.core import binomial_schedule
    binom = {}
    for snaps in (nsteps // max(ckpt, 1) + 1, 4):
        mb = AD.StreamModel(hmax, p, modes=modes, c0=TC.gaussian_humps(51), adj_kernel='fused', adj_binomial=True)
        mb.gradient(mb.c0, a, 2, tau, cObs, ckpt=2)  # warm-up
        r = mb.gradient(mb.c0, a, nsteps, tau, cObs, ckpt=snaps)[3]
        adv = binomial_schedule(nsteps, snaps)[0]
        out['binomial_%d_snapshots' % snaps] = {
            'forward_steps': nsteps + adv, 'adjoint_forward_ms': r['forward_ms'],
            'adjoint_reverse_ms': r['backward_ms'], 'cost_vs_forward': (r['forward_ms'] + r['backward_ms']) / fwd_ms}
        binom[snaps] = out['binomial_%d_snapshots' % snaps]
        del mb
    # snapshots for <= 3x: the binomial schedule runs nsteps + advances(c)
    # forward steps (each step once more, with its stage states, right before
    # its transposed stages) and the reverse sweep, one after the other, so its
    # cost is at least 2 + reverse / forward whatever c
    rev_per_fwd = out['ckpt%d' % ckpt]['reverse_vs_forward'] if ckpt else out['all_states']['reverse_vs_forward']
    floor = 2.0 + rev_per_fwd
    best = None
    for c in range(2, nsteps + 1):
        adv = binomial_schedule(nsteps, c)[0]
        if 1.0 + adv / nsteps + rev_per_fwd <= 3.0:
            best = c
            break
    out['binomial'] = {'measured': binom, 'cost_floor_vs_forward': floor,
                       'snapshots_for_3x': best,
                       'note': 'predicted cost 1 + advances(c)/nsteps + reverse/forward; the fixed per-step '
                               'recomputation makes it >= %.2fx for every snapshot count' % floor
                       if best is None else 'smallest snapshot count whose predicted cost is <= 3x'}
    rng = np.random.default_rng(54)
    errs = []
    for _ in range(ndir):
        d0 = rng.standard_normal(g0.shape) * np.abs(m.c0).max()
        da = rng.standard_normal(ga.shape) * 0.05
        Jp = m.misfit(m.c0 + h * d0, a + h * da, nsteps, tau, cObs)
        Jm = m.misfit(m.c0 - h * d0, a - h * da, nsteps, tau, cObs)
        fd = (Jp - Jm) / (2 * h)
        an = float(np.sum(g0 * d0) + np.sum(ga * da))
        errs.append(abs(fd - an) / abs(an))
    out.update(J=J, K=m.prob.K, fd_max_rel_err=max(errs), fd_directions=ndir, fd_h=h, tau=tau)
    return out


def run_step(hmax, p, ordRK, lim, testcase, steps, warmup, kernel_steps, copy_mb):
    """The drop-in step workload on domainSquare(hmax): set-up through the
    step functions, the autotuner, `steps` timed steps from captured graphs
    (inputs resident), then a per-kernel timing pass (eager launches with HIP
    events at each kernel's own start and end) and, if copy_mb > 0, a device
    copy of copy_mb MiB per buffer for the attainable bandwidth."""
    import numpy as np
    import torch
    from festung_amd.advection import steps as S
    t_setup = time.perf_counter()
    pd = S.Struct(testcase=testcase, p=p, hmax=hmax, ordRK=ordRK, isSlopeLim=lim, numSteps=steps, useGraph=True)
    pd = S.configureProblem(pd)
    # tau = CFL 0.1 h / (max|u| (2p+1)), max|u| over the vertices
    g_dim = int(np.ceil(1.0 / hmax))
    pd.g = pd.generateGridData(pd.hmax)
    xy = pd.g.coordV
    umax = float(np.max(np.hypot(pd.u1Cont(0, xy[:, 0], xy[:, 1]), pd.u2Cont(0, xy[:, 0], xy[:, 1]))))
    tau = 0.1 * (1.0 / g_dim) / (umax * (2 * p + 1))
    pd.tEnd = tau * steps
    pd.generateGridData = lambda h: pd.g
    pd = S.preprocessProblem(pd)
    pd = S.initializeProblem(pd)
    pd = S.solveStep(pd, 1)          # first step sets the device velocity
    prob = pd.problem
    K, N, numV = prob.K, prob.N, pd.g.numV
    maxDeg = int(np.max(np.diff(pd.g.v2tPtr)))
    nstage = {1: 1, 2: 2, 3: 3, 4: 4}[ordRK]
    # elements per workgroup: the autotuner times 64 / 128 / 256 and keeps the fastest
    best_bs, tune_ms = prob.autotune(tau, 10)
    prob.step(warmup, tau)
    prob.sync()
    torch.cuda.synchronize()
    t_setup = time.perf_counter() - t_setup

    # timed region: `steps` steps from the captured graph, events on the problem's stream
    stream = torch.cuda.ExternalStream(prob.stream)
    e0 = torch.cuda.Event(enable_timing=True)
    e1 = torch.cuda.Event(enable_timing=True)
    torch.cuda.synchronize()
    t0 = time.perf_counter()
    e0.record(stream)
    prob.step(steps, tau)
    e1.record(stream)
    prob.sync()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t0
    ev_ms = e0.elapsed_time(e1)
    state = prob.get_state()
    if not np.all(np.isfinite(state)):
        print('non-finite state after the timed steps', file=sys.stderr)
        sys.exit(1)

    # per-kernel timing pass (eager launches bracketed by HIP events)
    prob.profile(True)
    prob.step(kernel_steps, tau)
    kt = prob.profile_read()
    prob.profile(False)
    # fraction of elements the solution limiter changed in the last stage
    # (it stores only those at p >= 2)
    lim_frac = float(np.mean(prob.limiter_info()[0] < 1.0)) if lim else 0.0
    kernels = {}
    for kind, (ms, cnt) in kt.items():
        if cnt == 0:
            continue
        per_step_bytes = kernel_bytes(kind, p, K, numV, lim, ordRK, maxDeg, lim_frac)
        launches_per_step = cnt / kernel_steps
        avg_ms = ms / cnt
        bytes_per_launch = per_step_bytes / launches_per_step if launches_per_step else 0.0
        kernels[kind] = dict(launches_per_step=launches_per_step, avg_us=1e3 * avg_ms,
                             bytes_per_launch=bytes_per_launch,
                             GBs=bytes_per_launch / (avg_ms * 1e-3) / 1e9 if avg_ms > 0 else 0.0,
                             share=0.0)
    tot_ms = sum(v['avg_us'] * v['launches_per_step'] for v in kernels.values())
    for v in kernels.values():
        v['share'] = v['avg_us'] * v['launches_per_step'] / tot_ms
    dom = max(kernels, key=lambda k: kernels[k]['share'])
    step_bytes = sum(kernel_bytes(k, p, K, numV, lim, ordRK, maxDeg, lim_frac)
                     for k in ('stage', 'vertex', 'limit_dot', 'limit'))
    secs = wall
    roof = dict(bound='hbm', kernel=dom, achieved=kernels[dom]['GBs'], peak=HBM_PEAK_GBS, unit='GB/s',
                frac=kernels[dom]['GBs'] / HBM_PEAK_GBS, traffic=None,
                step_GBs=step_bytes / (secs / steps) / 1e9,
                step_frac=step_bytes / (secs / steps) / 1e9 / HBM_PEAK_GBS)
    traffic = traffic_from_profile(dom, p, K)
    if traffic is not None:
        roof['traffic'] = traffic
    del pd, prob, state
    torch.cuda.synchronize()
    torch.cuda.empty_cache()

    if copy_mb > 0:
        # plain device copy of a buffer larger than the 256 MiB Infinity Cache:
        # the attainable read+write bandwidth on this box, and the reference
        # kernel for scaling the FETCH_SIZE / WRITE_SIZE counters
        nb = copy_mb * (1 << 20)
        src = torch.ones(nb // 8, dtype=torch.float64, device='cuda')
        dst = torch.empty_like(src)
        dst.copy_(src)
        torch.cuda.synchronize()
        c0 = torch.cuda.Event(enable_timing=True)
        c1 = torch.cuda.Event(enable_timing=True)
        c0.record()
        for _ in range(10):
            dst.copy_(src)
        c1.record()
        torch.cuda.synchronize()
        copy_GBs = 2 * nb * 10 / (c0.elapsed_time(c1) * 1e-3) / 1e9
        del src, dst
        torch.cuda.empty_cache()
        roof['copy_GBs'] = copy_GBs
        roof['frac_of_copy'] = kernels[dom]['GBs'] / copy_GBs
    return dict(K=K, N=N, stages=nstage, tau=tau, value=K * N * nstage * steps / secs,
                ms_per_step=1e3 * secs / steps, event_ms_per_step=ev_ms / steps, elem_block=best_bs,
                autotune_ms_per_step=tune_ms, setup_s=t_setup, roofline=roof, kernels=kernels,
                limited_fraction=lim_frac)


def traffic_from_profile(kernel, p, K):
    """HBM bytes per launch of the dominant kernel measured by the counter
    passes (FETCH_SIZE, WRITE_SIZE scaled by a plain copy) recorded in
    profiles/<kernel>_p<p>_K<K>_traffic.txt on a line 'bytes_per_launch: B'
    (None if there is no such summary)."""
    path = os.path.join(ROOT, 'profiles', 'stage_p%d_K%d_traffic.txt' % (p, K)) if kernel == 'stage' else None
    if not path or not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith('bytes_per_launch:'):
                return float(line.split(':', 1)[1])
    return None


def main():
    ap = argparse.ArgumentParser(description=__doc__, formatter_class=argparse.RawDescriptionHelpFormatter)
    ap.add_argument('--workload', choices=('step', 'batched', 'adjoint'), default='step')
    ap.add_argument('--ckpt', type=int, default=20)
    ap.add_argument('--nprob', type=int, default=512)
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=100)
    ap.add_argument('--warmup', type=int, default=10)
    ap.add_argument('--hmax', type=float, default=2.0 ** -9)
    ap.add_argument('--p', type=int, default=2)
    ap.add_argument('--ordRK', type=int, default=3)
    ap.add_argument('--no-limiter', action='store_true')
    ap.add_argument('--testcase', choices=('rotating_humps', 'random_discontinuous'), default='rotating_humps',
                    help='rotating_humps: rotation + 16-mode perturbation, Gaussian humps (p=2 drop-in); '
                         'random_discontinuous: 32-mode unit-speed velocity, disc/square indicators (p=4)')
    ap.add_argument('--config', choices=('c1', 'c2', 'c3'), default=None,
                    help='preset: c1 = hmax 2^-5 p=1 SSP-RK2 200 steps no limiter; c2 = the default; '
                         'c3 = hmax 2^-8 p=4 RK4 200 steps limiter, random_discontinuous')
    ap.add_argument('--kernel-steps', type=int, default=5, help='eager steps timed per kernel')
    ap.add_argument('--cpu-hmax', type=float, default=2.0 ** -9, help='mesh of the CPU sample (default: the workload mesh)')
    ap.add_argument('--no-cpu-baseline', action='store_true')
    ap.add_argument('--no-extras', action='store_true',
                    help='skip the K = 1 048 352 step, batched and adjoint measurements of the default run')
    ap.add_argument('--copy-mb', type=int, default=1024, help='size of the device copy probe (MiB per buffer)')
    a = ap.parse_args()
    if a.config == 'c1':
        a.hmax, a.p, a.ordRK, a.no_limiter, a.testcase = 2.0 ** -5, 1, 2, True, 'rotating_humps'
        a.steps = 200 if a.steps == 100 else a.steps
        a.cpu_hmax = 2.0 ** -5
    elif a.config == 'c3':
        a.hmax, a.p, a.ordRK, a.no_limiter, a.testcase = 2.0 ** -8, 4, 4, False, 'random_discontinuous'
        a.steps = 200 if a.steps == 100 else a.steps
        a.cpu_hmax = 2.0 ** -8
    if a.gpus != 1:
        print('bench.py measures one GPU only: the path runs on a single MI355X; '
              'multi-GPU sharding is not run.', file=sys.stderr)
        sys.exit(2)

    import numpy as np
    import torch
    if not torch.cuda.is_available():
        print('no GPU visible', file=sys.stderr)
        sys.exit(1)
    if a.workload == 'adjoint':
        nsteps = a.steps if a.steps != 100 else 400
        hm = a.hmax if a.hmax != 2.0 ** -9 else 2.0 ** -7
        r = run_adjoint(hm, a.p, nsteps, a.ckpt)
        print(json.dumps({
            'metric': 'adjoint gradient cost / forward cost (and FD agreement)',
            'value': r['ckpt%d' % a.ckpt]['cost_vs_forward'] if a.ckpt else r['all_states']['cost_vs_forward'],
            'unit': 'x forward', 'n_gpus': 1, 'steps': nsteps, 'warmup': 1,
            'ms_per_step': r['forward_ms'] / nsteps, 'higher_is_better': False, 'scaling': 'weak',
            'vs_baseline': None, 'dtype': 'f64', 'data': 'synthetic',
            'config': {'workload': 'adjoint: domainSquare hmax=2^%g K=%d p=%d SSP-RK3 %d steps, limiter off, '
                                   'J=||c(T)-c_obs||^2_M, grad wrt c0 and 32 stream-function coefficients, '
                                   'checkpoint every %d steps' % (np.log2(hm), r['K'], a.p, nsteps, a.ckpt)},
            'adjoint': r}))
        return
    if a.workload == 'batched':
        nsteps = a.steps if a.steps != 100 else 50
        rb = run_batched(a.nprob, nsteps)
        tb, ts, upd, setup, Ktot = rb['batched_s'], rb['single_s'], rb['updates'], rb['setup_s'], rb['K_total']
        print(json.dumps({
            'metric': 'DoF-stage updates/s, batched independent problems (one graph launch) vs one launch per problem',
            'value': upd / tb, 'unit': 'DoF-stage updates/s', 'n_gpus': 1, 'steps': nsteps, 'warmup': 1,
            'ms_per_step': 1e3 * tb / nsteps, 'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
            'dtype': 'f64', 'data': 'synthetic',
            'config': {'workload': '%d problems, domainSquare K log-uniform [1k, 64k] (total K=%d), p in {1,2,3}, '
                                   'SSP-RK min(p+1,3), %d steps, limiter off' % (a.nprob, Ktot, nsteps)},
            'batched_s': tb, 'one_launch_per_problem_s': ts, 'speedup': ts / tb, 'setup_s': setup}))
        return
    from festung_amd.advection import testcases as TC
    lim = not a.no_limiter
    p, ordRK = a.p, a.ordRK
    r = run_step(a.hmax, p, ordRK, lim, a.testcase, a.steps, a.warmup, a.kernel_steps, a.copy_mb)
    K, N, nstage = r['K'], r['N'], r['stages']
    log = lambda m: print('[bench] %s' % m, file=sys.stderr, flush=True)
    log('headline: %.4f ms/step, %s kernel %.1f us (%.3f of HBM peak); limited fraction %.4f'
        % (r['ms_per_step'], r['roofline']['kernel'], r['kernels'][r['roofline']['kernel']]['avg_us'],
           r['roofline']['frac'], r['limited_fraction']))

    extra = {}
    if not a.no_extras and a.config is None and a.hmax == 2.0 ** -9 and p == 2:
        # the same p = 2 step at about 1M triangles (hmax 1/724: K = 1 048 352),
        # outside the 256 MB Infinity Cache
        r1 = run_step(1.0 / 724, p, ordRK, lim, a.testcase, 40, 5, a.kernel_steps, 0)
        extra['step_K1048352'] = {k: r1[k] for k in ('K', 'ms_per_step', 'value', 'elem_block', 'limited_fraction')}
        extra['step_K1048352'].update(
            step_GBs=r1['roofline']['step_GBs'], step_frac=r1['roofline']['step_frac'],
            stage_GBs=r1['kernels']['stage']['GBs'], stage_frac=r1['kernels']['stage']['GBs'] / HBM_PEAK_GBS,
            stage_avg_us=r1['kernels']['stage']['avg_us'],
            kernels_us={k: v['avg_us'] for k, v in r1['kernels'].items()})
        log('K=1048352: %.4f ms/step, step %.3f / stage %.3f of HBM peak'
            % (r1['ms_per_step'], r1['roofline']['step_frac'], extra['step_K1048352']['stage_frac']))
        del r1
        # the batched workload (512 problems, one graph launch vs one launch per problem)
        rb = run_batched(a.nprob, 50)
        tb, ts, upd = rb['batched_s'], rb['single_s'], rb['updates']
        extra['batched'] = dict(problems=a.nprob, steps=50, K_total=rb['K_total'], batched_s=tb,
                                one_launch_per_problem_s=ts, speedup=ts / tb, dof_stage_updates_per_s=upd / tb,
                                setup_s=rb['setup_s'])
        log('batched: %.2fx (%.3f s vs %.3f s)' % (ts / tb, tb, ts))
        # the adjoint workload (K = 32 768, 400 SSP-RK3 steps, checkpoints every 20)
        ra = run_adjoint(2.0 ** -7, 2, 400, 20, variants=False)
        extra['adjoint'] = dict(K=ra['K'], steps=400, ckpt=20, forward_ms=ra['forward_ms'],
                                cost_vs_forward=ra['ckpt20']['cost_vs_forward'],
                                cost_vs_forward_all_states=ra['all_states']['cost_vs_forward'],
                                fd_max_rel_err=ra['fd_max_rel_err'], fd_directions=ra['fd_directions'],
                                fd_h=ra['fd_h'], binomial=ra['binomial'])
        log('adjoint: %.2fx forward, FD max rel err %.2e' % (ra['ckpt20']['cost_vs_forward'], ra['fd_max_rel_err']))

    cpu = None
    if not a.no_cpu_baseline:
        log('CPU baseline (oracle) on the host ...')
        cpu = cpu_baseline(a.cpu_hmax, p, ordRK, lim, TC.TESTCASES[a.testcase])

    out = {
        'metric': METRIC, 'value': r['value'], 'unit': 'DoF-stage updates/s', 'n_gpus': 1,
        'steps': a.steps, 'warmup': a.warmup, 'ms_per_step': r['ms_per_step'],
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f64',
        'data': 'synthetic',
        'config': {'workload': 'advection drop-in step: domainSquare hmax=2^%g K=%d p=%d N=%d %s '
                               'limiter=%s, %s'
                               % (np.log2(a.hmax), K, p, N, 'RK4' if ordRK == 4 else 'SSP-RK%d' % ordRK,
                                  'on' if lim else 'off',
                                  'rotation+16-mode stream function, 8 Gaussian humps'
                                  if a.testcase == 'rotating_humps' else
                                  '32-mode stream function (unit max speed), disc/square indicators'),
                   'K': K, 'p': p, 'N': N, 'ordRK': ordRK, 'limiter': lim, 'tau': r['tau'],
                   'stages_per_step': nstage, 'event_ms_per_step': r['event_ms_per_step'],
                   'elem_block': r['elem_block'], 'autotune_ms_per_step': r['autotune_ms_per_step'],
                   'setup_s': r['setup_s'], 'limited_fraction': r['limited_fraction']},
        'roofline': r['roofline'],
        'kernels': r['kernels'],
        'cpu_baseline': cpu,
    }
    out.update(extra)
    print(json.dumps(out))


if __name__ == '__main__':
    main()

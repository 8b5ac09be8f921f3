"""
Benchmark of the delayed-coupling network integrator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--plain NAME|none] [--variants A,B|none] [--no-grad]
                    [--no-cpu] [--no-fc]

A "step" is one integration step of the whole sweep batch (all P instances):
delayed CSR coupling + model dfun + Euler(-Maruyama) update + history write
+ tavg monitor + FP64 online FC.  Inputs are resident in HBM before the
timed region; the timed region is exactly K steps (the launches of the last
block are shortened to end on K) bracketed by stream synchronisations.
Every leg runs on its own: a leg that fails leaves its field null and its
error text under "errors", and the JSON line is still printed.

value        instance*steps/s of the batched mixed-precision sweep at N=998
             (Jansen-Rit, P=1024, delays quantised to 32 bins, bfloat16
             history of the 1/cmax-scaled sigmoid, FP32 state and sums, FP64
             FC): P * K / wall time
variants     the same sweep with a float16 history, with an 8-bit
             fixed-point history, and with the delay-binned sum over bins
             of W_d X(t - d) on the matrix cores (bfloat16 / float16 MFMA,
             float32 accumulation) instead of the CSR row gathers
plain        the same sweep through the plain FP32 kernel (exact delays,
             float32 history); vs_plain = value / plain value (wall clock),
             vs_plain_kernel the same ratio of step-kernel times
fc_err_vs_fp64
             max |FC_device - FC_fp64| over 8 instances spread over the
             sweep, after an untimed run of the workload's full length
             (10,000 steps), for the headline, the plain kernel and the
             variants, against the float64 oracle (exact delays) on the host
full_run     wall-clock rate of those full-length runs
grad         gradient evaluations/s of the FC-fit inversion workload (MPR
             N=84, P=4096, 20000 steps, checkpoints every 1000 steps replayed
             in the backward pass): forward + FC loss on the device + adjoint
             reverse sweep for all instances
roofline     the step kernel against HBM bandwidth: algorithmic bytes per
             launch (1, 2 or 4 B per edge gather + per node write, per instance
             and step, times the steps of the launch) / mean launch duration
             (HIP events on the simulation's stream); `compulsory` counts
             each distinct (source, delay) history value once per step
             instead of once per edge; `traffic` is the measured HBM bytes per
             launch from the committed rocprofv3 PMC summary for this kernel,
             workload and launch shape; `device_copy` the device's own
             streaming rates (grid-stride read and copy over 1 GiB,
             tvbx_stream_rates), the practical ceiling that traffic rate
             compares with
cpu_baseline the float32 CPU oracle (the reference's loop structure: per
             instance, sequential steps, CSR row reduction, OpenMP over
             instances), built on the host with -O3 -march=native with and
             without -ffast-math, timed on a bounded sample of the plain
             workload on all cores of this process's share
"""

import argparse
import json
import os
import re
import subprocess
import sys
import time
import traceback

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'tvb-hpc_amd'))

METRIC = ('instance·steps/s at N=998 sweep vs plain kernel; FC err vs '
          'FP64; grad evals/s')
HBM_PEAK_GBS = 8000.0


_T0 = time.perf_counter()


def log(*a):
    print('[%6.1f s]' % (time.perf_counter() - _T0), *a, file=sys.stderr,
          flush=True)


def step_kernel(sim):
    return 'tvbx_step_blocked' if sim.blocked else 'tvbx_step'


def traffic_from_profile(kernel, workload, launch):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    PMC summary (profiles/<kernel>_<workload>_pmc.txt), if it was measured
    with this launch shape."""
    path = os.path.join(ROOT, 'profiles', '%s_%s_pmc.txt' % (kernel, workload))
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r'hbm_bytes_per_dispatch\s*=\s*([0-9.eE+]+)', txt)
    la = re.search(r'^launch\s*=\s*(\{.*\})\s*$', txt, re.M)
    if not m or not la or json.loads(la.group(1)) != launch:
        return None
    return float(m.group(1))


def device_copy_rate(mib=1024):
    """The device's practical streaming rates, the ceiling the step
    kernels' measured traffic compares with: the best of a grid-stride read
    and a grid-stride copy of `mib` MiB (libtvbx's tvbx_stream_rates; the
    copy counts its reads and writes)."""
    from tvb_hpc.hip import Context
    read, copy = Context.default().stream_rates(mib)
    return {'read_GBs': read, 'GBs': copy,
            'what': 'tvbx_stream_rates over %d MiB: grid-stride read of '
                    '16-byte words, and grid-stride copy (read + write), '
                    'best of 3 grid sizes x 10 calls' % mib}


def distinct_sources(csr):
    """Distinct (source, delay) pairs per delay group: the history values a
    step must read at least once per instance (the compulsory gathers)."""
    import numpy as np
    col = np.asarray(csr.col, np.int64)
    out = []
    for g in range(csr.n_group):
        d = np.asarray(csr.delays[g], np.int64)
        out.append(int(np.unique(d * (1 << 32) + col).size))
    return max(out) if out else 0


def time_sweep(cfg, launch, K, W, graph_window, profile_steps):
    """W untimed steps, then K timed steps (wall clock between stream
    synchronisations), then `profile_steps` eager launches timed one by one
    with HIP events on the simulation's stream.  Returns a result dict."""
    sim = cfg.simulation(**launch)
    N, P, nnz = cfg.csr.nnode, cfg.n_inst, cfg.csr.nnz
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    if getattr(cfg.scheme, 'sigma', None) is not None:
        sim.set_noise(seed=cfg.seed + 1)
    sim.set_exec(graph=graph_window > 0, window=max(graph_window, 1),
                 profile=False)
    sim.run(W)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(K)
    sim.sync()
    el = time.perf_counter() - t0
    if sim.flags().any():
        raise RuntimeError('non-finite state in the timed run')
    # kernel time: direct launches bracketed by events on the simulation's
    # stream (the timed region above replays graph windows)
    sim.set_exec(graph=False, profile=True)
    sim.timing(reset=True)
    sim.run(profile_steps)
    sim.sync()
    ms_kern, launches = sim.timing(reset=True)
    esz = {'f32': 4, 'bf16': 2, 'f16': 2, 'u8': 1}[cfg.ring_dtype]
    C = sim.spec.C
    bytes_inst_step = esz * (nnz + N * C)
    comp_inst_step = esz * (distinct_sources(cfg.csr) * C + N * C)
    steps_per_launch = profile_steps / max(launches, 1)
    launch_ms = ms_kern / max(launches, 1)
    per_s = steps_per_launch / (launch_ms * 1e-3)
    achieved = bytes_inst_step * P * per_s / 1e9
    comp = comp_inst_step * P * per_s / 1e9
    kernel = step_kernel(sim)
    del sim
    return dict(value=P * K / el, ms_per_step=el / K * 1e3, steps=K,
                roofline={
                    'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                    'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                    'traffic': None, 'kernel': kernel,
                    'bytes_per_instance_step': bytes_inst_step,
                    'compulsory': {
                        'bytes_per_instance_step': comp_inst_step,
                        'achieved': comp, 'frac': comp / HBM_PEAK_GBS,
                        'what': 'distinct (source, delay) history values '
                                'read once + one ring write per node'},
                    'steps_per_launch': steps_per_launch,
                    'launch_ms': launch_ms, 'launches': launches,
                    'kernel_ms_per_step': launch_ms / steps_per_launch})


def fc_run(cfg, launch, n_steps, fc_idx, graph_window):
    """Untimed full-length run of n_steps from the initial state; returns the
    FC of instances fc_idx (FP64 co-moments on the device) and the wall
    time of the steps."""
    import numpy as np
    sim = cfg.simulation(**launch)
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    if getattr(cfg.scheme, 'sigma', None) is not None:
        sim.set_noise(seed=cfg.seed + 1)
    sim.set_exec(graph=graph_window > 0, window=max(graph_window, 1),
                 profile=False)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(n_steps)
    sim.sync()
    el = time.perf_counter() - t0
    if sim.flags().any():
        raise RuntimeError('non-finite state in the full-length run')
    fc = np.stack([sim.fc(int(p), 1)[0] for p in fc_idx])
    ns = sim.n_fc_samples
    del sim
    return fc, ns, el


def time_gradient(n_evals=2, window=1000):
    """Gradient evaluations/s of the FC-fit workload, end to end on the
    device: checkpointed forward of all instances (a checkpoint every
    `window` steps, replayed in the backward pass; window = 0 records every
    step instead), FC loss against a target FC (from a run at G = 0.5) on
    the device (autograd.fc_loss_device), and the adjoint."""
    import numpy as np
    import torch
    from tvb_hpc import autograd as AG
    from tvb_hpc import autotune, configs
    cfg = configs.ALL['mpr84']()
    la = dict(autotune.best_launch(cfg))
    T, K = cfg.n_steps, cfg.tavg_period
    N, P = cfg.csr.nnode, cfg.n_inst
    sim = cfg.simulation(adjoint=True, fc=False, sample_slots=T // K, **la)
    x0 = cfg.initial_state()
    dev = torch.device('cuda:0')
    G = torch.tensor(np.asarray(cfg.params['c:a'], np.float32), device=dev)
    eta = torch.full((N, P), float(cfg.model.const['eta']), device=dev)
    with torch.no_grad():
        s_t = AG.integrate(sim, {'c:a': torch.full_like(G, 0.5),
                                 'm:eta': eta}, x0, T, window, seed=5)
        fct = AG.fc(s_t).mean(dim=0)
        del s_t
    times = []
    for e in range(n_evals + 1):
        Gt = G.clone().requires_grad_(True)
        et = eta.clone().requires_grad_(True)
        torch.cuda.synchronize()
        t = time.perf_counter()
        s = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, window,
                         seed=5)
        loss = AG.fc_loss_device(s, fct).sum()
        loss.backward()
        torch.cuda.synchronize()
        times.append(time.perf_counter() - t)
        assert torch.isfinite(Gt.grad).all() and torch.isfinite(et.grad).all()
        del s, loss
    ms = float(np.median(times[1:])) * 1e3
    del sim
    return {'workload': '%s: %s, %d steps, %s' %
                        (cfg.name, cfg.description, T,
                         'every step recorded in HBM (no replay)'
                         if window == 0 else
                         'checkpoint every %d' % window),
            'evals_per_s': 1e3 / ms, 'ms_per_eval': ms,
            'instance_gradients_per_s': P * 1e3 / ms,
            'params': ['c:a (P)', 'm:eta (N x P)'], 'launch': la,
            'note': 'one evaluation = checkpointed forward, FC loss on the '
                    'device (FP64 co-moments, float32 gradient), replay of '
                    'each window recording x(n) and the coupling sums, '
                    'reverse sweep, for all instances'}


def host_info():
    """Cores this process may run on, the machine's logical CPUs and the
    CPU model (from lscpu)."""
    try:
        avail = len(os.sched_getaffinity(0))
    except AttributeError:
        avail = os.cpu_count() or 1
    model = None
    try:
        out = subprocess.run(['lscpu'], stdout=subprocess.PIPE, text=True,
                             timeout=20).stdout
        m = re.search(r'^Model name:\s*(.+)$', out, re.M)
        model = m.group(1).strip() if m else None
    except Exception:
        pass
    omp = os.environ.get('OMP_NUM_THREADS')
    return {'nproc': avail, 'machine_cpus': os.cpu_count(),
            'omp_num_threads': int(omp) if omp and omp.isdigit() else None,
            'cpu_model': model}


def cpu_threads(info):
    """All cores of this process's share: the affinity set, capped by
    OMP_NUM_THREADS when the machine sets it (a shared host's quota)."""
    n = info['nproc']
    if info['omp_num_threads']:
        n = min(n, info['omp_num_threads'])
    return max(1, n)


def cpu_baseline(plain_cfg, n_inst, n_steps, threads):
    """The float32 oracle (the reference's loop structure: per instance,
    sequential steps, CSR row reduction, OpenMP over instances) built on
    this host with -O3 -march=native, without and with -ffast-math
    (examples/hcp100.py:65-72 builds its C path with both), timed on a
    bounded sample of the plain workload on `threads` cores."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import numpy as np
    import oracle as O
    seed = None
    if getattr(plain_cfg.scheme, 'sigma', None) is not None:
        seed = plain_cfg.seed + 1
    idx = np.arange(n_inst)
    res = {}
    for variant in ('native', 'native_fastmath'):
        lib = O.load_library(O.build_native(
            fast_math=variant.endswith('fastmath')))
        ref = O.for_config(plain_cfg, idx, np.float32,
                           n_threads=threads, library=lib)
        ref.run(5, seed=seed)
        t = time.perf_counter()
        ref.run(n_steps, seed=seed)
        el = time.perf_counter() - t
        res[variant] = (n_inst * n_steps / el, el)
    best = max(res, key=lambda k: res[k][0])
    info = host_info()
    return {'value': res[best][0], 'unit': 'instance*steps/s',
            'cores': threads, 'kind': 'port',
            'sample': '%d instances x %d steps of %s (float32 oracle, '
                      '-O3 -march=native -fopenmp%s, %d threads, %.1f s)'
                      % (n_inst, n_steps, plain_cfg.name,
                         ' -ffast-math' if best.endswith('fastmath') else '',
                         threads, res[best][1]),
            'no_fast_math': res['native'][0],
            'fast_math': res['native_fastmath'][0],
            'nproc': info['nproc'], 'machine_cpus': info['machine_cpus'],
            'cpu_model': info['cpu_model']}


def fc64_reference(plain_cfg, fc_idx, n_steps, threads):
    """FC of instances fc_idx from the float64 oracle over n_steps (exact
    delays), numpy.corrcoef of its monitor samples."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import numpy as np
    import oracle as O
    seed = None
    if getattr(plain_cfg.scheme, 'sigma', None) is not None:
        seed = plain_cfg.seed + 1
    ref = O.for_config(plain_cfg, np.asarray(fc_idx), np.float64,
                       n_threads=threads)
    t = time.perf_counter()
    s = ref.run(n_steps, seed=seed)['samples']
    log('float64 oracle FC reference: %d instances x %d steps, %.1f s' %
        (len(fc_idx), n_steps, time.perf_counter() - t))
    return np.stack([np.corrcoef(s[:, :, k].T) for k in range(len(fc_idx))])


RMAP_WORKLOADS = {
    # surface vertices (the reference's 515,056-vertex surface), regions,
    # state variables, instances
    'rmap515k': (515056, 400, 2, 64),
}


def bench_rmap(args):
    """Region mapping of a 515k-vertex surface: to_avg (region sums) and
    from_avg (vertex gather), device-resident arrays, K calls each."""
    import ctypes
    import numpy as np
    import torch
    from tvb_hpc import hip
    nnode, nroi, nvar, P = RMAP_WORKLOADS[args.workload]
    K = args.steps or 200
    W = args.warmup
    rng = np.random.default_rng(5)
    rmap = rng.integers(0, nroi, size=nnode).astype(np.int32)
    node_h = rng.standard_normal((nnode, nvar, P)).astype(np.float32)
    m = hip.RegionMap(rmap, nroi)
    L = hip.lib()
    node = torch.from_numpy(node_h).cuda()
    roi = torch.empty((nroi, nvar, P), dtype=torch.float32, device='cuda')
    back = torch.empty_like(node)
    torch.cuda.synchronize()
    vp = lambda t: ctypes.c_void_p(t.data_ptr())

    def run(fn, src, dst, n):
        for _ in range(n):
            hip.check(fn(m.h, nvar, P, vp(src), vp(dst)))
        m.ctx.sync()

    res = {}
    for name, fn, src, dst in (('to_avg', L.tvbx_rmap_to_avg, node, roi),
                               ('from_avg', L.tvbx_rmap_from_avg, roi, back)):
        run(fn, src, dst, W)
        t0 = time.perf_counter()
        run(fn, src, dst, K)
        res[name] = (time.perf_counter() - t0) / K
    # algorithmic bytes: every vertex value read (to_avg) or written
    # (from_avg) once, plus the region values and the index arrays
    vals = nnode * nvar * P
    b_to = 4 * vals + 4 * nroi * nvar * P + 4 * nnode * nvar
    b_from = 4 * vals + 4 * nnode
    ach = b_to / res['to_avg'] / 1e9
    out = {
        'metric': 'region map vertex values/s (to_avg)',
        'value': vals / res['to_avg'], 'unit': 'vertex*var*instance/s',
        'n_gpus': 1, 'steps': K, 'warmup': W,
        'ms_per_step': res['to_avg'] * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': '%s: %d vertices, %d regions, %d variables, '
                               '%d instances' % (args.workload, nnode, nroi,
                                                 nvar, P)},
        'roofline': {'bound': 'hbm', 'achieved': ach, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': ach / HBM_PEAK_GBS,
                     'traffic': None, 'kernel': 'k_rmap_to_avg',
                     'bytes_per_call': b_to},
        'from_avg': {'ms_per_call': res['from_avg'] * 1e3,
                     'achieved_gbs': b_from / res['from_avg'] / 1e9,
                     'kernel': 'k_rmap_from_avg'},
    }
    roi_h = roi.cpu().numpy()
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import oracle as O
    t0 = time.perf_counter()
    ref = O.rmap_to_avg(rmap, node_h, nroi)
    el = time.perf_counter() - t0
    out['exact_vs_oracle'] = bool(np.array_equal(roi_h, ref))
    if not args.no_cpu:
        out['cpu_baseline'] = {'value': vals / el,
                               'unit': 'vertex*var*instance/s', 'cores': 1,
                               'kind': 'port',
                               'sample': 'the whole workload once (numpy '
                                         'unbuffered add.at, one thread)'}
    print(json.dumps(out), flush=True)


LOGP_WORKLOADS = {'logp_mpr84': 'mpr84'}


def bench_logp(args):
    """Euler-Maruyama log-likelihood of a device-resident MPR trajectory
    (N=84, P=4096, teacher forcing) and its parameter gradients."""
    import numpy as np
    import torch
    from tvb_hpc import autotune, configs
    cfg = configs.ALL[LOGP_WORKLOADS[args.workload]]()
    T = args.steps or 1000
    N, P = cfg.csr.nnode, cfg.n_inst
    la = dict(autotune.best_launch(cfg))
    sim = cfg.simulation(adjoint=True, **la)
    S = sim.spec.S
    x0 = cfg.initial_state()
    frames = torch.empty((T + 1, S, N, P), dtype=torch.float32, device='cuda')
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.set_noise(seed=cfg.seed + 1)
    sim.state_into(frames[0].data_ptr())
    for k in range(T):
        sim.run(1)
        sim.state_into(frames[k + 1].data_ptr())
    sim.ctx.sync()
    res = {}
    for mode in ('logp', 'grad'):
        for rep in range(1 + max(1, args.warmup // 100)):
            sim.step = 0
            sim.set_state(x0)
            sim.init_history('zeros')
            t0 = time.perf_counter()
            out = sim.logp(frames, grad=(mode == 'grad'))
            el = time.perf_counter() - t0
        res[mode] = el
    lp = out[0]
    nnz = cfg.csr.nnz
    bytes_step = 4 * nnz * P + 4 * S * N * P + 4 * N * P + 16 * N * P
    ach = bytes_step * T / res['logp'] / 1e9
    out = {
        'metric': 'EM log-likelihood transition*instances/s',
        'value': P * T / res['logp'], 'unit': 'transition*instance/s',
        'n_gpus': 1, 'steps': T, 'warmup': args.warmup,
        'ms_per_step': res['logp'] / T * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None,
        'dtype': 'f32 state, compensated f32 sums', 'data': 'synthetic',
        'config': {'workload': '%s: EM log-likelihood of a %d-transition '
                               'trajectory of %s' % (args.workload, T,
                                                     cfg.description),
                   'launch': la},
        'roofline': {'bound': 'hbm', 'achieved': ach, 'peak': HBM_PEAK_GBS,
                     'unit': 'GB/s', 'frac': ach / HBM_PEAK_GBS,
                     'traffic': None, 'kernel': 'tvbx_step (forced)',
                     'bytes_per_step': bytes_step,
                     'note': 'wall time of the whole call'},
        'with_gradients': {'value': P * T / res['grad'],
                           'ms_per_step': res['grad'] / T * 1e3},
        'logp_mean': float(np.mean(lp)),
    }
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        n_c = min(args.cpu_inst, P)
        idx = np.arange(n_c)
        fr = frames[:, :, :, :n_c].cpu().numpy()
        threads = cpu_threads(host_info())
        csr = cfg.csr
        o = O.OracleSim('MontbrioPazoRoxin', 'Linear', N, csr.row, csr.col,
                        csr.weight, csr.delays, n_c, csr.max_delay + 2,
                        cfg.scheme.dt, scheme='em',
                        cparams={'a': np.asarray(cfg.params['c:a'])[idx]},
                        sigma=cfg.scheme.sigma, dtype=np.float32,
                        n_threads=threads)
        o.set_state(fr[0])
        o.init_history('zeros')
        t0 = time.perf_counter()
        ref = o.logp(fr)
        el = time.perf_counter() - t0
        out['cpu_baseline'] = {
            'value': n_c * T / el, 'unit': 'transition*instance/s',
            'cores': threads, 'kind': 'port',
            'sample': '%d instances x %d transitions (float32 oracle)'
                      % (n_c, T)}
        out['max_rel_err_vs_oracle'] = float(
            np.abs(lp[:n_c] - ref).max() / np.abs(ref).max())
    print(json.dumps(out), flush=True)


def bench_batch(args):
    """The uneven-size sweep: subjects with different connectomes in one
    launch per block of steps."""
    import numpy as np
    from tvb_hpc import autotune, configs
    cfg = configs.BATCHES[args.workload]()
    K = args.steps or cfg.n_steps
    W = args.warmup
    la = dict(autotune.best_batch_launch(cfg))
    log('batch', cfg.name, 'subjects', len(cfg.subjects), 'P', cfg.n_inst,
        'launch', la)
    gw = 0 if args.no_graph else args.graph_window
    b = cfg.batch(**la)
    for k, x in enumerate(cfg.initial_states()):
        b.set_state(k, x)
    b.init_history('zeros')
    b.set_noise(seed=cfg.seed + 1)
    b.set_exec(graph=gw > 0, window=max(gw, 1))
    b.run(W)
    b.sync()
    t0 = time.perf_counter()
    b.run(K)
    b.sync()
    el = time.perf_counter() - t0
    assert not b.flags().any(), 'non-finite state in the benchmark run'
    # launch time: back-to-back eager launches of one block of steps each
    L = la['steps_per_launch']
    n_l = max(1, args.profile_steps // L)
    b.set_exec(graph=False)
    b.sync()
    t1 = time.perf_counter()
    b.run(n_l * L)
    b.sync()
    launch_ms = (time.perf_counter() - t1) / n_l * 1e3
    nnz = sum(c.nnz * p for c, p in cfg.subjects)
    nodes = sum(c.nnode * p for c, p in cfg.subjects)
    bytes_step = 4 * (nnz + nodes)      # gathers + ring writes, all subjects
    achieved = bytes_step * L / (launch_ms * 1e-3) / 1e9
    Ns = sorted({c.nnode for c, _ in cfg.subjects})
    out = {
        'metric': METRIC, 'value': cfg.n_inst * K / el,
        'unit': 'instance*steps/s', 'n_gpus': 1, 'steps': K, 'warmup': W,
        'ms_per_step': el / K * 1e3, 'higher_is_better': True,
        'scaling': 'weak', 'vs_baseline': None, 'dtype': 'f32',
        'data': 'synthetic',
        'config': {'workload': '%s: %s' % (cfg.name, cfg.description),
                   'reading': '2048 instances = %d subjects x %d instances: '
                              'each subject its own connectome (N drawn '
                              'from the set, own seed, own conduction '
                              'speed U(1, 10) m/s), its instances sweeping '
                              'G, J_NMDA and sigma; not one connectome per '
                              'instance' % (len(cfg.subjects),
                                            cfg.subjects[0][1]),
                   'subjects': len(cfg.subjects), 'n_inst': cfg.n_inst,
                   'node_counts': Ns, 'nnz_total': sum(
                       c.nnz for c, _ in cfg.subjects),
                   'dt_ms': cfg.scheme.dt, 'launch': la,
                   'graph_window': gw},
        'roofline': {'bound': 'hbm', 'achieved': achieved,
                     'peak': HBM_PEAK_GBS, 'unit': 'GB/s',
                     'frac': achieved / HBM_PEAK_GBS,
                     'traffic': traffic_from_profile('tvbx_step_batch',
                                                     args.workload, la),
                     'kernel': 'tvbx_step_batch',
                     'bytes_per_step': bytes_step, 'steps_per_launch': L,
                     'launch_ms': launch_ms,
                     'launch_time': 'wall time of back-to-back launches'},
    }
    del b
    if not args.no_cpu:
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        threads = cpu_threads(host_info())
        # float32 oracle on a bounded sample: the first instances of every
        # subject
        t = 0.0
        work = 0
        n_s = args.cpu_steps
        off = 0
        for k, (csr, p) in enumerate(cfg.subjects):
            idx = np.arange(off, off + min(p, threads))
            off += p
            prm = {'J': np.asarray(cfg.params['m:J'])[idx]}
            ref = O.OracleSim('ReducedWongWang', 'Linear', csr.nnode, csr.row,
                              csr.col, csr.weight, csr.delays, len(idx),
                              csr.max_delay + 2, cfg.scheme.dt, scheme='em',
                              params=prm, cparams={'a': np.asarray(
                                  cfg.params['c:a'])[idx]},
                              sigma=np.asarray(cfg.params['sigma'])[idx],
                              dtype=np.float32, n_threads=threads)
            ref.set_state(cfg.model.initial_state(
                csr.nnode, len(idx), np.random.default_rng(k)))
            ref.init_history('zeros')
            t1 = time.perf_counter()
            ref.run(n_s, seed=1)
            t += time.perf_counter() - t1
            work += len(idx) * n_s
        out['cpu_baseline'] = {
            'value': work / t, 'unit': 'instance*steps/s', 'cores': threads,
            'kind': 'port',
            'sample': '%d instances of each of the %d subjects x %d steps on '
                      '%d threads (float32 oracle, %.1f s)' % (
                          threads, len(cfg.subjects), n_s, threads, t)}
    print(json.dumps(out), flush=True)


def launch_of(cfg, arg):
    from tvb_hpc import autotune
    if arg:
        ipl, w, nb, spl = [int(v) for v in arg.split(',')]
        return dict(ipl=ipl, waves=w, node_block=nb, steps_per_launch=spl)
    return dict(autotune.best_launch(cfg))


class Legs:
    """Runs the benchmark's legs one by one; a leg that raises leaves its
    field null and its error text in ``errors``, and the line still
    prints."""

    def __init__(self):
        self.errors = {}

    def __call__(self, name, fn, *a, **kw):
        try:
            return fn(*a, **kw)
        except Exception as e:  # noqa: BLE001 -- reported, not hidden
            self.errors[name] = '%s: %s' % (type(e).__name__, e)
            log('leg %s failed:\n%s' % (name, traceback.format_exc()))
            return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None,
                    help='timed steps (default: the workload\'s run length)')
    ap.add_argument('--warmup', type=int, default=200)
    ap.add_argument('--workload', default='jansenrit998_bf16')
    ap.add_argument('--variants',
                    default='jansenrit998_f16,jansenrit998_u8,'
                            'jansenrit998_bf16_mfma,jansenrit998_f16_mfma',
                    help='other mixed-precision workloads timed alongside '
                         "(comma separated, or 'none')")
    ap.add_argument('--plain', default='jansenrit998',
                    help="plain-kernel workload to compare with, or 'none'")
    ap.add_argument('--no-grad', action='store_true')
    ap.add_argument('--grad-no-replay', action='store_true',
                    help='also time gradients with every step recorded')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--no-fc', action='store_true',
                    help='skip the full-length FC error leg')
    ap.add_argument('--fc-instances', type=int, default=8,
                    help='instances whose FC is checked against float64')
    ap.add_argument('--fc-steps', type=int, default=None,
                    help='length of the FC error leg (default: the '
                         'workload\'s run length)')
    ap.add_argument('--cpu-inst', type=int, default=64)
    ap.add_argument('--cpu-steps', type=int, default=1000)
    ap.add_argument('--graph-window', type=int, default=100,
                    help='steps per captured graph window')
    ap.add_argument('--no-graph', action='store_true')
    ap.add_argument('--profile-steps', type=int, default=320)
    ap.add_argument('--launch', default=None,
                    help='ipl,waves,node_block,steps_per_launch '
                         '(default: tuned)')
    ap.add_argument('--copy-calibration', type=int, default=0,
                    metavar='MIB',
                    help='only run 10 elementwise copies of MIB MiB (the '
                         'known-size kernel that scales PMC byte counters)')
    args = ap.parse_args()
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        x = torch.ones(n, device='cuda:0')
        y = torch.empty_like(x)
        for _ in range(10):
            torch.mul(x, 1.0, out=y)   # reads and writes 4 n bytes
        torch.cuda.synchronize()
        log('copied %d MiB x 10' % args.copy_calibration)
        return
    if args.gpus != 1:
        print('bench.py measures one GPU: the integrator shards sweep '
              'instances across GPUs without communication, and this build '
              'runs on one GPU only (--gpus %d requested)' % args.gpus,
              file=sys.stderr)
        sys.exit(2)

    import numpy as np
    from tvb_hpc import configs

    if args.workload in configs.BATCHES:
        return bench_batch(args)
    if args.workload in RMAP_WORKLOADS:
        return bench_rmap(args)
    if args.workload in LOGP_WORKLOADS:
        return bench_logp(args)
    legs = Legs()
    gw = 0 if args.no_graph else args.graph_window
    if not args.no_cpu:
        # the CPU baseline's -march=native builds compile on this host
        # (one g++ each) while the GPU legs run
        import threading
        sys.path.insert(0, os.path.join(ROOT, 'oracle'))
        import oracle as O
        for fm in (False, True):
            threading.Thread(target=lambda fm=fm: legs(
                'cpu_build', O.build_native, fast_math=fm),
                daemon=True).start()
    cfg = configs.ALL[args.workload]()
    K = args.steps or cfg.n_steps
    W = args.warmup
    N, P, nnz = cfg.csr.nnode, cfg.n_inst, cfg.csr.nnz
    launch = launch_of(cfg, args.launch)
    log('workload', cfg.name, 'N', N, 'P', P, 'nnz', nnz, 'launch', launch)
    dtype = 'f32' if cfg.ring_dtype == 'f32' else \
        '%s history, f32 state and sums' % cfg.ring_dtype
    out = {
        'metric': METRIC, 'value': None, 'unit': 'instance*steps/s',
        'n_gpus': 1, 'steps': K, 'warmup': W, 'ms_per_step': None,
        'higher_is_better': True, 'scaling': 'weak', 'vs_baseline': None,
        'dtype': dtype, 'data': 'synthetic',
        'config': {
            'workload': '%s: %s' % (cfg.name, cfg.description),
            'n_node': N, 'n_inst': P, 'nnz': nnz,
            'delay_bins': cfg.delay_bins, 'dt_ms': cfg.scheme.dt,
            'launch': launch, 'graph_window': gw,
            'timed': 'K steps of the whole sweep (coupling, dfun, update, '
                     'history write, tavg monitor, FP64 FC merges), inputs '
                     'resident in HBM, wall clock between stream '
                     'synchronisations',
        },
        'roofline': None,
    }
    main_r = legs('main', time_sweep, cfg, launch, K, W, gw,
                  args.profile_steps)
    if main_r is not None:
        rf = main_r['roofline']
        rf['traffic'] = traffic_from_profile(rf['kernel'], args.workload,
                                             launch)
        cp = legs('copy', device_copy_rate)
        if cp is not None:
            if rf['traffic']:
                # measured HBM bytes per launch over the measured launch time
                cp['step_kernel_traffic_GBs'] = \
                    rf['traffic'] / (rf['launch_ms'] * 1e-3) / 1e9
                cp['step_kernel_vs_copy'] = \
                    cp['step_kernel_traffic_GBs'] / cp['GBs']
                cp['step_kernel_vs_read'] = \
                    cp['step_kernel_traffic_GBs'] / cp['read_GBs']
            rf['device_copy'] = cp
        out['value'] = main_r['value']
        out['ms_per_step'] = main_r['ms_per_step']
        out['roofline'] = rf
    plain_r, pcfg, pl = None, None, None
    if args.plain != 'none' and args.plain != args.workload:
        pcfg = configs.ALL[args.plain]()
        pl = launch_of(pcfg, None)
        log('plain workload', pcfg.name, 'launch', pl)
        plain_r = legs('plain', time_sweep, pcfg, pl, K, W, gw,
                       args.profile_steps)
        if plain_r is not None:
            plain_r['roofline']['traffic'] = traffic_from_profile(
                plain_r['roofline']['kernel'], args.plain, pl)
            out['plain'] = {'workload': '%s: %s' % (pcfg.name,
                                                    pcfg.description),
                            'value': plain_r['value'],
                            'ms_per_step': plain_r['ms_per_step'],
                            'launch': pl, 'roofline': plain_r['roofline']}
            if main_r is not None:
                out['vs_plain'] = main_r['value'] / plain_r['value']
                out['vs_plain_kernel'] = (
                    plain_r['roofline']['kernel_ms_per_step'] /
                    main_r['roofline']['kernel_ms_per_step'])
    variants, vcfgs = {}, {}
    for vname in ([] if args.variants == 'none' else
                  args.variants.split(',')):
        if vname in (args.workload, args.plain):
            continue
        vcfg = configs.ALL[vname]()
        vl = launch_of(vcfg, None)
        log('variant', vcfg.name, 'launch', vl)
        vr = legs('variant:' + vname, time_sweep, vcfg, vl, K, W, gw,
                  args.profile_steps)
        if vr is None:
            continue
        vcfgs[vname] = (vcfg, vl)
        variants[vname] = {
            'workload': '%s: %s' % (vcfg.name, vcfg.description),
            'value': vr['value'], 'ms_per_step': vr['ms_per_step'],
            'launch': vl, 'dtype': '%s history, f32 state and sums'
                                   % vcfg.ring_dtype,
            'roofline_frac': vr['roofline']['frac'],
            'roofline_frac_compulsory': vr['roofline']['compulsory']['frac'],
            'kernel_ms_per_step': vr['roofline']['kernel_ms_per_step']}
        if plain_r is not None:
            variants[vname]['vs_plain'] = vr['value'] / plain_r['value']
    if variants:
        out['variants'] = variants

    # FC after a full-length run (untimed) against the float64 oracle
    deterministic = getattr(cfg.scheme, 'sigma', None) is None
    if cfg.fc and deterministic and args.fc_instances > 0 and \
            not args.no_fc:
        n_fc = args.fc_steps or cfg.n_steps
        fc_idx = np.linspace(0, P - 1, args.fc_instances).astype(int)
        runs = [(cfg.name, cfg, launch)]
        if pcfg is not None:
            runs.append(('plain', pcfg, pl))
        runs += [(k, c, l) for k, (c, l) in vcfgs.items()]
        fcs = {}
        full = {}
        for key, c, la in runs:
            r = legs('fc:' + key, fc_run, c, la, n_fc, fc_idx, gw)
            if r is not None:
                fcs[key] = r[0]
                full[key] = {'steps': n_fc, 'fc_samples': r[1],
                             'value': P * n_fc / r[2],
                             'ms_per_step': r[2] / n_fc * 1e3}
        ref_cfg = pcfg if pcfg is not None and not pcfg.delay_bins else \
            configs.ALL['jansenrit998']()
        threads = cpu_threads(host_info())
        fc64 = legs('fc:float64', fc64_reference, ref_cfg, fc_idx, n_fc,
                    threads) if fcs else None
        if fc64 is not None:
            err = {'instances': [int(p) for p in fc_idx], 'steps': n_fc,
                   'samples': n_fc // cfg.tavg_period,
                   'reference': 'float64 oracle, exact delays, '
                                'numpy.corrcoef of its samples'}
            for key, f in fcs.items():
                err[key] = float(np.abs(f - fc64).max())
            out['fc_err_vs_fp64'] = err
        if full:
            out['full_run'] = full

    if not args.no_grad:
        g = legs('grad', time_gradient, window=1000)
        if g is not None:
            if args.grad_no_replay:
                g0 = legs('grad:no_replay', time_gradient, window=0)
                if g0 is not None:
                    g['variant_no_replay'] = g0
            out['grad'] = g
    if not args.no_cpu:
        pname = args.plain if args.plain != 'none' else args.workload
        bcfg = configs.ALL[pname]()
        if bcfg.delay_bins:
            bcfg = configs.ALL['jansenrit998']()
        threads = cpu_threads(host_info())
        out['cpu_baseline'] = legs('cpu_baseline', cpu_baseline, bcfg,
                                   min(args.cpu_inst, bcfg.n_inst),
                                   args.cpu_steps, threads)
    if legs.errors:
        out['errors'] = legs.errors
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()

"""
Benchmark of the delayed-coupling network integrator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--plain NAME|none] [--no-grad] [--no-cpu]

A "step" is one integration step of the whole sweep batch (all P instances):
delayed CSR coupling + model dfun + Euler(-Maruyama) update + history write
+ tavg monitor + FP64 online FC.  Inputs are resident in HBM before the
timed region; the timed region is K steps bracketed by stream
synchronisations.  Prints one JSON line.

value        instance*steps/s of the batched mixed-precision sweep at N=998
             (Jansen-Rit, P=1024, 32 delay bins, 8-bit fixed-point history of
             the [0, 1]-scaled sigmoid, FP32 state and sums, FP64 FC):
             P * K / wall time
variants     the same sweep with a bfloat16 history (and others on request)
plain        the same sweep through the plain FP32 kernel (exact delays,
             float32 history); vs_plain = value / plain value
fc_err_vs_fp64
             max |FC_device - FC_fp64| over the sampled instances, FC of all
             W + K steps, for the mixed and the plain path, against the
             float64 oracle run on the host (exact delays)
grad         gradient evaluations/s of the FC-fit inversion workload (MPR
             N=84, P=4096, 20000 steps, checkpoints every 1000): checkpointed
             forward + adjoint reverse sweep for all instances
roofline     the step kernel against HBM bandwidth: algorithmic bytes per
             launch (1, 2 or 4 B per edge gather + per node write, per instance
             and step, times the steps of the launch) / mean launch duration
             (HIP events on the simulation's stream); `traffic` is the
             measured HBM bytes per launch from the committed rocprofv3 PMC
             summary for this kernel, workload and launch shape
cpu_baseline the float32 CPU oracle (the reference's loop structure: per
             instance, sequential steps, CSR row reduction, OpenMP over
             instances) timed on a bounded sample of the plain workload
"""

import argparse
import json
import os
import re
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, 'tvb-hpc_amd'))

METRIC = ('instance·steps/s at N=998 sweep vs plain kernel; FC err vs '
          'FP64; grad evals/s')
HBM_PEAK_GBS = 8000.0


def log(*a):
    print(*a, file=sys.stderr, flush=True)


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


def time_sweep(cfg, launch, K, W, graph_window, profile_steps, fc_idx):
    """Run W + K steps (K timed), read the FC of instances fc_idx, then time
    the step launches one by one with events.  Returns a result dict."""
    import numpy as np
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
    assert not sim.flags().any(), 'non-finite state in the benchmark run'
    fc = None
    if cfg.fc and len(fc_idx):
        fc = np.stack([sim.fc(int(p), 1)[0] for p in fc_idx])
    # kernel time: direct launches bracketed by events on the simulation's
    # stream (the timed region above replays graph windows)
    sim.set_exec(graph=False, profile=True)
    sim.timing(reset=True)
    sim.run(profile_steps)
    sim.sync()
    ms_kern, launches = sim.timing(reset=True)
    esz = {'f32': 4, 'bf16': 2, 'f16': 2, 'u8': 1}[cfg.ring_dtype]
    bytes_inst_step = esz * (nnz + N * sim.spec.C)
    steps_per_launch = profile_steps / max(launches, 1)
    launch_ms = ms_kern / max(launches, 1)
    achieved = bytes_inst_step * P * steps_per_launch / (launch_ms * 1e-3) / 1e9
    kernel = step_kernel(sim)
    del sim
    return dict(value=P * K / el, ms_per_step=el / K * 1e3, fc=fc,
                roofline={
                    'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
                    'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
                    'traffic': None, 'kernel': kernel,
                    'bytes_per_instance_step': bytes_inst_step,
                    'steps_per_launch': steps_per_launch,
                    'launch_ms': launch_ms, 'launches': launches})


def time_gradient(n_evals=2, window=0):
    """Gradient evaluations/s of the FC-fit workload, end to end on the
    device: checkpointed forward of all instances, FC loss against a
    target FC (from a run at G = 0.5) in torch, and the adjoint."""
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
        loss = AG.fc_loss(s, fct).sum()
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
            'note': 'one evaluation = forward (recording x(n) and the '
                    'coupling sums), FC loss in torch on the device, reverse '
                    'sweep, for all instances'}


def cpu_legs(plain_cfg, fc_idx, n_fc_steps, n_inst, n_steps, threads):
    """The oracle on the host: (1) float32 on a bounded sample of the plain
    workload, timed -- the CPU baseline; (2) float64 over the benchmark's
    W + K steps for instances fc_idx -- the reference FC."""
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import numpy as np
    import oracle as O

    def oracle(cfg, idx, dtype):
        csr = cfg.csr
        # delay group of every sampled instance (one group per instance)
        gs = cfg.group_size or cfg.n_inst
        delays = csr.delays[np.asarray(idx) // gs]
        weight = csr.weight[np.asarray(idx) // gs] if csr.weight_groups \
            else csr.weight
        mp, cp = {}, {}
        for k, v in cfg.params.items():
            if k == 'sigma':
                continue
            if ':' in k:
                scope, nm = k.split(':', 1)
            else:
                scope, nm = ('c' if k in cfg.cfun.param else 'm'), k
            a = np.asarray(v)
            val = a[..., idx] if a.ndim >= 1 and a.shape[-1] == cfg.n_inst \
                else a
            (mp if scope == 'm' else cp)[nm] = val
        cp.update({k: v for k, v in cfg.cfun.param.items()
                   if k not in cp and not hasattr(v, 'name')})
        sigma = getattr(cfg.scheme, 'sigma', None)
        if sigma is not None and 'sigma' in cfg.params:
            sigma = np.asarray(cfg.params['sigma'])[idx]
        ref = O.OracleSim(type(cfg.model).__name__, type(cfg.cfun).__name__,
                          csr.nnode, csr.row, csr.col, weight, delays,
                          len(idx), csr.max_delay + 2, cfg.scheme.dt,
                          scheme='euler' if sigma is None else 'em',
                          params=mp, cparams=cp, sigma=sigma, group_size=1,
                          dtype=dtype, tavg_period=cfg.tavg_period,
                          bold=cfg.bold, bold_obs=cfg.bold_obs,
                          n_threads=threads)
        ref.set_state(cfg.initial_state()[..., idx])
        ref.init_history('zeros')
        return ref

    seed = None
    if getattr(plain_cfg.scheme, 'sigma', None) is not None:
        seed = plain_cfg.seed + 1

    idx = np.arange(n_inst)
    ref = oracle(plain_cfg, idx, np.float32)
    ref.run(5, seed=seed)
    t = time.perf_counter()
    ref.run(n_steps, seed=seed)
    el = time.perf_counter() - t
    base = {'value': n_inst * n_steps / el, 'unit': 'instance*steps/s',
            'cores': threads, 'kind': 'port',
            'sample': '%d instances x %d steps of %s on %d threads (float32 '
                      'oracle, %.1f s)' % (n_inst, n_steps, plain_cfg.name,
                                           threads, el)}
    fc64 = None
    if len(fc_idx):
        ref = oracle(plain_cfg, np.asarray(fc_idx), np.float64)
        t = time.perf_counter()
        s = ref.run(n_fc_steps, seed=seed)['samples']
        log('float64 oracle FC reference: %d instances x %d steps, %.1f s' %
            (len(fc_idx), n_fc_steps, time.perf_counter() - t))
        fc64 = np.stack([np.corrcoef(s[:, :, k].T) for k in range(len(fc_idx))])
    return base, fc64


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
        'dtype': 'f32 state, f64 sums', 'data': 'synthetic',
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
        threads = min(16, os.cpu_count() or 1)
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
    b = cfg.batch(**la)
    for k, x in enumerate(cfg.initial_states()):
        b.set_state(k, x)
    b.init_history('zeros')
    b.set_noise(seed=cfg.seed + 1)
    b.run(W)
    b.sync()
    t0 = time.perf_counter()
    b.run(K)
    b.sync()
    el = time.perf_counter() - t0
    assert not b.flags().any(), 'non-finite state in the benchmark run'
    # launch time: back-to-back launches of one block of steps each
    L = la['steps_per_launch']
    n_l = max(1, args.profile_steps // L)
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
                   'subjects': len(cfg.subjects), 'n_inst': cfg.n_inst,
                   'node_counts': Ns, 'nnz_total': sum(
                       c.nnz for c, _ in cfg.subjects),
                   'dt_ms': cfg.scheme.dt, 'launch': la},
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
        threads = min(16, os.cpu_count() or 1)
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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None)
    ap.add_argument('--warmup', type=int, default=200)
    ap.add_argument('--workload', default='jansenrit998_u8')
    ap.add_argument('--variants', default='jansenrit998_bf16',
                    help='other mixed-precision workloads timed alongside '
                         "(comma separated, or 'none')")
    ap.add_argument('--plain', default='jansenrit998',
                    help="plain-kernel workload to compare with, or 'none'")
    ap.add_argument('--no-grad', action='store_true')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--fc-instances', type=int, default=8,
                    help='instances whose FC is checked against float64')
    ap.add_argument('--cpu-inst', type=int, default=64)
    ap.add_argument('--cpu-steps', type=int, default=2500)
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
    gw = 0 if args.no_graph else args.graph_window
    cfg = configs.ALL[args.workload]()
    K = args.steps or cfg.n_steps
    W = args.warmup
    N, P, nnz = cfg.csr.nnode, cfg.n_inst, cfg.csr.nnz
    # FC of a deterministic sweep is compared with float64 (a stochastic
    # sweep's sample paths separate from the float64 ones)
    deterministic = getattr(cfg.scheme, 'sigma', None) is None
    fc_idx = np.linspace(0, P - 1, args.fc_instances).astype(int) \
        if cfg.fc and deterministic and args.fc_instances > 0 \
        else np.zeros(0, int)
    launch = launch_of(cfg, args.launch)
    log('workload', cfg.name, 'N', N, 'P', P, 'nnz', nnz, 'launch', launch)
    main_r = time_sweep(cfg, launch, K, W, gw, args.profile_steps, fc_idx)
    main_r['roofline']['traffic'] = traffic_from_profile(
        main_r['roofline']['kernel'], args.workload, launch)
    out = {
        'metric': METRIC,
        'value': main_r['value'],
        'unit': 'instance*steps/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': main_r['ms_per_step'],
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32' if cfg.ring_dtype == 'f32' else
                 'f32 state, %s history' % cfg.ring_dtype,
        'data': 'synthetic',
        'config': {
            'workload': '%s: %s' % (cfg.name, cfg.description),
            'n_node': N, 'n_inst': P, 'nnz': nnz,
            'delay_bins': cfg.delay_bins, 'dt_ms': cfg.scheme.dt,
            'launch': launch, 'graph_window': gw,
        },
        'roofline': main_r['roofline'],
    }
    plain_r = None
    if args.plain != 'none' and args.plain != args.workload:
        pcfg = configs.ALL[args.plain]()
        pl = launch_of(pcfg, None)
        log('plain workload', pcfg.name, 'launch', pl)
        plain_r = time_sweep(pcfg, pl, K, W, gw, args.profile_steps, fc_idx)
        plain_r['roofline']['traffic'] = traffic_from_profile(
            plain_r['roofline']['kernel'], args.plain, pl)
        out['plain'] = {'workload': '%s: %s' % (pcfg.name, pcfg.description),
                        'value': plain_r['value'],
                        'ms_per_step': plain_r['ms_per_step'],
                        'launch': pl, 'roofline': plain_r['roofline']}
        out['vs_plain'] = main_r['value'] / plain_r['value']
    variants = {}
    for vname in ([] if args.variants == 'none' else
                  args.variants.split(',')):
        if vname in (args.workload, args.plain):
            continue
        vcfg = configs.ALL[vname]()
        vl = launch_of(vcfg, None)
        log('variant', vcfg.name, 'launch', vl)
        vr = time_sweep(vcfg, vl, K, W, gw, args.profile_steps, fc_idx)
        variants[vname] = {'workload': '%s: %s' % (vcfg.name,
                                                   vcfg.description),
                           'value': vr['value'],
                           'ms_per_step': vr['ms_per_step'], 'launch': vl,
                           'dtype': 'f32 state, %s history' % vcfg.ring_dtype,
                           'roofline_frac': vr['roofline']['frac'],
                           'fc': vr['fc']}
    if not args.no_grad:
        out['grad'] = time_gradient()
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        pname = args.plain if args.plain != 'none' else args.workload
        pcfg = configs.ALL[pname]()
        if pcfg.delay_bins:
            pcfg = configs.ALL['jansenrit998']()
        base, fc64 = cpu_legs(pcfg, fc_idx, W + K, min(args.cpu_inst, P),
                              args.cpu_steps, threads)
        out['cpu_baseline'] = base
        if fc64 is not None and main_r['fc'] is not None:
            err = {'instances': [int(p) for p in fc_idx],
                   'samples': (W + K) // cfg.tavg_period,
                   'mixed' if cfg.delay_bins else 'main':
                       float(np.abs(main_r['fc'] - fc64).max())}
            if plain_r is not None and plain_r['fc'] is not None:
                err['plain'] = float(np.abs(plain_r['fc'] - fc64).max())
            for vname, v in variants.items():
                if v['fc'] is not None:
                    err[vname] = float(np.abs(v['fc'] - fc64).max())
            out['fc_err_vs_fp64'] = err
    if variants:
        out['variants'] = {k: {kk: vv for kk, vv in v.items() if kk != 'fc'}
                           for k, v in variants.items()}
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()

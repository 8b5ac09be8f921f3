"""
Benchmark of the delayed-coupling network integrator on one MI355X.

    python bench.py [--gpus 1] [--steps K] [--warmup W] [--workload NAME]
                    [--no-cpu] [--path plain]

A "step" is one integration step of the whole sweep batch (all P instances):
delayed CSR coupling + model dfun + Euler(-Maruyama) update + history write
(+ monitor / FC work when the workload has it).  Inputs are resident in HBM
before the timed region; the timed region is K steps bracketed by stream
synchronisations.  Prints one JSON line.

value        instance*steps/s of the whole batch (P * K / wall time)
roofline     the step kernel against HBM bandwidth: algorithmic bytes per
             launch (4 B per edge gather + 4 B per node write, per instance
             and step) / mean launch duration (HIP events on the kernel's
             stream); `traffic` is the measured HBM bytes per launch from the
             committed rocprofv3 PMC summary for this kernel, if present
cpu_baseline the float32 CPU oracle (the reference's loop structure: per
             instance, sequential steps, CSR row reduction, OpenMP over
             instances) timed on a bounded sample of the same workload
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


def traffic_from_profile(kernel, workload):
    """HBM bytes per dispatch of `kernel` on `workload` from the committed
    PMC summary (profiles/<kernel>_<workload>_pmc.txt), or None."""
    path = os.path.join(ROOT, 'profiles', '%s_%s_pmc.txt' % (kernel, workload))
    if not os.path.exists(path):
        return None
    txt = open(path).read()
    m = re.search(r'hbm_bytes_per_dispatch\s*=\s*([0-9.eE+]+)', txt)
    return float(m.group(1)) if m else None


def cpu_baseline(cfg, x0, n_inst, n_steps, threads):
    sys.path.insert(0, os.path.join(ROOT, 'oracle'))
    import numpy as np
    import oracle as O
    from tvb_hpc import configs
    sl = cfg.sliced(n_inst)
    csr = sl.csr
    name = type(sl.model).__name__
    cname = type(sl.cfun).__name__
    mp, cp = {}, {}
    for k, v in sl.params.items():
        if k == 'sigma':
            continue
        scope, nm = k.split(':', 1) if ':' in k else ('m', k)
        (mp if scope == 'm' else cp)[nm] = v
    cp.update({k: v for k, v in sl.cfun.param.items()
               if k not in cp and not hasattr(v, 'name')})
    sigma = getattr(sl.scheme, 'sigma', None)
    ref = O.OracleSim(name, cname, csr.nnode, csr.row, csr.col, csr.weight,
                      csr.delays, n_inst, csr.max_delay + 2, sl.scheme.dt,
                      scheme='em' if sigma is not None else 'euler',
                      params=mp, cparams=cp, sigma=sigma, dtype=np.float32,
                      n_threads=threads)
    ref.set_state(x0[..., :n_inst])
    ref.init_history('zeros')
    ref.run(5, seed=1 if sigma is not None else None)
    t = time.perf_counter()
    ref.run(n_steps, seed=1 if sigma is not None else None)
    el = time.perf_counter() - t
    return n_inst * n_steps / el, el


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument('--gpus', type=int, default=1)
    ap.add_argument('--steps', type=int, default=None)
    ap.add_argument('--warmup', type=int, default=200)
    ap.add_argument('--workload', default='jansenrit998')
    ap.add_argument('--path', default='plain')
    ap.add_argument('--no-cpu', action='store_true')
    ap.add_argument('--cpu-steps', type=int, default=400)
    ap.add_argument('--graph-window', type=int, default=100)
    ap.add_argument('--profile-steps', type=int, default=300)
    ap.add_argument('--launch', default=None,
                    help='ipl,waves,node_block,steps_per_launch '
                         '(default: tuned)')
    ap.add_argument('--no-graph', action='store_true',
                    help='time direct launches instead of graph windows')
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
    from tvb_hpc import autotune, configs, hip

    cfg = configs.ALL[args.workload]()
    K = args.steps or cfg.n_steps
    W = args.warmup
    N, P, nnz = cfg.csr.nnode, cfg.n_inst, cfg.csr.nnz
    x0 = cfg.initial_state()
    if args.launch:
        ipl, w, nb, spl = [int(v) for v in args.launch.split(',')]
        launch = dict(ipl=ipl, waves=w, node_block=nb, steps_per_launch=spl)
    else:
        launch = autotune.best_launch(cfg)
    log('workload', cfg.name, 'N', N, 'P', P, 'nnz', nnz, 'H', cfg.horizon,
        'launch', launch)
    sim = cfg.simulation(**launch)
    sim.set_state(x0)
    sim.init_history('zeros')
    sigma = getattr(cfg.scheme, 'sigma', None)
    if sigma is not None:
        sim.set_noise(seed=cfg.seed + 1)
    sim.set_exec(graph=not args.no_graph, window=args.graph_window,
                 profile=False)
    sim.run(W)
    sim.sync()
    t0 = time.perf_counter()
    sim.run(K)
    sim.sync()
    el = time.perf_counter() - t0
    value = P * K / el
    assert not sim.flags().any(), 'non-finite state in the benchmark run'
    # kernel time for the roofline: direct launches bracketed by events on
    # the simulation's stream (the timed region above uses graph windows)
    sim.set_exec(graph=False, profile=True)
    sim.timing(reset=True)
    sim.run(args.profile_steps)
    sim.sync()
    ms_kern, launches = sim.timing(reset=True)
    K_prof = args.profile_steps

    C = sim.spec.C
    esz = 4 if cfg.ring_dtype == 'f32' else 2
    bytes_inst_step = esz * (nnz + N * C)
    steps_per_launch = K_prof / max(launches, 1)
    achieved = bytes_inst_step * P * steps_per_launch / (
        ms_kern / launches * 1e-3) / 1e9
    traffic = traffic_from_profile('tvbx_step', args.workload)
    out = {
        'metric': METRIC,
        'value': value,
        'unit': 'instance*steps/s',
        'n_gpus': 1,
        'steps': K,
        'warmup': W,
        'ms_per_step': el / K * 1e3,
        'higher_is_better': True,
        'scaling': 'weak',
        'vs_baseline': None,
        'dtype': 'f32' if cfg.ring_dtype == 'f32' else
                 'f32 (%s history)' % cfg.ring_dtype,
        'data': 'synthetic',
        'config': {
            'workload': '%s: %s' % (cfg.name, cfg.description),
            'n_node': N, 'n_inst': P, 'nnz': nnz, 'horizon': cfg.horizon,
            'dt_ms': cfg.scheme.dt, 'path': args.path, 'launch': launch,
        },
        'roofline': {
            'bound': 'hbm', 'achieved': achieved, 'peak': HBM_PEAK_GBS,
            'unit': 'GB/s', 'frac': achieved / HBM_PEAK_GBS,
            'traffic': traffic,
            'kernel': 'tvbx_step',
            'bytes_per_instance_step': bytes_inst_step,
            'launch_ms': ms_kern / launches,
            'launches': launches,
        },
    }
    if not args.no_cpu:
        threads = min(16, os.cpu_count() or 1)
        n_inst = min(P, threads)
        v, el_c = cpu_baseline(cfg, x0, n_inst, args.cpu_steps, threads)
        out['cpu_baseline'] = {
            'value': v, 'unit': 'instance*steps/s', 'cores': threads,
            'kind': 'port',
            'sample': '%d instances x %d steps of %s on %d threads (float32 '
                      'oracle, %.1f s)' % (n_inst, args.cpu_steps, cfg.name,
                                           threads, el_c),
        }
    print(json.dumps(out), flush=True)


if __name__ == '__main__':
    main()

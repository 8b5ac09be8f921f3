"""
Launch-shape autotuner for the step kernels.

The choices:

* ``ipl``: sweep instances per lane (1, 2, 4).  A wavefront covers
  64 * ipl consecutive instances, so each delayed-observable gather is one
  contiguous 256 B .. 1 KiB read (float32 ring) -- longer reads use HBM
  better -- at the cost of ipl-fold registers per lane and fewer
  workgroups.  Delay groups must be multiples of 64 * ipl instances.
* ``node_block`` / ``waves``: nodes per workgroup and waves per workgroup.
  Each wave takes one node at a time.  With node_block < N the node axis is
  split over workgroups and every step is one launch (grid-wide dependency
  between steps); with node_block >= N one workgroup holds a whole instance
  tile and ``steps_per_launch`` steps run inside one launch, separated by
  workgroup barriers.
* ``unroll``: delayed gathers issued together per CSR row group (8, 16,
  32): more loads in flight per wave against more registers.
* ``min_waves``: waves per SIMD the blocked kernel's register budget must
  allow (more resident waves against register spills).
* ``pipe``: the next group of gathers is issued before the current one is
  consumed (two groups in flight per wave, two register sets).
* ``prio`` (on unless a launch says otherwise): in multi-step launches
  without barriers the wave priority falls with the launch's progress, so
  the waves sharing a SIMD end together instead of the last ones running
  alone.  It helps where the waves are issue-bound (the 8- and 16-bit
  rings) and can cost a few per cent where they are purely HBM-bound.
* ``xcd_map``: workgroup numbering that gives each of the 8 XCDs whole
  instance tiles, so the rows that read the same ring line share an L2.
* temporal blocking: with one node per wave (node_block == waves < N),
  ``steps_per_launch`` up to d_min + 1 (d_min the smallest delay) steps run
  in one launch with no synchronisation at all, the state and the monitor
  accumulator staying in registers, because no step of the launch reads a
  ring slot another workgroup writes in it.

``tune(cfg)`` times candidates on the device with short runs and records
the best per (workload, N, P, nnz, device) in a JSON table next to the
library; ``best_launch(cfg)`` reads that table, falling back to measured
defaults for MI355X.
"""

import json
import os
import time

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
TABLE = os.path.join(HERE, '_lib', 'tuned.json')


def valid_ipls(cfg):
    P = cfg.n_inst
    out = []
    for ipl in (1, 2, 4):
        tile = 64 * ipl
        if cfg.group_size is not None and cfg.group_size % tile:
            continue
        # keep at least a few instance tiles per node block
        if ipl > 1 and P < 2 * tile:
            continue
        out.append(ipl)
    return out


def candidates(cfg):
    N = cfg.csr.nnode
    out = []
    if getattr(cfg, 'coupling', 'gather') == 'mfma':
        # blocked advance shapes only, at most mfma_steps(ipl) steps
        for ipl in valid_ipls(cfg):
            for spl in sorted({min(b, mfma_steps(ipl))
                               for b in block_lengths(cfg)}):
                for waves in (1, 2, 4, 8):
                    out.append(dict(ipl=ipl, waves=waves, node_block=waves,
                                    steps_per_launch=spl))
        return out
    for ipl in valid_ipls(cfg):
        for waves, nb in ((4, 4), (2, 2), (8, 8), (4, 8), (8, 16), (1, 1),
                          (16, 16)):
            if nb <= N:
                out.append(dict(ipl=ipl, waves=waves, node_block=nb,
                                steps_per_launch=1))
        if N <= 400:
            for waves in (8, 16):
                out.append(dict(ipl=ipl, waves=waves, node_block=N,
                                steps_per_launch=50))
        for spl in block_lengths(cfg):
            for waves in (1, 2, 4, 8):
                if waves < N:
                    out.append(dict(ipl=ipl, waves=waves, node_block=waves,
                                    steps_per_launch=spl))
    return out


def mfma_steps(ipl):
    "steps per launch the matrix-core sums kernel holds accumulators for"
    return 8 if ipl == 4 else 16


def block_lengths(cfg):
    "Temporal-blocking lengths the delays allow (up to d_min + 1, <= 64)."
    if cfg.csr.nnz == 0:
        return []
    lmax = int(np.min(cfg.csr.delays)) + 1
    return sorted({min(lmax, b) for b in (8, 16, 32, 64)} - {1})


def default_launch(cfg):
    """Defaults for a configuration with no measured entry, after what won
    for the built-in ones: the widest ipl the instance count allows; when the
    smallest delay allows blocks of 8 steps or more, one node per wave in
    blocks of up to 32 steps with the 4-waves-per-SIMD register budget (16
    gathers per group from 2 instances per lane, 8 pipelined below), else
    per-step launches with 4-node workgroups."""
    ipls = valid_ipls(cfg)
    ipl = 4 if 4 in ipls and cfg.n_inst >= 1024 else max(
        [i for i in ipls if i <= 2] or [1])
    bl = block_lengths(cfg)
    if getattr(cfg, 'coupling', 'gather') == 'mfma':
        # the sums come from the matrix-core kernel: blocks of up to
        # mfma_steps(ipl) steps, a light advance kernel
        return dict(ipl=ipl, waves=4, node_block=4,
                    steps_per_launch=min(max(bl or [1]), mfma_steps(ipl)))
    if bl and max(bl) >= 8:
        la = dict(ipl=ipl, waves=1, node_block=1,
                  steps_per_launch=min(max(bl), 32), min_waves=4)
        la.update(dict(unroll=16) if ipl >= 2 else dict(unroll=8, pipe=True))
        return la
    return dict(ipl=ipl, waves=4, node_block=4, steps_per_launch=1)


def _prefix(cfg):
    # the workload (name and sizes), not the kernel source: a launch shape
    # stays a good starting point across source revisions
    if hasattr(cfg, 'subjects'):
        return '%s|S%d|P%d|nnz%d|' % (cfg.name, len(cfg.subjects), cfg.n_inst,
                                      sum(c.nnz for c, _ in cfg.subjects))
    return '%s|N%d|P%d|nnz%d|' % (cfg.name, cfg.csr.nnode, cfg.n_inst,
                                  cfg.csr.nnz)


def _key(cfg, ctx_info):
    return _prefix(cfg) + ctx_info.get('arch', '?')


def _load():
    try:
        with open(TABLE) as fd:
            return json.load(fd)
    except (OSError, ValueError):
        return {}


def best_launch(cfg, ctx=None):
    table = _load()
    if table:
        from .hip import Context
        ctx = ctx or Context.default()
        hit = table.get(_key(cfg, ctx.info()))
        if hit:
            return hit['launch']
    return default_launch(cfg)


def table_launches(cfg):
    "Launches recorded for this configuration (on any device)."
    prefix = _prefix(cfg)
    return [v['launch'] for k, v in _load().items() if k.startswith(prefix)]


def _time(cfg, cand, x0, steps, warmup, ctx):
    sim = cfg.simulation(context=ctx, **cand)
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.run(warmup)
    sim.sync()
    t = time.perf_counter()
    sim.run(steps)
    sim.sync()
    ms = (time.perf_counter() - t) / steps * 1e3
    del sim
    return ms


def _emaj_allowed(cfg):
    "edge-major sums need a pre() that ignores post_syn (the CUDA order)"
    return not cfg.cfun.pre_uses_post_syn()


def tune(cfg, steps=100, warmup=10, ctx=None, log=None,
         variants=(dict(unroll=16), dict(unroll=4), dict(min_waves=4),
                   dict(min_waves=4, unroll=4), dict(min_waves=4, unroll=16),
                   dict(min_waves=2), dict(min_waves=2, unroll=16),
                   dict(xcd_map=True), dict(xcd_map=True, min_waves=4),
                   dict(pipe=True), dict(pipe=True, min_waves=4),
                   dict(pipe=True, min_waves=4, unroll=4),
                   dict(pipe=True, min_waves=4, unroll=16),
                   dict(prio=False), dict(prio=False, min_waves=4),
                   dict(prio=False, min_waves=4, unroll=4)),
         top=3):
    """Time every launch shape for `steps` steps (and the launch recorded
    before, if any), then the `top` best of each launch width with the
    kernel variants; record and return the best."""
    from .hip import Context
    ctx = ctx or Context.default()
    x0 = cfg.initial_state()
    results = []
    cands = candidates(cfg)
    prev = _load().get(_key(cfg, ctx.info()))
    if prev and prev['launch'] not in cands:
        cands.append(dict(prev['launch']))
    # one untimed pass first: the device leaves its idle clocks
    _time(cfg, cands[0], x0, steps, warmup, ctx)
    for cand in cands:
        ms = _time(cfg, cand, x0, steps, warmup, ctx)
        results.append((ms, cand))
        if log:
            log('%s %.4f ms/step' % (cand, ms))
    results.sort(key=lambda r: r[0])
    bases = []
    for ipl in sorted({c['ipl'] for _, c in results}):
        bases += [c for _, c in results if c['ipl'] == ipl][:top]
    blocked = [b for b in bases if b['steps_per_launch'] > 1 and
               b['node_block'] < cfg.csr.nnode]
    emaj_ok = _emaj_allowed(cfg)
    for base in bases:
        extra = []
        if emaj_ok and base in blocked and base['steps_per_launch'] <= 16:
            e = min(v for v in (4, 8, 16) if v >= base['steps_per_launch'])
            extra = [dict(emaj=e), dict(emaj=e, xcd_map=True),
                     dict(emaj=e, min_waves=4), dict(emaj=e, unroll=8),
                     dict(emaj=e, xcd_map=True, min_waves=4)]
        for var in list(variants) + extra:
            if all(base.get(k) == v for k, v in var.items()):
                continue
            if 'min_waves' in var and base not in blocked:
                continue   # only the blocked kernel takes the bound
            cand = dict(base, **var)
            ms = _time(cfg, cand, x0, steps, warmup, ctx)
            results.append((ms, cand))
            if log:
                log('%s %.4f ms/step' % (cand, ms))
    results.sort(key=lambda r: r[0])
    best = results[0][1]
    table = _load()
    table[_key(cfg, ctx.info())] = {'launch': best, 'ms_per_step':
                                    results[0][0], 'tried': len(results)}
    os.makedirs(os.path.dirname(TABLE), exist_ok=True)
    tmp = TABLE + '.%d.tmp' % os.getpid()
    with open(tmp, 'w') as fd:
        json.dump(table, fd, indent=1, sort_keys=True)
    os.replace(tmp, TABLE)
    return best, results


# -- batches of subjects -------------------------------------------------------

def batch_candidates(cfg):
    "Launch shapes of a subject batch: one node per wave, temporal blocks."
    dmin = min(int(np.min(c.delays)) for c, _ in cfg.subjects if c.nnz)
    out = []
    for spl in sorted({1} | {min(dmin + 1, b) for b in (8, 16, 32, 64)}):
        for waves in (1, 2, 4, 8):
            for mw in (1, 4):
                for pipe in (False, True):
                    out.append(dict(ipl=1, waves=waves, steps_per_launch=spl,
                                    min_waves=mw, pipe=pipe))
    out += [dict(c, prio=False) for c in out if c['pipe']]
    return out


def default_batch_launch(cfg):
    dmin = min(int(np.min(c.delays)) for c, _ in cfg.subjects if c.nnz)
    return dict(ipl=1, waves=4, steps_per_launch=min(dmin + 1, 16))


def best_batch_launch(cfg, ctx=None):
    table = _load()
    if table:
        from .hip import Context
        ctx = ctx or Context.default()
        hit = table.get(_key(cfg, ctx.info()))
        if hit:
            return hit['launch']
    return default_batch_launch(cfg)


def _time_batch(cfg, cand, x0, steps, warmup, ctx):
    b = cfg.batch(context=ctx, **cand)
    for k, x in enumerate(x0):
        b.set_state(k, x)
    b.init_history('zeros')
    b.set_noise(seed=1)
    b.set_exec(graph=True, window=100)
    b.run(warmup)
    b.sync()
    t = time.perf_counter()
    b.run(steps)
    b.sync()
    ms = (time.perf_counter() - t) / steps * 1e3
    del b
    return ms


def tune_batch(cfg, steps=200, warmup=20, ctx=None, log=None, top=6):
    """Time the batch launch shapes (graph windows of 100 steps), then the
    `top` best again with persistent work-queue workgroups; record and
    return the best."""
    from .hip import Context
    ctx = ctx or Context.default()
    x0 = cfg.initial_states()
    results = []
    cands = batch_candidates(cfg)
    # the first shape twice: its first, untimed pass raises the clocks
    _time_batch(cfg, cands[0], x0, steps, warmup, ctx)
    for cand in cands:
        ms = _time_batch(cfg, cand, x0, steps, warmup, ctx)
        results.append((ms, cand))
        if log:
            log('%s %.4f ms/step' % (cand, ms))
    results.sort(key=lambda r: r[0])
    for _, base in list(results[:top]):
        cand = dict(base, queue=True)
        ms = _time_batch(cfg, cand, x0, steps, warmup, ctx)
        results.append((ms, cand))
        if log:
            log('%s %.4f ms/step' % (cand, ms))
    results.sort(key=lambda r: r[0])
    best = results[0][1]
    table = _load()
    table[_key(cfg, ctx.info())] = {'launch': best, 'ms_per_step':
                                    results[0][0], 'tried': len(results)}
    os.makedirs(os.path.dirname(TABLE), exist_ok=True)
    tmp = TABLE + '.%d.tmp' % os.getpid()
    with open(tmp, 'w') as fd:
        json.dump(table, fd, indent=1, sort_keys=True)
    os.replace(tmp, TABLE)
    return best, results

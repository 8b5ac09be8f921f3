"""Batches of subjects with different connectomes in one launch (the
uneven-size sweep), through libtvbx's C ABI, against the float32 oracle run
per subject, and against one simulation per subject."""

import numpy as np
import pytest

from harness import oracle_for, scaled_err, small_csr
from tvb_hpc import coupling, model, scheme
from tvb_hpc.hip import Batch, Simulation

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _subjects():
    """three connectomes (different N, seeds, speeds) with 64, 100 and 64
    instances: the middle one pads to 128 lanes"""
    out = []
    for N, seed, v, P in ((12, 3, 1.0, 64), (20, 4, 2.0, 100),
                          (16, 5, 1.5, 64)):
        csr = small_csr(N=N, seed=seed, speeds=(v, ), dt=0.1, lmax=50.0)
        out.append((csr, P))
    return out


def _params(P_total):
    rng = np.random.default_rng(7)
    return {'c:a': rng.uniform(0.0, 0.5, P_total).astype(np.float32),
            'm:J': rng.uniform(0.1, 0.3, P_total).astype(np.float32),
            'sigma': rng.uniform(1e-3, 1e-2, P_total).astype(np.float32)}


def _block_len(cap=8):
    "longest temporal block every test subject's delays allow (<= cap)"
    return min(min(int(csr.delays.min()) for csr, _ in _subjects()) + 1, cap)


def _launch(launch):
    la = dict(launch)
    if la.get('steps_per_launch') == 'L':
        la['steps_per_launch'] = _block_len()
    return la


def _batch(launch, em=True):
    launch = _launch(launch)
    subj = _subjects()
    m = model.ReducedWongWang()
    c = coupling.Linear(m)
    sc = scheme.EulerMaryuyamaStep(0.1) if em else scheme.EulerStep(0.1)
    P_total = sum(p for _, p in subj)
    prm = _params(P_total)
    if not em:
        prm.pop('sigma')
    kinds = {'c:a': 'inst', 'm:J': 'inst'}
    if em:
        kinds['sigma'] = 'inst'
    b = Batch(m, c, sc, subj, params=prm, kinds=kinds, **launch)
    rng = np.random.default_rng(11)
    x0 = [m.initial_state(csr.nnode, p, rng) for csr, p in subj]
    for k, x in enumerate(x0):
        b.set_state(k, x)
    b.init_history('zeros')
    return b, subj, prm, x0, m, c


def _inst0(subj, ipl=1):
    tile = 64 * ipl
    off, out = 0, []
    for _, p in subj:
        out.append(off)
        off += (p + tile - 1) // tile * tile
    return out


@pytest.mark.parametrize('launch', [
    dict(waves=4, steps_per_launch=1),
    dict(waves=2, steps_per_launch='L'),
    dict(waves=1, steps_per_launch='L', min_waves=4)])
def test_batch_matches_oracle_per_subject(launch):
    """Philox noise: subject k's instances use the counters of their
    batch-wide indices, which the oracle takes as an instance offset."""
    b, subj, prm, x0, m, c = _batch(launch)
    b.set_noise(seed=21)
    b.run(800)
    off = np.cumsum([0] + [p for _, p in subj])
    i0 = _inst0(subj)
    for k, (csr, p) in enumerate(subj):
        sl = slice(off[k], off[k + 1])
        ref = oracle_for('ReducedWongWang', 'Linear', csr, p, b.H[k], 0.1,
                         'em', params={'J': prm['m:J'][sl]},
                         cparams={'a': prm['c:a'][sl]},
                         sigma=prm['sigma'][sl], inst_off=i0[k])
        ref.set_state(x0[k])
        ref.init_history('zeros')
        ref.run(800, seed=21)
        err = scaled_err(b.get_state(k), ref.get_state())
        print('subject %d (N %d, P %d): %.2e' % (k, csr.nnode, p, err))
        assert err < TOL
    assert not b.flags().any()


def test_batch_equals_one_simulation_per_subject():
    """Without noise each subject of the batch computes exactly what a
    simulation of that subject alone computes with the same launch."""
    launch = dict(waves=2, steps_per_launch='L')
    b, subj, prm, x0, m, c = _batch(launch, em=False)
    b.run(400)
    off = np.cumsum([0] + [p for _, p in subj])
    for k, (csr, p) in enumerate(subj):
        sl = slice(off[k], off[k + 1])
        sim = Simulation(m, c, scheme.EulerStep(0.1), csr, p,
                         params={'c:a': prm['c:a'][sl], 'm:J': prm['m:J'][sl]},
                         kinds={'c:a': 'inst', 'm:J': 'inst'}, waves=2,
                         node_block=2, steps_per_launch=_block_len())
        sim.set_state(x0[k])
        sim.init_history('zeros')
        sim.run(400)
        np.testing.assert_array_equal(b.get_state(k), sim.get_state())


def test_batch_blocked_equals_single_step_launches():
    bs = []
    for launch in (dict(waves=4, steps_per_launch=1),
                   dict(waves=4, steps_per_launch='L')):
        b, subj, prm, x0, m, c = _batch(launch)
        b.set_noise(seed=3)
        b.run(300)
        bs.append([b.get_state(k) for k in range(len(subj))])
    for a, c in zip(*bs):
        np.testing.assert_array_equal(a, c)


def test_batch_validation():
    subj = _subjects()
    m = model.ReducedWongWang()
    c = coupling.Linear(m)
    dmin = min(int(csr.delays.min()) for csr, _ in subj)
    with pytest.raises(ValueError):
        Batch(m, c, scheme.EulerStep(0.1), subj, steps_per_launch=dmin + 2)
    b = Batch(m, c, scheme.EulerStep(0.1), subj, waves=2,
              steps_per_launch=dmin + 1)
    with pytest.raises(ValueError):
        b.set_state(0, np.zeros((1, 5, 64), np.float32))


@pytest.mark.parametrize('graph', [False, True])
def test_batch_work_queue_and_graphs_equal_direct_launches(graph):
    """Persistent work-queue workgroups (items claimed with an atomic
    counter) and captured graph windows compute exactly what one workgroup
    per item launched from the host computes, over several launches (the
    queue counter must come back to zero after each)."""
    L = _block_len()
    ref = None
    for queue, g in ((False, False), (True, graph)):
        b, subj, prm, x0, m, c = _batch(dict(waves=2, steps_per_launch=L))
        b.set_noise(seed=5)
        b.set_exec(graph=g, window=3 * L, queue=queue)
        b.run(10 * L + 3)
        got = [b.get_state(k) for k in range(len(subj))]
        if ref is None:
            ref = got
        else:
            for a, r in zip(got, ref):
                np.testing.assert_array_equal(a, r)
        assert not b.flags().any()

"""The drop-in path on the device: ``transforms.network_time_step`` with the
HIP target, called with the reference kernel's arguments, and the
reference's own acceptance check (examples/hackathon.py) run on the
device."""

import itertools

import numpy as np
import pytest

import oracle as O
from harness import scaled_err, small_csr
from kat_hackathon import initial_theta, trend_stats, workload
from tvb_hpc import coupling, model, scheme, transforms
from tvb_hpc import expr as E
from tvb_hpc.hip import HipTarget, Simulation
from tvb_hpc.network import CSR

pytestmark = pytest.mark.gpu


def _hackathon_like(batched):
    """examples/hackathon.py:23-55 on a small connectome: Kuramoto, Euler,
    coupling a as a run-time argument, windows of 10 steps."""
    osc = model.Kuramoto()
    osc.dt = 1.0
    osc.const['omega'] = 10.0 * 2.0 * np.pi / 1e3
    cfun = coupling.Kuramoto(osc)
    cfun.param['a'] = E.var('a')
    scm = scheme.EulerStep(osc.dt)
    knl = transforms.network_time_step(osc, cfun, scm, target=HipTarget())
    csr = small_csr(N=20, dt=1.0, speeds=(0.7, ), zero_delays=6)
    N = csr.nnode
    P = 70 if batched else 1
    delays = csr.delays[0].astype(np.uintc)
    ntime = int(delays.max()) + 3 + 400
    rng = np.random.default_rng(2)
    shape = (N, 1, P) if batched else (N, 1)
    state = rng.uniform(0, 2 * np.pi, size=shape).astype(np.float32)
    inp = np.zeros(shape, np.float32)
    param = np.full((N, 1, P) if batched else (N, 1), osc.const['omega'],
                    np.float32)
    drift = np.zeros(shape, np.float32)
    diffs = np.zeros(shape, np.float32)
    obsrv = np.zeros((ntime, N, 2, P) if batched else (ntime, N, 2),
                     np.float32)
    a = np.linspace(0.5, 3, P).astype(np.float32) if batched else 2.0
    ref = O.OracleSim('Kuramoto', 'Kuramoto', N, csr.row, csr.col,
                      csr.weight, delays, P, ntime, 1.0, scheme='euler',
                      params={'omega': osc.const['omega']},
                      cparams={'a': a}, R=2, dtype=np.float32,
                      order='reference')
    x0 = state.reshape(N, 1, P)
    ref.set_state(np.transpose(x0, (1, 0, 2)))
    ref.set_ring(np.transpose(obsrv.reshape(ntime, N, 2, P), (0, 2, 1, 3)))
    tol = 1e-4 if batched else 1e-5
    for i in range(30):
        knl(nstep=10, nnode=N, ntime=ntime, state=state, input=inp,
            param=param, drift=drift, diffs=diffs, obsrv=obsrv,
            nnz=csr.nnz, delays=delays, row=csr.row.astype(np.uintc),
            col=csr.col.astype(np.uintc), weights=csr.weight, a=a,
            i_step_0=i * 10)
        if i == 0:
            # the first call fills obsrv[0] from the state; later calls
            # continue the oracle's own history
            ref.init_history('current')
        ref.run(10)
        xr = ref.get_state()
        xg = np.transpose(state.reshape(N, 1, P), (1, 0, 2))
        dphi = np.angle(np.exp(1j * (xg.astype(np.float64) - xr)))
        # device sinf differs from the host's in the last bit, and Kuramoto
        # at coupling a up to 3 amplifies that ~10x per 100 steps
        assert np.abs(dphi).max() / (2 * np.pi) < tol, i
    rr = ref.get_ring()  # (H, R, N, P)
    og = np.transpose(obsrv.reshape(ntime, N, 2, P), (0, 2, 1, 3))
    th_g, th_r = og[:, 0], rr[:, 0]
    dphi = np.angle(np.exp(1j * (th_g.astype(np.float64) - th_r)))
    assert np.abs(dphi).max() / (2 * np.pi) < tol
    assert scaled_err(og[:, 1], rr[:, 1]) < 10 * tol
    # slots written by the 300 steps are exactly 0 .. 299, as the
    # reference's obsrv after 30 windows of 10 steps
    written = np.abs(og).sum(axis=(1, 2, 3)) > 0
    assert written[:300].all() and not written[300:].any()


def test_network_time_step_executor_unbatched():
    _hackathon_like(batched=False)


def test_network_time_step_executor_batched():
    _hackathon_like(batched=True)


def test_hackathon_known_answer_on_device():
    """examples/hackathon.py:58-89 on the device: 8 x 8 (speed x coupling)
    Kuramoto sweep on hcp0, 4000 Euler steps of 1 ms, 10-step obsrv sums;
    the FC-vs-SC trend statistics must satisfy the reference's assertions
    derr_speed > 350 and derr_coupl < -500, and the device trace must match
    the oracle's."""
    conn, couplings, speeds = workload()
    N = conn.nnode
    osc = model.Kuramoto()
    cfun = coupling.Kuramoto(osc)
    # one delay group per speed; 64 lanes per group, the first 8 used
    delays = np.stack([(conn.lnz / s / 1.0).astype(np.uintc)
                       for s in speeds]).astype(np.int32)
    csr = CSR(N, conn.row, conn.col, conn.wnz, delays)
    G = 64
    P = G * speeds.size
    a = np.tile(np.resize(couplings, G), speeds.size).astype(np.float32)
    H = int(delays.max()) + 3 + 4000
    sim = Simulation(osc, cfun, scheme.EulerStep(1.0), csr, P, horizon=H,
                     params={'omega': 2 * np.pi * 10 / 1e3, 'c:a': a},
                     kinds={'m:omega': 'glob'}, group_size=G, tavg_period=10,
                     sample_slots=400, order='reference', waves=8)
    th0 = np.zeros((1, N, P), np.float32)
    items = list(itertools.product(range(speeds.size), range(couplings.size)))
    th_real = initial_theta(N, len(items))
    for j, (si, ci) in enumerate(items):
        th0[0, :, si * G + ci] = th_real[:, j]
    sim.set_state(th0)
    sim.init_history('zeros')
    sim.run(4000)
    smp, first = sim.samples()
    assert first == 0 and smp.shape[0] == 400
    cols = [si * G + ci for si, ci in items]
    trace = (smp[:, :, cols] * 10).transpose(2, 0, 1)
    ds, dc = trend_stats(trace, conn.weights, speeds, couplings)
    print('device derr_speed=%.1f derr_coupl=%.1f' % (ds, dc))
    assert ds > 350.0
    assert dc < -500.0


def test_hackathon_known_answer_fc_sc_on_device():
    """The same acceptance check with the windowed FC and the FC-SC error
    computed on the device (FP64 co-moments restarted per 2 s window): only
    the (windows, P) errors come back; they equal the host statistic of the
    same samples, and the trend assertions hold."""
    from tvb_hpc import metric
    conn, couplings, speeds = workload()
    N = conn.nnode
    osc = model.Kuramoto()
    cfun = coupling.Kuramoto(osc)
    delays = np.stack([(conn.lnz / s / 1.0).astype(np.uintc)
                       for s in speeds]).astype(np.int32)
    csr = CSR(N, conn.row, conn.col, conn.wnz, delays)
    G = 64
    P = G * speeds.size
    a = np.tile(np.resize(couplings, G), speeds.size).astype(np.float32)
    H = int(delays.max()) + 3 + 4000
    kw = dict(horizon=H, params={'omega': 2 * np.pi * 10 / 1e3, 'c:a': a},
              kinds={'m:omega': 'glob'}, group_size=G, tavg_period=10,
              order='reference', waves=8)
    th0 = np.zeros((1, N, P), np.float32)
    items = list(itertools.product(range(speeds.size), range(couplings.size)))
    th_real = initial_theta(N, len(items))
    for j, (si, ci) in enumerate(items):
        th0[0, :, si * G + ci] = th_real[:, j]
    cols = [si * G + ci for si, ci in items]
    sim = Simulation(osc, cfun, scheme.EulerStep(1.0), csr, P, fc=True,
                     sample_slots=64, **kw)
    sim.set_state(th0)
    sim.init_history('zeros')
    err = metric.windowed_fc_sc_error(sim, conn.weights, 2, 2000)
    assert err.shape == (2, P)
    # the same run keeping its samples: host statistic per window
    sim2 = Simulation(osc, cfun, scheme.EulerStep(1.0), csr, P,
                      sample_slots=400, **kw)
    sim2.set_state(th0)
    sim2.init_history('zeros')
    sim2.run(4000)
    smp, _ = sim2.samples()
    for w in range(2):
        fc64 = O.fc_f64(smp[w * 200:(w + 1) * 200, :, cols].astype(np.float64))
        ref = metric.fc_sc_error(fc64, conn.weights.astype(np.float64))
        np.testing.assert_allclose(err[w, cols], ref, rtol=1e-9, atol=1e-9)
    ds, dc = metric.sweep_trends(err[-1, cols], speeds.size, couplings.size)
    print('device FC-SC derr_speed=%.1f derr_coupl=%.1f' % (ds, dc))
    assert ds > 350.0
    assert dc < -500.0

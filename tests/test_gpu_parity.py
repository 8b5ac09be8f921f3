"""Parity of the HIP integrator (called through libtvbx's C ABI) with the
CPU oracle, on the device.

Bar: float32 trajectories within 1e-5 of the float32 oracle, relative to each
state variable's scale (``harness.state_err``; Kuramoto phases modulo 2 pi,
relative to 2 pi), over 1000 steps with identical injected noise.  The two
sides differ only in the last-bit rounding of the transcendental functions
(libm vs. the device math library).
"""

import numpy as np
import pytest

import oracle as O
from harness import (CASES, CFUNS, MODELS, oracle_for, scaled_err, small_csr,
                     state_err)
from tvb_hpc import scheme as SC
from tvb_hpc.hip import Simulation, TvbxError
from tvb_hpc.network import CSR

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _pair(model, cfun, dt, cparams, P=96, N=24, steps=1000, em=True,
          launch=None, ring_pre=None, init='zeros', seed_noise=11,
          order='cuda', zero_delays=0, ring_dtype='f32', ring_scale=1.0):
    csr = small_csr(N=N, dt=dt, zero_delays=zero_delays)
    m = MODELS[model]()
    c = CFUNS[cfun](m)
    scheme = SC.EulerMaryuyamaStep(dt) if em else SC.EulerStep(dt)
    sigma = 0.01 if em else None
    params = {'c:' + k: np.full(P, v, np.float32) if k == 'a' else v
              for k, v in cparams.items()}
    if em:
        params['sigma'] = np.full(P, sigma, np.float32)
    sim = Simulation(m, c, scheme, csr, P, params=params,
                     kinds={'sigma': 'inst'} if em else None,
                     ring_pre=ring_pre, order=order, ring_dtype=ring_dtype,
                     ring_scale=ring_scale, **(launch or {}))
    rng = np.random.default_rng(seed_noise)
    x0 = m.initial_state(N, P, rng)
    sim.set_state(x0)
    sim.init_history(init)
    ref = oracle_for(model, cfun, csr, P, sim.H, dt, 'em' if em else 'euler',
                     cparams=cparams, sigma=sigma, order=order,
                     ring_dtype=ring_dtype, ring_pre=sim.spec.ring_pre,
                     ring_scale=ring_scale)
    ref.set_state(x0)
    ref.init_history(init)
    S = len(m.state_sym)
    noise = rng.standard_normal((steps, S, N, P)).astype(np.float32) \
        if em else None
    return sim, ref, noise


LAUNCHES = [dict(waves=4, node_block=4, steps_per_launch=1),
            dict(waves=8, node_block=10 ** 6, steps_per_launch=100)]


@pytest.mark.parametrize('model,cfun,dt,cparams', CASES)
@pytest.mark.parametrize('launch', LAUNCHES)
def test_trajectory_parity_injected_noise(model, cfun, dt, cparams, launch):
    sim, ref, noise = _pair(model, cfun, dt, cparams, launch=launch)
    for k0 in range(0, noise.shape[0], 250):
        sim.run(250, noise=noise[k0:k0 + 250])
        ref.run(250, noise=noise[k0:k0 + 250])
        xg, xr = sim.get_state(), ref.get_state()
        assert np.isfinite(xg).all()
        err = state_err(model, xg, xr)
        assert err < TOL, (model, cfun, k0 + 250, err)
    assert not sim.flags().any()


@pytest.mark.parametrize('model,cfun,dt,cparams',
                         [CASES[0], CASES[3], CASES[4]])
@pytest.mark.parametrize('launch', [
    dict(ipl=2, waves=4, node_block=4),
    dict(ipl=4, waves=4, node_block=2),
    dict(ipl=4, waves=8, node_block=10 ** 6, steps_per_launch=50)])
def test_trajectory_parity_wide_lanes(model, cfun, dt, cparams, launch):
    """Two and four instances per lane (vector loads of the ring), with an
    instance count that is not a multiple of the lane width."""
    sim, ref, noise = _pair(model, cfun, dt, cparams, P=300, launch=launch,
                            steps=500)
    sim.run(500, noise=noise)
    ref.run(500, noise=noise)
    assert state_err(model, sim.get_state(), ref.get_state()) < TOL
    assert not sim.flags().any()


@pytest.mark.parametrize('ring_dtype', ['bf16', 'f16', 'u8'])
@pytest.mark.parametrize('ipl', [1, 2, 4])
def test_reduced_precision_ring_parity(ring_dtype, ipl):
    """bf16 / f16 / 8-bit fixed-point history holding 200 * sigmoid(y)
    (1 / cmax scaling, values in [0, 1]) vs the float32 oracle emulating the
    same rounding.  Rounding can resolve a last-bit difference of expf
    either way, so the bar is 1e-4."""
    sim, ref, noise = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3},
                            P=200, steps=600, ring_dtype=ring_dtype,
                            ring_scale=200.0,
                            launch=dict(ipl=ipl, waves=4, node_block=4))
    assert sim.spec.ring_pre
    sim.run(600, noise=noise)
    ref.run(600, noise=noise)
    err = state_err('JansenRit', sim.get_state(), ref.get_state())
    print('JR %s ring vs emulating oracle: %.2e' % (ring_dtype, err))
    assert err < 1e-4


def test_reduced_precision_raw_ring_parity():
    "Kuramoto phases stored in bf16 (pre needs post_syn: no pre-transform)"
    sim, ref, noise = _pair('Kuramoto', 'Kuramoto', 0.1, {'a': 1.5}, P=128,
                            steps=400, ring_dtype='bf16')
    assert not sim.spec.ring_pre
    sim.run(400, noise=noise)
    ref.run(400, noise=noise)
    assert state_err('Kuramoto', sim.get_state(), ref.get_state()) < 1e-4


@pytest.mark.parametrize('model,cfun,dt,cparams',
                         [c for c in CASES if c[1] in ('Kuramoto', 'Diff')])
@pytest.mark.parametrize('launch', LAUNCHES)
def test_reference_order_parity(model, cfun, dt, cparams, launch):
    """The loopy kernel's order (post_syn and zero delays read the slot
    before it is rewritten), with some zero-delay edges."""
    sim, ref, noise = _pair(model, cfun, dt, cparams, launch=launch,
                            order='reference', zero_delays=12)
    sim.run(600, noise=noise[:600])
    ref.run(600, noise=noise[:600])
    assert state_err(model, sim.get_state(), ref.get_state()) < TOL
    # the ring as the reference's obsrv array would hold it
    assert scaled_err(sim.get_ring(), ref.get_ring()) < TOL


@pytest.mark.parametrize('launch', LAUNCHES)
def test_kuramoto_far_phases_take_the_checked_path(launch):
    """Phases far outside [0, 2 pi) (the model wraps them by one period
    per step) put the coupling's sine arguments beyond the reduced range in
    some lanes and not others: those gather groups are redone through the
    library's sine, and the trajectory still follows the oracle."""
    sim, ref, noise = _pair('Kuramoto', 'Kuramoto', 0.1, {'a': 1.5},
                            launch=launch, steps=100)
    rng = np.random.default_rng(4)
    x0 = sim.get_state()
    far = rng.random(x0.shape) < 0.3
    x0 = np.where(far, x0 + 2 * np.pi * rng.integers(20, 200, x0.shape),
                  x0).astype(np.float32)
    sim.set_state(x0)
    ref.set_state(x0)
    sim.init_history('current')
    ref.init_history('current')
    sim.run(100, noise=noise)
    ref.run(100, noise=noise)
    xg, xr = sim.get_state(), ref.get_state()
    assert np.isfinite(xg).all()
    assert state_err('Kuramoto', xg, xr) < TOL


@pytest.mark.parametrize('model,cfun,dt,cparams', CASES[:4])
def test_trajectory_parity_deterministic(model, cfun, dt, cparams):
    """Without noise; 500 steps, because the delayed Kuramoto network is
    chaotic enough deterministically that last-bit differences of sinf grow
    past the bar by step 1000 (8e-5 of 2 pi)."""
    sim, ref, _ = _pair(model, cfun, dt, cparams, em=False,
                        launch=dict(waves=2, node_block=3), zero_delays=5)
    sim.run(500)
    ref.run(500)
    assert state_err(model, sim.get_state(), ref.get_state()) < TOL


def test_ring_contents_match_oracle():
    sim, ref, noise = _pair('Generic2dOscillator', 'Sigmoidal', 0.1,
                            {'a': 0.5}, ring_pre=False, steps=300)
    sim.run(300, noise=noise)
    ref.run(300, noise=noise)
    rg, rr = sim.get_ring(), ref.get_ring()
    assert rg.shape == rr.shape
    assert scaled_err(rg, rr) < TOL


def test_philox_mode_matches_oracle():
    """Counter-based noise: the same (step, node, instance, svar) counters on
    both sides; normals differ only by the rounding of log / sin / cos."""
    sim, ref, _ = _pair('Generic2dOscillator', 'Sigmoidal', 0.1, {'a': 0.5})
    sim.set_noise(seed=12345)
    sim.run(1000)
    ref.run(1000, seed=12345)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL


@pytest.mark.parametrize('launch', [
    dict(ipl=1, waves=4, node_block=4),
    dict(ipl=2, waves=8, node_block=10 ** 6, steps_per_launch=25)])
def test_graph_windows_match_eager_launches(launch):
    """hipGraph windows (step counter on the device) reproduce eager
    launches bit for bit over several replays, with host reads, monitor
    samples and a step reset between them."""
    model, cfun, dt, cp = 'Generic2dOscillator', 'Sigmoidal', 0.1, {'a': 0.5}
    simG, _, _ = _pair(model, cfun, dt, cp, P=200, launch=launch)
    simE, _, _ = _pair(model, cfun, dt, cp, P=200, launch=launch)
    simG.set_exec(graph=True, window=50)
    simE.set_exec(graph=False)
    for s in (simG, simE):
        s.set_noise(seed=3)
    for rep in range(4):
        simG.run(150)   # three replays of the 50-step window
        simE.run(150)
        np.testing.assert_array_equal(simG.get_state(), simE.get_state())
        assert simG.step == simE.step == 150 * (rep + 1)
    np.testing.assert_array_equal(simG.get_ring(), simE.get_ring())
    # restart from another step: the device clock follows the host
    for s in (simG, simE):
        s.step = 17
    simG.run(120)
    simE.run(120)
    np.testing.assert_array_equal(simG.get_state(), simE.get_state())


def _block_len(dt, cap=8):
    "longest temporal block the small test connectome allows (<= cap)"
    return min(int(small_csr(N=24, dt=dt).delays.min()) + 1, cap)


@pytest.mark.parametrize('model,cfun,dt,cparams', CASES)
@pytest.mark.parametrize('ipl', [1, 4])
def test_temporal_blocking_matches_per_step_launches(model, cfun, dt, cparams,
                                                     ipl):
    """Launches of d_min + 1 steps with one node per wave and no
    synchronisation reproduce one-step launches bit for bit (same arithmetic,
    state kept in registers instead of HBM), and the oracle within TOL."""
    L = _block_len(dt)
    assert L > 1
    P = 300
    simA, ref, noise = _pair(model, cfun, dt, cparams, P=P, steps=600,
                             launch=dict(ipl=ipl, waves=4, node_block=4))
    simB, _, _ = _pair(model, cfun, dt, cparams, P=P, steps=600,
                       launch=dict(ipl=ipl, waves=4, node_block=4,
                                   steps_per_launch=L))
    assert simB.blocked and simB.H == simA.H + L - 1
    simA.run(600, noise=noise)
    simB.run(600, noise=noise)
    np.testing.assert_array_equal(simB.get_state(), simA.get_state())
    ref.run(600, noise=noise)
    assert state_err(model, simB.get_state(), ref.get_state()) < TOL
    # Philox noise through graph windows (window not a multiple of L: the
    # window is rounded to one)
    for s in (simA, simB):
        s.set_noise(seed=5)
    simB.set_exec(graph=True, window=50)
    simA.set_exec(graph=False)
    simA.run(3 * 56)
    simB.run(3 * 56)
    np.testing.assert_array_equal(simB.get_state(), simA.get_state())
    assert not simB.flags().any()


@pytest.mark.parametrize('model,cfun,dt,cparams', CASES[:4])
@pytest.mark.parametrize('launch', [
    dict(ipl=1, waves=4, node_block=4, unroll=4),
    dict(ipl=4, waves=1, node_block=1, steps_per_launch='L', unroll=4,
         min_waves=4)])
def test_pipelined_gathers_match_grouped(model, cfun, dt, cparams, launch):
    """Issuing the next group of gathers before consuming the current one
    changes only when loads are issued: bit-identical results."""
    launch = dict(launch)
    if launch.get('steps_per_launch') == 'L':
        launch['steps_per_launch'] = _block_len(dt)
    P = 256
    simA, _, noise = _pair(model, cfun, dt, cparams, P=P, steps=400,
                           launch=launch)
    simB, _, _ = _pair(model, cfun, dt, cparams, P=P, steps=400,
                       launch=dict(launch, pipe=True))
    assert simB.spec.pipe and not simA.spec.pipe
    simA.run(400, noise=noise)
    simB.run(400, noise=noise)
    np.testing.assert_array_equal(simB.get_state(), simA.get_state())


@pytest.mark.parametrize('unroll', [16, 32])
@pytest.mark.parametrize('pipe', [False, True])
def test_remainder_group_parity(unroll, pipe):
    """Rows shorter than a gather group (the small test connectome has
    rows of a few to ~20 edges): every row ends in a partial group, issued
    whole with weight-0 slots; the sums match the oracle and are
    bit-identical to 4-edge groups, whose remainders are shorter."""
    model, cfun, dt, cp = 'JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}
    L = _block_len(dt)
    launch = dict(ipl=1, waves=2, node_block=2, steps_per_launch=L,
                  min_waves=4, unroll=unroll, pipe=pipe)
    simA, ref, noise = _pair(model, cfun, dt, cp, P=128, steps=300,
                             launch=launch)
    simB, _, _ = _pair(model, cfun, dt, cp, P=128, steps=300,
                       launch=dict(launch, unroll=4))
    for s in (simA, simB, ref):
        s.run(300, noise=noise)
    np.testing.assert_array_equal(simA.get_state(), simB.get_state())
    assert state_err(model, simA.get_state(), ref.get_state()) < TOL


@pytest.mark.parametrize('model,cfun,dt,cparams', CASES[:4])
def test_wave_priority_changes_nothing_but_order(model, cfun, dt, cparams):
    """Progress-ordered wave priorities in temporally blocked launches change
    only which wave issues first: bit-identical to equal priorities, and to
    the oracle within the usual bar."""
    launch = dict(ipl=1, waves=2, node_block=2,
                  steps_per_launch=_block_len(dt), min_waves=4, pipe=True)
    simA, ref, noise = _pair(model, cfun, dt, cparams, P=256, steps=400,
                             launch=dict(launch, prio=False))
    simB, _, _ = _pair(model, cfun, dt, cparams, P=256, steps=400,
                       launch=launch)
    assert simB.spec.prio and not simA.spec.prio
    simA.run(400, noise=noise)
    simB.run(400, noise=noise)
    ref.run(400, noise=noise)
    np.testing.assert_array_equal(simB.get_state(), simA.get_state())
    assert state_err(model, simB.get_state(), ref.get_state()) < TOL


@pytest.mark.parametrize('P', [512, 96])
@pytest.mark.parametrize('blocked', [False, True])
def test_xcd_block_map_matches_identity(P, blocked):
    """The XCD-aware workgroup numbering only permutes the blocks: results
    are bit for bit those of the identity numbering (P = 96 gives a block
    count that is not a multiple of 8, where the numbering stays the
    identity)."""
    model, cfun, dt, cp = 'JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}
    L = _block_len(dt) if blocked else 1
    out = []
    for xm in (False, True):
        sim, ref, noise = _pair(model, cfun, dt, cp, P=P, steps=300,
                                launch=dict(waves=4, node_block=4,
                                            steps_per_launch=L,
                                            xcd_map=xm))
        sim.run(300, noise=noise)
        out.append(sim.get_state())
    np.testing.assert_array_equal(out[0], out[1])
    ref.run(300, noise=noise)
    assert state_err(model, out[1], ref.get_state()) < TOL


def test_temporal_blocking_monitor_samples():
    """tavg + BOLD samples and FC of blocked launches equal those of
    one-step launches."""
    dt = 0.1
    L = _block_len(dt)
    sims = []
    for launch in (dict(waves=4, node_block=4),
                   dict(waves=4, node_block=4, steps_per_launch=L)):
        sim, _ = _monitor_pair(fc=False, slots=100, launch=launch)
        sim.run(2000)
        sims.append(sim)
    a, fa = sims[0].samples()
    b, fb = sims[1].samples()
    assert fa == fb == 0
    np.testing.assert_array_equal(a, b)
    np.testing.assert_array_equal(sims[0].get_state(), sims[1].get_state())


def test_temporal_blocking_validation():
    dt = 0.1
    csr = small_csr(N=24, dt=dt)
    dmin = int(csr.delays.min())
    m = MODELS['Linear']()
    c = CFUNS['Linear'](m)
    sc = SC.EulerStep(dt)
    with pytest.raises(ValueError):    # a block longer than d_min + 1
        Simulation(m, c, sc, csr, 64, waves=4, node_block=4,
                   steps_per_launch=dmin + 2)
    with pytest.raises(ValueError):    # not one node per wave
        Simulation(m, c, sc, csr, 64, waves=4, node_block=8,
                   steps_per_launch=2)
    with pytest.raises(ValueError):    # the loopy ordering
        Simulation(m, c, sc, csr, 64, waves=4, node_block=4,
                   steps_per_launch=2, order='reference')
    zd = small_csr(N=24, dt=dt, zero_delays=3)
    with pytest.raises(ValueError):    # zero delays
        Simulation(m, c, sc, zd, 64, waves=4, node_block=4,
                   steps_per_launch=2)
    # the C ABI checks the delays against the horizon on its own
    sim = Simulation(m, c, sc, csr, 64, waves=4, node_block=4,
                     steps_per_launch=dmin + 1)
    assert sim.H == csr.max_delay + dmin + 2


def test_rng_normals_match_oracle(ctx):
    z = ctx.rng_normals(1 << 16, seed=99)
    zr = O.normals_f32(1 << 16, seed=99)
    np.testing.assert_allclose(z, zr, rtol=2e-6, atol=2e-6)


def test_instance_independence_and_determinism():
    """Instance p computes the same bits whatever else runs beside it, and
    two runs from the same state agree bit for bit."""
    model, cfun, dt, cp = 'JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}
    simA, _, noise = _pair(model, cfun, dt, cp, P=200, steps=200)
    x0 = simA.get_state()
    simA.run(200, noise=noise)
    xa = simA.get_state()
    simB, _, _ = _pair(model, cfun, dt, cp, P=64, steps=200)
    simB.set_state(x0[..., :64])
    simB.init_history('zeros')
    simB.run(200, noise=noise[..., :64])
    np.testing.assert_array_equal(simB.get_state(), xa[..., :64])
    simA.set_state(x0)
    simA.step = 0
    simA.init_history('zeros')
    simA.run(200, noise=noise)
    np.testing.assert_array_equal(simA.get_state(), xa)


def test_delay_groups_match_oracle():
    """Several conduction speeds: one delay row per group of 64 instances."""
    dt = 0.1
    csr = small_csr(N=20, dt=dt, speeds=(1.0, 2.5, 7.0))
    P, gs = 192, 64
    m = MODELS['Generic2dOscillator']()
    c = CFUNS['Sigmoidal'](m)
    G = np.linspace(0, 1, P).astype(np.float32)
    sim = Simulation(m, c, SC.EulerStep(dt), csr, P, params={'c:a': G},
                     group_size=gs, waves=4, node_block=5)
    x0 = m.initial_state(20, P, np.random.default_rng(3))
    sim.set_state(x0)
    sim.init_history('zeros')
    ref = oracle_for('Generic2dOscillator', 'Sigmoidal', csr, P, sim.H, dt,
                     'euler', cparams={'a': G}, group_size=gs)
    ref.set_state(x0)
    ref.init_history('zeros')
    sim.run(800)
    ref.run(800)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL


def _monitor_pair(fc, slots, launch=None):
    dt = 0.1
    csr = small_csr(N=16, dt=dt)
    P = 64
    m = MODELS['Generic2dOscillator']()
    c = CFUNS['Sigmoidal'](m)
    sc = SC.EulerMaryuyamaStep(dt, sigma=0.005)
    G = np.linspace(0, 1, P).astype(np.float32)
    sim = Simulation(m, c, sc, csr, P, params={'c:a': G}, tavg_period=20,
                     bold=True, fc=fc, sample_slots=slots,
                     **(launch or dict(waves=4)))
    x0 = m.initial_state(16, P, np.random.default_rng(5))
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.set_noise(seed=7)
    ref = oracle_for('Generic2dOscillator', 'Sigmoidal', csr, P, sim.H, dt,
                     'em', cparams={'a': G}, sigma=0.005, tavg_period=20,
                     bold=True)
    ref.set_state(x0)
    ref.init_history('zeros')
    return sim, ref


def test_monitor_bold_samples_match_oracle():
    """tavg -> Balloon-Windkessel samples of the device match the oracle's,
    with identical injected noise.  (With Philox noise the normals differ in
    their last bit -- device sincospif vs. host cos -- and a few excitable
    instances of this network amplify that into different spikes.)"""
    sim, ref = _monitor_pair(fc=False, slots=100)
    noise = np.random.default_rng(9).standard_normal(
        (2000, 2, 16, 64)).astype(np.float32)
    sim.run(2000, noise=noise)
    out = ref.run(2000, noise=noise)
    s_dev, first = sim.samples()
    assert first == 0 and s_dev.shape == (100, 16, 64)
    s_ref = out['samples']
    # BOLD = V0 (k1 (1 - q) + k2 (1 - q/v) + k3 (1 - v)) with q, v ~ 1 and
    # V0 k1 ~ 11: one float32 ulp of q (1.2e-7, from powf on either side)
    # moves the signal by 1.3e-6, whatever the signal's own size.
    err = np.abs(s_dev - s_ref)
    print('max |dBOLD| %.2e of max |BOLD| %.2e' % (err.max(),
                                                   np.abs(s_ref).max()))
    np.testing.assert_array_equal(sim.get_state(), ref.get_state())
    assert err.max() < 4e-6
    assert err.max() < 2e-3 * np.abs(s_ref).max()


def test_device_fc_matches_fp64_fc_of_samples():
    """FP64 co-moments folded in blocks of 7 samples (Chan et al. merges)
    give the correlation of the same samples computed by the oracle's
    one-pass FP64 Welford and by numpy.corrcoef."""
    sim, ref = _monitor_pair(fc=True, slots=7)
    sim2, _ = _monitor_pair(fc=False, slots=100)
    sim.run(2000)
    sim2.run(2000)
    fc = sim.fc()
    assert sim.n_fc_samples == 100
    s_dev, _ = sim2.samples()
    s64 = s_dev.astype(np.float64)
    fc_ref = O.fc_f64(s64)
    assert np.abs(fc - fc_ref).max() < 2e-6
    for p in (0, 17, 63):
        np.testing.assert_allclose(np.corrcoef(s64[:, :, p].T), fc_ref[p],
                                   atol=1e-9)


def test_corr_kernel_matches_numpy(ctx):
    rng = np.random.default_rng(0)
    x = rng.standard_normal((300, 12, 70)).astype(np.float32)
    x += 0.5 * x[:, :1]
    c = ctx.corr(x)
    for p in (0, 33, 69):
        np.testing.assert_allclose(c[p], np.corrcoef(x[:, :, p].T), atol=2e-6)


def test_validation_errors():
    csr = small_csr(N=10)
    m = MODELS['Linear']()
    c = CFUNS['Linear'](m)
    with pytest.raises(ValueError):
        Simulation(m, c, SC.EulerStep(0.1), csr, 64,
                   horizon=csr.max_delay + 1)
    sim = Simulation(m, c, SC.EulerStep(0.1), csr, 64)
    from tvb_hpc import hip
    bad_row = csr.row.copy()
    bad_row[-1] += 1
    with pytest.raises(TvbxError, match='row'):
        hip.check(hip.lib().tvbx_sim_set_connectome(
            sim.h, bad_row.ctypes.data, csr.col.ctypes.data,
            csr.weight.ctypes.data, csr.delays.ctypes.data))
    big = csr.delays.copy()
    big[0, 0] = sim.H - 1
    with pytest.raises(TvbxError, match='horizon'):
        hip.check(hip.lib().tvbx_sim_set_connectome(
            sim.h, csr.row.ctypes.data, csr.col.ctypes.data,
            csr.weight.ctypes.data, big.ctypes.data))
    bad_col = csr.col.copy()
    bad_col[3] = 10
    with pytest.raises(TvbxError, match='col'):
        hip.check(hip.lib().tvbx_sim_set_connectome(
            sim.h, csr.row.ctypes.data, bad_col.ctypes.data,
            csr.weight.ctypes.data, csr.delays.ctypes.data))
    nanp = np.full(64, np.nan, np.float32)
    with pytest.raises(TvbxError, match='finite'):
        sim2 = Simulation(m, c, SC.EulerStep(0.1), csr, 64,
                          params={'m:lam': nanp}, kinds={'m:lam': 'inst'})
        del sim2


@pytest.mark.parametrize('blocked', [False, True])
def test_nonfinite_flag_marks_only_bad_instance(blocked):
    """Per-step launches check every step; temporally blocked launches
    check the state at the end of each launch (a non-finite state stays
    non-finite under the unforced update)."""
    dt = 0.1
    csr = small_csr(N=12, dt=dt)
    P = 64
    m = MODELS['Linear']()
    c = CFUNS['Linear'](m)
    lam = np.full(P, -1.0, np.float32)
    lam[5] = 1e30  # blows up instance 5 only
    L = min(int(csr.delays.min()) + 1, 8)
    launch = dict(waves=1, node_block=1, steps_per_launch=L) if blocked \
        else {}
    sim = Simulation(m, c, SC.EulerStep(dt), csr, P,
                     params={'m:lam': lam, 'c:a': 0.01},
                     kinds={'m:lam': 'inst'}, **launch)
    sim.set_state(np.ones((1, 12, P), np.float32))
    sim.init_history('constant')
    sim.run(50)
    f = sim.flags()
    assert f[5] == 1 and f.sum() == 1
    x = sim.get_state()
    assert np.isfinite(np.delete(x, 5, axis=2)).all()


@pytest.mark.parametrize('blocked', [False, True])
def test_reference_rww_singularity_is_flagged(blocked):
    """The reference RWW's h = (a x - b) / (1 - exp(-d (a x - b))) is 0/0 when
    a*x rounds to b; the device flags exactly the instances that hit it,
    the others stay finite, and ReducedWongWang has no such point."""
    launch = dict(waves=1, node_block=1, steps_per_launch=_block_len(0.1)) \
        if blocked else None
    sim, ref, _ = _pair('RWW', 'Linear', 0.1, {'a': 0.05}, em=False,
                        launch=launch)
    sim.run(1000)
    x = sim.get_state()
    f = sim.flags()
    bad = ~np.isfinite(x).all(axis=(0, 1))
    np.testing.assert_array_equal(f.astype(bool), bad)
    sim2, _, _ = _pair('ReducedWongWang', 'Linear', 0.1, {'a': 0.05},
                       em=False)
    sim2.run(1000)
    assert np.isfinite(sim2.get_state()).all() and not sim2.flags().any()


def _group_sparse_csr(n_subj=3, N=18, dt=0.1, seed=21):
    """subjects sharing the pattern of edges present in all of them, with
    their own weights and lengths (examples/hcp100.py:22-53)"""
    rng = np.random.default_rng(seed)
    base = rng.random((N, N)) < 0.45
    W = np.stack([np.where(base, rng.uniform(0.1, 1.0, (N, N)), 0.0)
                  for _ in range(n_subj)])
    L = np.stack([rng.uniform(1.0, 40.0, (N, N)) for _ in range(n_subj)])
    return CSR.group_sparse(W, L, [1.0, 2.0, 3.5][:n_subj], dt)


@pytest.mark.parametrize('launch', [dict(waves=4, node_block=5),
                                    'blocked'])
def test_grouped_weights_match_oracle(launch):
    """One weight row and one delay row per group of 64 instances (subjects
    on a shared pattern), per-step and temporally blocked launches."""
    dt = 0.1
    csr = _group_sparse_csr(dt=dt)
    assert csr.weight_groups and csr.weight.shape == csr.delays.shape
    N, P, gs = csr.nnode, 192, 64
    if launch == 'blocked':
        L = min(int(csr.delays.min()) + 1, 4)
        assert L >= 2
        launch = dict(waves=2, node_block=2, steps_per_launch=L)
    m = MODELS['Generic2dOscillator']()
    c = CFUNS['Sigmoidal'](m)
    G = np.linspace(0, 1, P).astype(np.float32)
    sim = Simulation(m, c, SC.EulerStep(dt), csr, P, params={'c:a': G},
                     group_size=gs, **launch)
    x0 = m.initial_state(N, P, np.random.default_rng(4))
    sim.set_state(x0)
    sim.init_history('zeros')
    ref = oracle_for('Generic2dOscillator', 'Sigmoidal', csr, P, sim.H, dt,
                     'euler', cparams={'a': G}, group_size=gs)
    ref.set_state(x0)
    ref.init_history('zeros')
    sim.run(600)
    ref.run(600)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL


@pytest.mark.parametrize('ipl', [1, 2, 4])
@pytest.mark.parametrize('big_ring', [False, True])
@pytest.mark.parametrize('launch', [dict(waves=4, node_block=4),
                                    dict(waves=1, node_block=1,
                                         steps_per_launch=4, pipe=True)])
def test_long_rows_match_oracle(ipl, big_ring, launch):
    """Rows of 150-250 edges, past the two 64-edge register rows: the rest
    is gathered in 64-edge chunks loaded at every step.  big_ring: a
    horizon of 60,000 slots makes the ring larger than 4 GiB, so gathers take
    the 64-bit address path (the oracle, with its own short ring, computes
    the same values)."""
    N, P, steps, dt = 300, 64 * ipl, 200, 0.1
    csr = small_csr(N=N, density=0.7, seed=23, dt=dt, lmax=40.0)
    deg = np.diff(csr.row)
    assert deg.max() > 192 and deg.min() > 128
    m = MODELS['ReducedWongWang']()
    c = CFUNS['Linear'](m)
    scheme = SC.EulerMaryuyamaStep(dt)
    la = dict(launch, ipl=ipl)
    if la.get('steps_per_launch', 1) > 1:
        assert int(csr.delays.min()) >= la['steps_per_launch'] - 1
    H = 60000 if big_ring else None
    sim = Simulation(m, c, scheme, csr, P, horizon=H,
                     params={'c:a': np.full(P, 0.02, np.float32),
                             'sigma': np.full(P, 0.01, np.float32)},
                     kinds={'sigma': 'inst'}, **la)
    ring_bytes = sim.H * N * sim.pitch * 4
    assert (ring_bytes >= 4 << 30) == big_ring
    rng = np.random.default_rng(3)
    x0 = m.initial_state(N, P, rng)
    sim.set_state(x0)
    sim.init_history('zeros')
    noise = rng.standard_normal((steps, 1, N, P)).astype(np.float32)
    sim.run(steps, noise=noise)
    ref = oracle_for('ReducedWongWang', 'Linear', csr, P, csr.max_delay + 2,
                     dt, 'em', cparams={'a': 0.02}, sigma=0.01)
    ref.set_state(x0)
    ref.init_history('zeros')
    ref.run(steps, noise=noise)
    err = scaled_err(sim.get_state(), ref.get_state())
    print('long rows (%d-%d edges) %s big_ring=%s: %.2e'
          % (deg.min(), deg.max(), la, big_ring, err))
    assert err < TOL


@pytest.mark.parametrize('ring_dtype', ['bf16', 'f16'])
@pytest.mark.parametrize('launch', [
    dict(ipl=1, waves=4, node_block=4, steps_per_launch=8),
    dict(ipl=4, waves=2, node_block=2, steps_per_launch=1)])
def test_mfma_coupling_parity(ring_dtype, launch):
    """Delay-binned sums on the matrix cores (16-target blocks, 32-pair
    chunks, weights as high + low 16-bit parts) against the float32 oracle
    emulating the ring's rounding, with injected noise: within 1e-4, as the
    other reduced-precision rings (the products are exact; the sums are
    accumulated in another order and the split weights carry 16 (bf16) or
    22 (f16) significant bits)."""
    sim, ref, noise = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3},
                            P=200, steps=600, ring_dtype=ring_dtype,
                            ring_scale=200.0, launch=launch)
    assert int(sim.csr.delays.min()) >= launch['steps_per_launch'] - 1
    sim.set_coupling('mfma')
    sim.run(600, noise=noise)
    ref.run(600, noise=noise)
    err = state_err('JansenRit', sim.get_state(), ref.get_state())
    print('JR %s ring, MFMA coupling %s vs emulating oracle: %.2e'
          % (ring_dtype, launch, err))
    assert err < 1e-4
    assert not sim.flags().any()


def test_mfma_coupling_graphs_equal_eager():
    """Captured graph windows replay the two-kernel (sums, advance) launches
    bit for bit."""
    out = []
    for graph in (False, True):
        sim, ref, noise = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3},
                                P=128, steps=10, em=False, ring_dtype='bf16',
                                ring_scale=200.0,
                                launch=dict(waves=4, node_block=4,
                                            steps_per_launch=8))
        sim.set_coupling('mfma')
        sim.set_exec(graph=graph, window=16)
        sim.run(8 * 13 + 5)
        out.append(sim.get_state())
    np.testing.assert_array_equal(out[0], out[1])


def test_mfma_coupling_refused_where_it_does_not_apply():
    sim, _, _ = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}, P=64,
                      steps=1, launch=dict(waves=4, node_block=4,
                                           steps_per_launch=4))
    with pytest.raises(TvbxError):
        sim.set_coupling('mfma')     # float32 ring: no matrix-core kernel


@pytest.mark.parametrize('ring_dtype,ipl', [('f32', 1), ('bf16', 4),
                                            ('bf16', 2), ('u8', 4)])
def test_edge_major_blocked_equals_step_major(ring_dtype, ipl):
    """Blocked launches that form all their steps' coupling sums first
    (edges outer, steps inner) give the state of the step-by-step blocked
    launches bit for bit, including a launch shorter than the block."""
    out = []
    for emaj in (0, 8):
        sim, _, _ = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}, P=256,
                          steps=1, em=False, ring_dtype=ring_dtype,
                          ring_scale=1.0 if ring_dtype == 'f32' else 200.0,
                          launch=dict(ipl=ipl, waves=2, node_block=2,
                                      steps_per_launch=8, emaj=emaj))
        assert int(sim.csr.delays.min()) >= 7
        sim.run(8 * 10 + 3)
        out.append(sim.get_state())
        assert not sim.flags().any()
    np.testing.assert_array_equal(out[0], out[1])


def test_edge_major_long_rows_equal_step_major():
    """The same for rows of 188-232 edges (chunks past the register rows)."""
    N, P, dt = 300, 64, 0.1
    csr = small_csr(N=N, density=0.7, seed=23, dt=dt, lmax=40.0)
    m = MODELS['ReducedWongWang']()
    c = CFUNS['Linear'](m)
    out = []
    for emaj in (0, 4):
        sim = Simulation(m, c, SC.EulerStep(dt), csr, P,
                         params={'c:a': np.full(P, 0.02, np.float32)},
                         waves=1, node_block=1, steps_per_launch=4,
                         emaj=emaj)
        x0 = m.initial_state(N, P, np.random.default_rng(3))
        sim.set_state(x0)
        sim.init_history('zeros')
        sim.run(4 * 25 + 2)
        out.append(sim.get_state())
    np.testing.assert_array_equal(out[0], out[1])


@pytest.mark.parametrize('ipl,xcd', [(1, False), (2, True)])
def test_pacing_changes_nothing(ipl, xcd):
    """Workgroups of one XCD waiting for each other after every step of a
    blocked launch only change the timing: the state equals the unpaced
    run's bit for bit, over several launches (the counters must come back
    to zero after each) and a shorter last launch."""
    out = []
    for pace in (False, True):
        sim, _, _ = _pair('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}, P=256,
                          steps=1, em=False, ring_dtype='bf16',
                          ring_scale=200.0,
                          launch=dict(ipl=ipl, waves=2, node_block=2,
                                      steps_per_launch=8, xcd_map=xcd,
                                      pace=pace))
        sim.run(8 * 12 + 5)
        out.append(sim.get_state())
        assert not sim.flags().any()
    np.testing.assert_array_equal(out[0], out[1])

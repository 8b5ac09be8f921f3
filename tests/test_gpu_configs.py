"""Parity at the five sweep workloads' own sizes, on the device, through
libtvbx's C ABI: each workload at its full instance count, connectome
generator, dt, delays and tuned launch (``autotune.best_launch``), with a
slice of instances compared against the CPU oracle.

* trajectories: float32 oracle, identical injected noise (or the Philox
  stream, for the batch), 1000 steps, scaled error below 1e-5
  (``harness.state_err``);
* reduced-precision histories: the float32 oracle emulating the same
  rounding, 1e-4;
* FC after the full 10,000-step Jansen-Rit run against the float64 oracle
  with exact delays, 1e-3;
* the MPR adjoint with checkpoints every 1000 steps against float64 central
  differences, 1e-4.

The step being checked is tvb_hpc/network.py:115-136 (delayed CSR
coupling) + model.py:105-140 (dfun) + scheme.py:88-97 (Euler-Maruyama).
These sizes reach kernel paths the small tests do not: rings of 4 GiB and
more (64-bit gather addresses; plain Jansen-Rit P=1024, G2DO P=8192),
rows of 64-127 edges (Jansen-Rit N=998) and of more than 128 edges (the
N=400 subjects of the uneven batch).
"""

import numpy as np
import pytest

import oracle as O
from harness import scaled_err, state_err
from tvb_hpc import autotune, configs

pytestmark = pytest.mark.gpu

TOL = 1e-5
STEPS = 1000


def _injected_run(sim, cfg, slices, steps=STEPS, chunk=100, seed=77):
    """Run `steps` steps with injected normals; instances outside `slices`
    get zeros.  Returns the slices' noise (steps, S, N, n) per slice."""
    S, N, P = sim.spec.S, cfg.csr.nnode, cfg.n_inst
    rng = np.random.default_rng(seed)
    per = [rng.standard_normal((steps, S, N, sl.stop - sl.start)).astype(
        np.float32) for sl in slices]
    buf = np.zeros((chunk, S, N, P), np.float32)
    for k0 in range(0, steps, chunk):
        for sl, nz in zip(slices, per):
            buf[:, :, :, sl] = nz[k0:k0 + chunk]
        sim.run(chunk, noise=buf)
    return per


def _ring_bytes(sim, cfg):
    esz = {'f32': 4, 'bf16': 2, 'f16': 2, 'u8': 1}[cfg.ring_dtype]
    return sim.H * sim.spec.R * cfg.csr.nnode * sim.pitch * esz


def test_config0_kuramoto76(ctx):
    cfg = configs.kuramoto76()
    la = autotune.best_launch(cfg, ctx)
    sim = cfg.simulation(**la)
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    sl = slice(0, cfg.n_inst)
    (nz, ) = _injected_run(sim, cfg, [sl])
    ref = O.for_config(cfg, np.arange(cfg.n_inst))
    ref.run(STEPS, noise=nz)
    err = state_err('Kuramoto', sim.get_state(), ref.get_state())
    print('kuramoto76 %s: %.2e' % (la, err))
    assert err < TOL
    assert not sim.flags().any()


def test_config1_g2do84_full_sweep(ctx):
    """P = 8192 (64 couplings x 128 speeds, one delay row per speed): the
    ring is over 4 GiB.  First, middle and last delay groups against the
    oracle, with their BOLD samples and the FC of those samples."""
    cfg = configs.g2do84()
    la = autotune.best_launch(cfg, ctx)
    sim = cfg.simulation(**la)
    assert _ring_bytes(sim, cfg) >= 4 << 30
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    gs = cfg.group_size
    slices = [slice(0, gs), slice(64 * gs, 65 * gs),
              slice(cfg.n_inst - gs, cfg.n_inst)]
    noise = _injected_run(sim, cfg, slices)
    x = sim.get_state()
    smp, _ = sim.samples()
    for sl, nz in zip(slices, noise):
        ref = O.for_config(cfg, np.arange(sl.start, sl.stop))
        out = ref.run(STEPS, noise=nz)
        err = state_err('Generic2dOscillator', x[..., sl], ref.get_state())
        ref_s = out['samples']
        assert smp.shape[0] == ref_s.shape[0] == STEPS // cfg.tavg_period
        # BOLD = V0 (k1 (1 - q) + k2 (1 - q/v) + k3 (1 - v)) with q, v ~ 1
        # and V0 k1 ~ 11: one float32 ulp of q (from powf on either side)
        # moves the signal by 1.3e-6 whatever its size (here below 1e-3,
        # where float32 and float64 BOLD differ by ~1e-5): an absolute bound
        err_b = float(np.abs(smp[..., sl] - ref_s).max())
        print('g2do84 %s instances %d-%d: state %.2e, |dBOLD| %.2e of '
              'max |BOLD| %.2e' % (la, sl.start, sl.stop - 1, err, err_b,
                                   np.abs(ref_s).max()))
        assert err < TOL and err_b < 4e-6
    fc = sim.fc(slices[-1].start, 4)
    fc_ref = O.fc_f64(smp[:, :, slices[-1].start:slices[-1].start + 4]
                      .astype(np.float64))
    assert np.abs(fc - fc_ref).max() < 1e-3
    assert not sim.flags().any()


def test_config2_rww_uneven_batch(ctx):
    """The uneven batch (32 subjects, N in {68 .. 400}, P = 2048) in one
    launch with its tuned shape and Philox noise; the first 8 instances of
    every subject against the oracle (instance offsets of the batch-wide
    numbering)."""
    cfg = configs.rww_uneven()
    la = autotune.best_batch_launch(cfg, ctx)
    b = cfg.batch(**la)
    x0 = cfg.initial_states()
    for k, x in enumerate(x0):
        b.set_state(k, x)
    b.init_history('zeros')
    b.set_noise(seed=21)
    b.run(STEPS)
    rowmax = max(int(np.diff(c.row).max()) for c, _ in cfg.subjects)
    assert rowmax > 128
    tile = 64 * la.get('ipl', 1)
    off, i0 = 0, []
    for _, p in cfg.subjects:
        i0.append(off)
        off += (p + tile - 1) // tile * tile
    poff = np.cumsum([0] + [p for _, p in cfg.subjects])
    worst = 0.0
    for k, (csr, p) in enumerate(cfg.subjects):
        n = 8
        g = np.arange(poff[k], poff[k] + n)
        ref = O.OracleSim('ReducedWongWang', 'Linear', csr.nnode, csr.row,
                          csr.col, csr.weight, csr.delays, n,
                          csr.max_delay + 2, cfg.scheme.dt, scheme='em',
                          params={'J': np.asarray(cfg.params['m:J'])[g]},
                          cparams={'a': np.asarray(cfg.params['c:a'])[g]},
                          sigma=np.asarray(cfg.params['sigma'])[g],
                          dtype=np.float32, inst_off=i0[k])
        ref.set_state(x0[k][..., :n])
        ref.init_history('zeros')
        ref.run(STEPS, seed=21)
        err = scaled_err(b.get_state(k)[..., :n], ref.get_state())
        worst = max(worst, err)
        assert err < TOL, (k, csr.nnode, int(np.diff(csr.row).max()), err)
    print('rww_uneven %s: worst subject %.2e (longest row %d edges)'
          % (la, worst, rowmax))
    assert not b.flags().any()


def _jr(name):
    return configs.ALL[name]()


JR_SLICE = slice(496, 512)


@pytest.mark.parametrize('name,tol', [('jansenrit998', TOL),
                                      ('jansenrit998_bf16', 1e-4),
                                      ('jansenrit998_f16', 1e-4),
                                      ('jansenrit998_u8', 1e-4),
                                      ('jansenrit998_bf16_mfma', 1e-4),
                                      ('jansenrit998_f16_mfma', 1e-4)])
def test_config3_jansenrit998_trajectories(ctx, name, tol):
    """N = 998, P = 1024, tuned launch, 1000 steps (deterministic Euler).
    Plain float32: the ring is over 4 GiB, rows hold 64-127 edges.  The
    32-bin reduced-precision histories against the oracle emulating their
    rounding (1e-4: a rounding step can resolve a last-bit difference of
    exp either way)."""
    cfg = _jr(name)
    la = autotune.best_launch(cfg, ctx)
    sim = cfg.simulation(**la)
    if cfg.ring_dtype == 'f32':
        assert _ring_bytes(sim, cfg) >= 4 << 30
    assert int(np.diff(cfg.csr.row).max()) > 64
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    sim.run(STEPS)
    idx = np.arange(JR_SLICE.start, JR_SLICE.stop)
    ref = O.for_config(cfg, idx)
    ref.run(STEPS)
    err = state_err('JansenRit', sim.get_state()[..., JR_SLICE],
                    ref.get_state())
    print('%s %s: %.2e' % (name, la, err))
    assert err < tol
    assert not sim.flags().any()


FC_IDX = np.linspace(0, 1023, 8).astype(int)
FC_STEPS = 10000


@pytest.fixture(scope='module')
def jr_fc64():
    "float64 oracle FC of FC_IDX over the full run, exact delays"
    cfg = configs.jansenrit998()
    ref = O.for_config(cfg, FC_IDX, np.float64)
    s = ref.run(FC_STEPS)['samples']
    return np.stack([np.corrcoef(s[:, :, k].T) for k in range(len(FC_IDX))])


@pytest.mark.parametrize('name', ['jansenrit998', 'jansenrit998_bf16',
                                  'jansenrit998_f16', 'jansenrit998_u8',
                                  'jansenrit998_bf16_mfma',
                                  'jansenrit998_f16_mfma'])
def test_config3_fc_after_full_run(ctx, jr_fc64, name):
    """FC (FP64 co-moments on the device) after the 10,000-step run with
    graph windows, against the float64 oracle with exact delays."""
    cfg = _jr(name)
    la = autotune.best_launch(cfg, ctx)
    sim = cfg.simulation(**la)
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    sim.set_exec(graph=True, window=100)
    sim.run(FC_STEPS)
    fc = np.stack([sim.fc(int(p), 1)[0] for p in FC_IDX])
    assert sim.n_fc_samples == FC_STEPS // cfg.tavg_period
    err = float(np.abs(fc - jr_fc64).max())
    print('%s FC after %d steps vs float64: %.2e' % (name, FC_STEPS, err))
    assert err < 1e-3


def test_config4_mpr84_forward(ctx):
    cfg = configs.mpr84()
    la = autotune.best_launch(cfg, ctx)
    sim = cfg.simulation(**la)
    sim.set_state(cfg.initial_state())
    sim.init_history('zeros')
    slices = [slice(0, 64), slice(cfg.n_inst - 64, cfg.n_inst)]
    noise = _injected_run(sim, cfg, slices, chunk=250)
    x = sim.get_state()
    for sl, nz in zip(slices, noise):
        ref = O.for_config(cfg, np.arange(sl.start, sl.stop))
        ref.run(STEPS, noise=nz)
        err = state_err('MontbrioPazoRoxin', x[..., sl], ref.get_state())
        print('mpr84 %s instances %d-%d: %.2e' % (la, sl.start, sl.stop - 1,
                                                  err))
        assert err < TOL
    assert not sim.flags().any()


def test_config4_mpr84_adjoint_checkpointed(ctx):
    """Gradients of 1 - corr(FC, target FC) w.r.t. G (P,) and eta (N, P) for
    all 4096 instances over 20,000 steps, checkpoints every 1000 steps
    replayed in the backward pass; a slice of instances against float64
    central differences of the oracle (converged h = 1e-5, checked against
    h / 2)."""
    import torch
    from tvb_hpc import autograd as AG
    cfg = configs.mpr84()
    T, K, N, P = cfg.n_steps, cfg.tavg_period, cfg.csr.nnode, cfg.n_inst
    la = dict(autotune.best_launch(cfg, ctx))
    sim = cfg.simulation(adjoint=True, fc=False, sample_slots=T // K, **la)
    x0 = cfg.initial_state()
    seed = 5
    sl = slice(2048, 2056)
    idx = np.arange(sl.start, sl.stop)
    G = np.asarray(cfg.params['c:a'], np.float64)
    eta0 = float(cfg.model.const['eta'])

    def oracle_loss(Gs, eta_s, fct):
        ref = O.OracleSim('MontbrioPazoRoxin', 'Linear', N, cfg.csr.row,
                          cfg.csr.col, cfg.csr.weight, cfg.csr.delays,
                          idx.size, cfg.csr.max_delay + 2, cfg.scheme.dt,
                          scheme='em', params={'eta': eta_s},
                          cparams={'a': Gs}, sigma=cfg.scheme.sigma,
                          dtype=np.float64, tavg_period=K,
                          inst_off=sl.start)
        ref.set_state(x0[..., sl])
        ref.init_history('zeros')
        s = ref.run(T, seed=seed)['samples']
        if fct is None:
            return s
        return AG.fc_loss(torch.from_numpy(s), fct).numpy()

    eta_s = np.full((N, idx.size), eta0)
    s_t = oracle_loss(np.full(idx.size, 0.5), eta_s, None)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()

    dev = torch.device('cuda:0')
    Gt = torch.tensor(G.astype(np.float32), device=dev, requires_grad=True)
    et = torch.full((N, P), eta0, device=dev, requires_grad=True)
    smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, 1000,
                       seed=seed)
    AG.fc_loss(smp, torch.as_tensor(fct, device=dev)).sum().backward()
    gG = Gt.grad.cpu().numpy()[sl]
    geta = et.grad.cpu().numpy()[:, sl]

    def fd(loss_of, h):
        return (loss_of(+h) - loss_of(-h)) / (2 * h)

    def reference(loss_of):
        d1 = fd(loss_of, 1e-5)
        d2 = fd(loss_of, 5e-6)
        assert np.abs(d1 - d2).max() < 1e-5 * np.abs(d1).max()
        return d1

    fdG = reference(lambda h: oracle_loss(G[sl] + h, eta_s, fct))
    errG = np.abs(gG - fdG).max() / np.abs(fdG).max()
    print('mpr84 %s dL/dG vs float64 central differences: %.2e' % (la, errG))
    assert errG < 1e-4
    for i in (0, 41, 83):
        def loss_i(h, i=i):
            e = eta_s.copy()
            e[i] += h
            return oracle_loss(G[sl], e, fct)
        fdi = reference(loss_i)
        err = np.abs(geta[i] - fdi).max() / np.abs(fdi).max()
        print('mpr84 dL/deta[%d] vs float64 central differences: %.2e'
              % (i, err))
        assert err < 1e-4

"""Gradients of an FC-fit loss through the delayed Euler-Maruyama recursion:
the hand-written adjoint (HIP) against converged central finite differences
of the float64 CPU oracle (h = 1e-5; the h = 1e-3 difference is reported
beside it) -- the float64 reference gradient -- and against float32 device
finite differences (h = 1e-3).

Model: Montbrio-Pazo-Roxin with linear delayed coupling, additive Philox
noise, tavg monitor of r; loss per instance 1 - corr(FC_sim, FC_target)
with the target FC from a run at G = 0.5."""

import numpy as np
import pytest
import torch

import oracle as O
from harness import small_csr
from tvb_hpc import autograd as AG
from tvb_hpc import coupling, model, scheme
from tvb_hpc.hip import Simulation

pytestmark = pytest.mark.gpu

N, P, T, WIN, K, DT, SEED = 12, 64, 400, 100, 10, 0.1, 2024


def _setup():
    csr = small_csr(N=N, density=0.5, seed=11, speeds=(1.0, ), dt=DT,
                    lmax=40.0)
    assert csr.delays.min() >= 1
    m = model.MontbrioPazoRoxin()
    c = coupling.Linear(m)
    sc = scheme.EulerMaryuyamaStep(DT, sigma=0.01)
    rng = np.random.default_rng(5)
    G = rng.uniform(0.2, 0.8, P).astype(np.float32)
    eta = (-4.6 + 0.2 * rng.standard_normal((N, P))).astype(np.float32)
    x0 = np.stack([rng.uniform(0.1, 1.0, (N, P)),
                   rng.uniform(-1.5, -0.5, (N, P))]).astype(np.float32)
    return csr, m, c, sc, G, eta, x0


def _device_sim(csr, m, c, sc, G, eta):
    return Simulation(m, c, sc, csr, P, params={'c:a': G, 'm:eta': eta},
                      kinds={'c:a': 'inst', 'm:eta': 'ni'}, tavg_period=K,
                      sample_slots=T // K, adjoint=True, waves=4,
                      node_block=4)


def _oracle_samples(csr, G, eta, x0):
    sim = O.OracleSim('MontbrioPazoRoxin', 'Linear', N, csr.row, csr.col,
                      csr.weight, csr.delays, P, csr.max_delay + 2, DT,
                      scheme='em', params={'eta': eta}, cparams={'a': G},
                      sigma=0.01, dtype=np.float64, tavg_period=K)
    sim.set_state(x0)
    sim.init_history('zeros')
    return sim.run(T, seed=SEED)['samples']


def _loss64(samples, fct):
    s = torch.from_numpy(np.asarray(samples, np.float64))
    return AG.fc_loss(s, fct).numpy()


def test_adjoint_matches_fp64_finite_differences():
    csr, m, c, sc, G, eta, x0 = _setup()
    # target FC from a run at G = 0.5
    s_t = _oracle_samples(csr, np.full(P, 0.5), eta, x0)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()

    sim = _device_sim(csr, m, c, sc, G, eta)
    Gt = torch.tensor(G, requires_grad=True)
    et = torch.tensor(eta, requires_grad=True)
    smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, WIN, seed=SEED)
    assert smp.shape == (T // K, N, P)
    L = AG.fc_loss(smp, fct).sum()
    L.backward()
    gG = Gt.grad.numpy()
    geta = et.grad.numpy()
    assert np.isfinite(gG).all() and np.isfinite(geta).all()

    # float64 reference gradient: central differences of the oracle.  The
    # loss is strongly nonlinear in some nodes' eta: at h = 1e-3 the central
    # difference still carries an O(h^2) truncation error of up to 2 %, so
    # the reference uses h = 1e-5 (truncation ~2e-6, float64 rounding
    # ~1e-11) and is checked for convergence against h / 2.
    def fd64(loss_of, h):
        return (loss_of(+h) - loss_of(-h)) / (2 * h)

    def reference(loss_of, name):
        d1 = fd64(loss_of, 1e-5)
        d2 = fd64(loss_of, 5e-6)
        assert np.abs(d1 - d2).max() < 1e-5 * np.abs(d1).max()
        d3 = fd64(loss_of, 1e-3)
        print('%s: h = 1e-3 central difference off the converged one by '
              '%.2e' % (name, np.abs(d3 - d1).max() / np.abs(d1).max()))
        return d1

    G64 = G.astype(np.float64)
    fdG = reference(lambda h: _loss64(_oracle_samples(csr, G64 + h, eta, x0),
                                      fct), 'dL/dG')
    errG = np.abs(gG - fdG).max() / np.abs(fdG).max()
    print('dL/dG: adjoint vs fp64 central differences, rel err %.2e '
          '(max |grad| %.3e)' % (errG, np.abs(fdG).max()))
    assert errG < 1e-4
    for i in (0, 5, 11):
        def loss_i(h, i=i):
            e = eta.astype(np.float64).copy()
            e[i] += h
            return _loss64(_oracle_samples(csr, G, e, x0), fct)
        fd = reference(loss_i, 'dL/deta[%d]' % i)
        err = np.abs(geta[i] - fd).max() / np.abs(fd).max()
        print('dL/deta[%d]: adjoint vs fp64 central differences, rel err '
              '%.2e (max |grad| %.3e)' % (i, err, np.abs(fd).max()))
        assert err < 1e-4


def test_adjoint_matches_device_finite_differences():
    """float32 central differences on the device (h = 1e-3) agree with the
    adjoint to the precision float32 differences allow."""
    csr, m, c, sc, G, eta, x0 = _setup()
    s_t = _oracle_samples(csr, np.full(P, 0.5), eta, x0)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()
    sim = _device_sim(csr, m, c, sc, G, eta)
    Gt = torch.tensor(G, requires_grad=True)
    et = torch.tensor(eta, requires_grad=True)
    smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, WIN, seed=SEED)
    AG.fc_loss(smp, fct).sum().backward()
    h = 1e-3

    def dev_loss(Gv):
        s = AG.integrate(sim, {'c:a': torch.tensor(Gv, dtype=torch.float64),
                               'm:eta': torch.tensor(eta)}, x0, T, WIN,
                         seed=SEED)
        return AG.fc_loss(s, fct).detach().numpy()

    fd = (dev_loss(G.astype(np.float64) + h) -
          dev_loss(G.astype(np.float64) - h)) / (2 * h)
    err = np.abs(Gt.grad.numpy() - fd).max() / np.abs(fd).max()
    print('dL/dG: adjoint vs fp32 device central differences %.2e' % err)
    assert err < 5e-3


def test_adjoint_repeatable_and_state_restored():
    csr, m, c, sc, G, eta, x0 = _setup()
    sim = _device_sim(csr, m, c, sc, G, eta)
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.set_noise(SEED)
    sim.checkpoint_run(T, WIN)
    x_end = sim.get_state()
    smp, first = sim.samples()
    dS = np.random.default_rng(0).standard_normal(smp.shape).astype(np.float32)
    g1 = sim.adjoint(dS, first)
    np.testing.assert_array_equal(sim.get_state(), x_end)
    g2 = sim.adjoint(dS, first)
    for k in g1:
        np.testing.assert_array_equal(g1[k], g2[k])
    assert sim.step == T


def test_adjoint_consuming_checkpoints_gives_the_same_gradients():
    """With checkpoint reuse off the adjoint replays each window inside its
    checkpoint's buffers: the gradients and the end-of-run state equal those
    of the copying adjoint, and a second adjoint refuses to run on the
    consumed checkpoints until checkpoint_run runs again."""
    from tvb_hpc.hip import TvbxError
    csr, m, c, sc, G, eta, x0 = _setup()
    out = []
    for keep in (True, False):
        sim = _device_sim(csr, m, c, sc, G, eta)
        sim.set_checkpoint_reuse(keep)
        sim.set_state(x0)
        sim.init_history('zeros')
        sim.set_noise(SEED)
        sim.checkpoint_run(T, WIN)
        x_end = sim.get_state()
        smp, first = sim.samples()
        dS = np.random.default_rng(0).standard_normal(smp.shape).astype(
            np.float32)
        g = sim.adjoint(dS, first)
        np.testing.assert_array_equal(sim.get_state(), x_end)
        assert sim.step == T
        if not keep:
            with pytest.raises(TvbxError, match='consumed'):
                sim.adjoint(dS, first)
            # a new checkpointed run makes it usable again
            sim.set_state(x0)
            sim.step = 0
            sim.init_history('zeros')
            sim.checkpoint_run(T, WIN)
            g2 = sim.adjoint(dS, first)
            for k in g:
                np.testing.assert_array_equal(g[k], g2[k])
        # the simulation continues from the end of the run
        sim.run(WIN)
        out.append((g, sim.get_state()))
    for k in out[0][0]:
        np.testing.assert_array_equal(out[0][0][k], out[1][0][k])
    np.testing.assert_array_equal(out[0][1], out[1][1])


def test_second_backward_of_one_integrate_raises():
    """integrate's backward replays inside the checkpoints: a second
    backward of the same forward fails loudly instead of reading them."""
    from tvb_hpc.hip import TvbxError
    csr, m, c, sc, G, eta, x0 = _setup()
    s_t = _oracle_samples(csr, np.full(P, 0.5), eta, x0)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()
    sim = _device_sim(csr, m, c, sc, G, eta)
    Gt = torch.tensor(G, device='cuda', requires_grad=True)
    et = torch.tensor(eta, device='cuda', requires_grad=True)
    smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, WIN, seed=SEED)
    L = AG.fc_loss(smp, fct).sum()
    L.backward(retain_graph=True)
    assert torch.isfinite(Gt.grad).all()
    with pytest.raises(TvbxError, match='consumed'):
        L.backward()


def test_adjoint_with_temporal_blocking():
    """Forward replay in blocked launches and reverse launches of several
    steps give the same samples and gradients, bit for bit, as one-step
    launches; so do reverse sweeps that gather their transposed rows in
    groups of 4 or 32 (the same additions in the same order), with one
    node per wave (lambda kept in registers) or four."""
    csr, m, c, sc, G, eta, x0 = _setup()
    L = min(int(csr.delays.min()) + 1, 8)
    assert L > 2
    dS = None
    out = []
    for launch in (dict(waves=4, node_block=4),
                   dict(waves=4, node_block=4, steps_per_launch=L),
                   dict(waves=4, node_block=4, steps_per_launch=L,
                        adj_unroll=32),
                   dict(waves=1, node_block=1, steps_per_launch=L,
                        adj_unroll=4),
                   dict(waves=1, node_block=1, steps_per_launch=L,
                        adj_unroll=32)):
        sim = Simulation(m, c, sc, csr, P, params={'c:a': G, 'm:eta': eta},
                         kinds={'c:a': 'inst', 'm:eta': 'ni'}, tavg_period=K,
                         sample_slots=T // K, adjoint=True, **launch)
        sim.set_state(x0)
        sim.init_history('zeros')
        sim.set_noise(SEED)
        sim.checkpoint_run(T, WIN)
        smp, first = sim.samples()
        if dS is None:
            dS = np.random.default_rng(1).standard_normal(
                smp.shape).astype(np.float32)
        out.append((smp, sim.adjoint(dS, first)))
    for smp, g in out[1:]:
        np.testing.assert_array_equal(out[0][0], smp)
        for k in out[0][1]:
            np.testing.assert_array_equal(out[0][1][k], g[k])


def test_adjoint_unroll_32_refuses_wide_workgroups():
    """The adjoint with 32 gathers in flight is built for at most 8 waves
    per workgroup: a 16-wave launch is refused before anything runs."""
    from tvb_hpc.hip import TvbxError
    csr, m, c, sc, G, eta, x0 = _setup()
    sim = Simulation(m, c, sc, csr, P, params={'c:a': G, 'm:eta': eta},
                     kinds={'c:a': 'inst', 'm:eta': 'ni'}, tavg_period=K,
                     sample_slots=T // K, adjoint=True, waves=16,
                     adj_unroll=32)
    sim.set_state(x0)
    sim.init_history('zeros')
    with pytest.raises(TvbxError, match='waves per workgroup'):
        sim.checkpoint_run(T, WIN)
    assert sim.step == 0


def test_autograd_device_tensors_match_host_tensors():
    """With CUDA tensors the samples and dL/dsamples stay on the device
    (strided device copies, the adjoint reading a device pointer); the
    gradients equal those of the host-tensor path."""
    csr, m, c, sc, G, eta, x0 = _setup()
    s_t = _oracle_samples(csr, np.full(P, 0.5), eta, x0)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()
    grads = []
    for dev in ('cpu', 'cuda'):
        sim = _device_sim(csr, m, c, sc, G, eta)
        Gt = torch.tensor(G, device=dev, requires_grad=True)
        et = torch.tensor(eta, device=dev, requires_grad=True)
        smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, WIN,
                           seed=SEED)
        assert smp.device.type == dev
        AG.fc_loss(smp, fct).sum().backward()
        grads.append((Gt.grad.cpu().numpy(), et.grad.cpu().numpy(),
                      smp.detach().cpu().numpy()))
    np.testing.assert_array_equal(grads[0][2], grads[1][2])
    for a, b in zip(grads[0][:2], grads[1][:2]):
        np.testing.assert_allclose(b, a, rtol=1e-5,
                                   atol=1e-6 * np.abs(a).max())


def test_device_fc_loss_gradients_match_torch_loss():
    """The parameter gradients through the device FC loss
    (autograd.fc_loss_device) equal those through the float64 torch loss
    (autograd.fc_loss) on the same device samples."""
    csr, m, c, sc, G, eta, x0 = _setup()
    s_t = _oracle_samples(csr, np.full(P, 0.5), eta, x0)
    fct = AG.fc(torch.from_numpy(s_t))[0].numpy()
    out = []
    for loss_fn in (AG.fc_loss, AG.fc_loss_device):
        sim = _device_sim(csr, m, c, sc, G, eta)
        Gt = torch.tensor(G, device='cuda', requires_grad=True)
        et = torch.tensor(eta, device='cuda', requires_grad=True)
        smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': et}, x0, T, WIN,
                           seed=SEED)
        L = loss_fn(smp, fct)
        L.sum().backward()
        out.append((L.detach().double().cpu().numpy(),
                    Gt.grad.cpu().numpy(), et.grad.cpu().numpy()))
    np.testing.assert_allclose(out[1][0], out[0][0], rtol=0, atol=1e-6)
    for a, b in zip(out[0][1:], out[1][1:]):
        np.testing.assert_allclose(b, a, rtol=0,
                                   atol=1e-4 * np.abs(a).max())


def test_adjoint_grouped_weights_matches_device_finite_differences():
    """Two delay groups of 64 instances with their own weights and delays
    (subjects on one sparsity pattern): the reverse sweep reads each group's
    transposed weights; dL/dG agrees with float32 device central
    differences."""
    from tvb_hpc.network import CSR
    csr1, m, c, sc, G1, eta1, x01 = _setup()
    rng = np.random.default_rng(9)
    w2 = np.stack([csr1.weight, csr1.weight * rng.uniform(
        0.3, 1.7, csr1.nnz)]).astype(np.float32)
    d2 = np.stack([csr1.delays[0], csr1.delays[0] + 2]).astype(np.int32)
    csr = CSR(N, csr1.row, csr1.col, w2, d2)
    P2 = 2 * P
    G = np.concatenate([G1, G1[::-1]]).astype(np.float32)
    eta = np.concatenate([eta1, eta1], axis=1)
    x0 = np.concatenate([x01, x01], axis=2)
    sim = Simulation(m, c, sc, csr, P2, params={'c:a': G, 'm:eta': eta},
                     kinds={'c:a': 'inst', 'm:eta': 'ni'}, tavg_period=K,
                     sample_slots=T // K, adjoint=True, waves=4,
                     node_block=4, group_size=P)
    ref = O.OracleSim('MontbrioPazoRoxin', 'Linear', N, csr.row, csr.col,
                      csr.weight, csr.delays, P2, csr.max_delay + 2, DT,
                      scheme='em', params={'eta': eta},
                      cparams={'a': np.full(P2, 0.5)}, sigma=0.01,
                      dtype=np.float64, tavg_period=K, group_size=P)
    ref.set_state(x0)
    ref.init_history('zeros')
    fct = AG.fc(torch.from_numpy(ref.run(T, seed=SEED)['samples']))[0]
    fct = fct.numpy()
    Gt = torch.tensor(G, requires_grad=True)
    smp = AG.integrate(sim, {'c:a': Gt, 'm:eta': torch.tensor(eta)}, x0, T,
                       WIN, seed=SEED)
    AG.fc_loss(smp, fct).sum().backward()
    h = 1e-3

    def dev_loss(Gv):
        s = AG.integrate(sim, {'c:a': torch.tensor(Gv, dtype=torch.float64),
                               'm:eta': torch.tensor(eta)}, x0, T, WIN,
                         seed=SEED)
        return AG.fc_loss(s, fct).detach().numpy()

    fd = (dev_loss(G.astype(np.float64) + h) -
          dev_loss(G.astype(np.float64) - h)) / (2 * h)
    g = Gt.grad.numpy()
    err = np.abs(g - fd).max() / np.abs(fd).max()
    print('grouped weights dL/dG: adjoint vs fp32 central differences %.2e'
          % err)
    assert err < 5e-3
    # the two groups' gradients differ (their weights and delays do)
    assert not np.allclose(g[:P], g[P:][::-1])


def test_adjoint_recorded_forward_matches_checkpoints():
    """window = 0 records every step during the forward run; the reverse
    sweep then runs without replay and gives the same samples and
    gradients, bit for bit, as checkpoints and replay."""
    csr, m, c, sc, G, eta, x0 = _setup()
    out = []
    dS = np.random.default_rng(3).standard_normal((T // K, N, P)).astype(
        np.float32)
    for window in (WIN, 0):
        sim = _device_sim(csr, m, c, sc, G, eta)
        sim.set_state(x0)
        sim.init_history('zeros')
        sim.set_noise(SEED)
        sim.checkpoint_run(T, window)
        smp, first = sim.samples()
        out.append((smp, sim.adjoint(dS, first), sim.get_state()))
    np.testing.assert_array_equal(out[0][0], out[1][0])
    np.testing.assert_array_equal(out[0][2], out[1][2])
    for k in out[0][1]:
        np.testing.assert_array_equal(out[0][1][k], out[1][1][k])

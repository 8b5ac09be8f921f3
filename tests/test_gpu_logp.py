"""Euler-Maruyama transition log density of a given trajectory
(scheme.py:42-48 euler_maruyama_logp, summed over the network's nodes and
state variables, with the trajectory's own delayed coupling) on the device,
against the float32 oracle, against the closed form -sum xi^2 / 2 of a
trajectory the recursion generated from normals xi, and across launch
shapes."""

import numpy as np
import pytest

import oracle as O
from harness import MODELS, CFUNS, small_csr
from tvb_hpc import scheme as SC
from tvb_hpc.hip import Simulation

pytestmark = pytest.mark.gpu

N, P, T, DT, SIG = 16, 64, 240, 0.1, 0.02


def _frames(csr, G, x0, xi):
    "float32 oracle trajectory x(0) .. x(T) driven by the normals xi"
    gen = O.OracleSim('Generic2dOscillator', 'Sigmoidal', N, csr.row, csr.col,
                      csr.weight, csr.delays, P, csr.max_delay + 2, DT,
                      scheme='em', cparams={'a': G}, sigma=SIG,
                      dtype=np.float32)
    gen.set_state(x0)
    gen.init_history('zeros')
    tr = gen.run(T, noise=xi, traj=True)['traj']
    return np.concatenate([x0[None], tr]).astype(np.float32)


def _device_logp(csr, G, frames, **launch):
    m = MODELS['Generic2dOscillator']()
    c = CFUNS['Sigmoidal'](m)
    sim = Simulation(m, c, SC.EulerMaryuyamaStep(DT, sigma=SIG), csr, P,
                     params={'c:a': G}, **launch)
    sim.set_state(frames[0])
    sim.init_history('zeros')
    lp = sim.logp(frames, window=64)
    return sim, lp


@pytest.mark.parametrize('launch', [dict(waves=4, node_block=4),
                                    'blocked'])
def test_logp_matches_oracle_and_closed_form(launch):
    csr = small_csr(N=N, dt=DT, speeds=(2.0, ))
    if launch == 'blocked':
        L = min(int(csr.delays.min()) + 1, 8)
        assert L >= 2
        launch = dict(waves=1, node_block=1, steps_per_launch=L)
    print('logp launch', launch, flush=True)
    rng = np.random.default_rng(3)
    G = rng.uniform(0.0, 0.8, P).astype(np.float32)
    x0 = MODELS['Generic2dOscillator']().initial_state(
        N, P, np.random.default_rng(4)).astype(np.float32)
    xi = rng.standard_normal((T, 2, N, P)).astype(np.float32)
    frames = _frames(csr, G, x0, xi)
    sim, lp = _device_logp(csr, G, frames, **launch)
    ev = O.OracleSim('Generic2dOscillator', 'Sigmoidal', N, csr.row, csr.col,
                     csr.weight, csr.delays, P, csr.max_delay + 2, DT,
                     scheme='em', cparams={'a': G}, sigma=SIG,
                     dtype=np.float32)
    ev.set_state(x0)
    ev.init_history('zeros')
    ref = ev.logp(frames)
    err = np.abs(lp - ref).max() / np.abs(ref).max()
    print('logp vs float32 oracle %.2e' % err)
    assert err < 2e-4
    closed = -0.5 * (xi.astype(np.float64) ** 2).sum(axis=(0, 1, 2))
    np.testing.assert_allclose(lp, closed, rtol=5e-3)
    np.testing.assert_array_equal(sim.get_state(), frames[-1])
    assert sim.step == T


def test_logp_same_across_launch_shapes_and_windows():
    csr = small_csr(N=N, dt=DT, speeds=(2.0, ))
    rng = np.random.default_rng(8)
    G = rng.uniform(0.0, 0.8, P).astype(np.float32)
    x0 = MODELS['Generic2dOscillator']().initial_state(
        N, P, np.random.default_rng(9)).astype(np.float32)
    xi = rng.standard_normal((T, 2, N, P)).astype(np.float32)
    frames = _frames(csr, G, x0, xi)
    L = min(int(csr.delays.min()) + 1, 8)
    _, a = _device_logp(csr, G, frames, waves=4, node_block=4)
    _, b = _device_logp(csr, G, frames, waves=2, node_block=2,
                        steps_per_launch=L)
    np.testing.assert_array_equal(a, b)


def test_logp_gradients_match_fp64_finite_differences():
    """d logp / d G (coupling, per instance) and d logp / d eta (model, per
    node and instance) of a forced MPR trajectory from the emitted drift
    Jacobians, against central differences of the float64 oracle's logp
    (which, x(n) being given, is a smooth function of the parameters)."""
    from tvb_hpc import coupling, model
    csr = small_csr(N=12, density=0.5, seed=11, speeds=(1.0, ), dt=DT,
                    lmax=40.0)
    n, P2, T2, sig = 12, 64, 150, 0.01
    m = model.MontbrioPazoRoxin()
    c = coupling.Linear(m)
    rng = np.random.default_rng(5)
    G = rng.uniform(0.2, 0.8, P2)
    eta = -4.6 + 0.2 * rng.standard_normal((n, P2))
    x0 = np.stack([rng.uniform(0.1, 1.0, (n, P2)),
                   rng.uniform(-1.5, -0.5, (n, P2))])
    xi = rng.standard_normal((T2, 2, n, P2))

    def oracle(Gv, etav, dtype):
        o = O.OracleSim('MontbrioPazoRoxin', 'Linear', n, csr.row, csr.col,
                        csr.weight, csr.delays, P2, csr.max_delay + 2, DT,
                        scheme='em', params={'eta': etav},
                        cparams={'a': Gv}, sigma=sig, dtype=dtype)
        o.set_state(x0)
        o.init_history('zeros')
        return o

    frames = np.concatenate([x0[None], oracle(G, eta, np.float32).run(
        T2, noise=xi, traj=True)['traj']]).astype(np.float32)
    sim = Simulation(m, c, SC.EulerMaryuyamaStep(DT, sigma=sig), csr, P2,
                     params={'c:a': G.astype(np.float32),
                             'm:eta': eta.astype(np.float32)},
                     kinds={'c:a': 'inst', 'm:eta': 'ni'}, adjoint=True,
                     waves=4, node_block=4)
    sim.set_state(frames[0])
    sim.init_history('zeros')
    lp, g = sim.logp(frames, window=50, grad=True)
    G32 = G.astype(np.float32).astype(np.float64)
    eta32 = eta.astype(np.float32).astype(np.float64)
    f64 = frames.astype(np.float64)
    ref = oracle(G32, eta32, np.float64).logp(f64)
    assert np.abs(lp - ref).max() < 1e-3 * np.abs(ref).max()
    h = 1e-6
    fd_G = (oracle(G32 + h, eta32, np.float64).logp(f64) -
            oracle(G32 - h, eta32, np.float64).logp(f64)) / (2 * h)
    errG = np.abs(g['c:a'] - fd_G).max() / np.abs(fd_G).max()
    print('d logp / dG vs fp64 central differences %.2e' % errG)
    assert errG < 2e-3
    for i in (0, 7):
        e_p, e_m = eta32.copy(), eta32.copy()
        e_p[i] += h
        e_m[i] -= h
        fd = (oracle(G32, e_p, np.float64).logp(f64) -
              oracle(G32, e_m, np.float64).logp(f64)) / (2 * h)
        err = np.abs(g['m:eta'][i] - fd).max() / np.abs(fd).max()
        print('d logp / d eta[%d] vs fp64 central differences %.2e' % (i, err))
        assert err < 2e-3


def test_logp_autograd_matches_direct_gradients():
    """autograd.logp: the backward of sum(w * logp) equals w times the
    direct per-instance gradients."""
    import torch
    from tvb_hpc import autograd as AG
    from tvb_hpc import coupling, model
    csr = small_csr(N=10, density=0.5, seed=11, speeds=(1.0, ), dt=DT,
                    lmax=30.0)
    n, P2, T2, sig = 10, 64, 80, 0.01
    m = model.MontbrioPazoRoxin()
    c = coupling.Linear(m)
    rng = np.random.default_rng(6)
    G = rng.uniform(0.2, 0.8, P2).astype(np.float32)
    eta = (-4.6 + 0.2 * rng.standard_normal((n, P2))).astype(np.float32)
    x0 = np.stack([rng.uniform(0.1, 1.0, (n, P2)),
                   rng.uniform(-1.5, -0.5, (n, P2))]).astype(np.float32)
    o = O.OracleSim('MontbrioPazoRoxin', 'Linear', n, csr.row, csr.col,
                    csr.weight, csr.delays, P2, csr.max_delay + 2, DT,
                    scheme='em', params={'eta': eta}, cparams={'a': G},
                    sigma=sig, dtype=np.float32)
    o.set_state(x0)
    o.init_history('zeros')
    frames = np.concatenate([x0[None], o.run(T2, seed=3, traj=True)['traj']])
    sim = Simulation(m, c, SC.EulerMaryuyamaStep(DT, sigma=sig), csr, P2,
                     params={'c:a': G, 'm:eta': eta},
                     kinds={'c:a': 'inst', 'm:eta': 'ni'}, adjoint=True,
                     waves=4, node_block=4)
    Gt = torch.tensor(G, dtype=torch.float64, requires_grad=True)
    et = torch.tensor(eta, dtype=torch.float64, requires_grad=True)
    w = torch.linspace(0.5, 1.5, P2, dtype=torch.float64)
    lp = AG.logp(sim, {'c:a': Gt, 'm:eta': et}, frames)
    (w * lp).sum().backward()
    sim.step = 0
    sim.set_state(frames[0])
    sim.init_history('zeros')
    lp2, g = sim.logp(frames, grad=True)
    np.testing.assert_array_equal(lp.detach().numpy(), lp2)
    np.testing.assert_allclose(Gt.grad.numpy(), w.numpy() * g['c:a'],
                               rtol=1e-12)
    np.testing.assert_allclose(et.grad.numpy(), w.numpy() * g['m:eta'],
                               rtol=1e-12)


def test_logp_device_frames_in_place_match_host_frames():
    """frames in device memory (a torch CUDA tensor, P a multiple of 64) are
    read in place; the result equals the host-frame path's"""
    import torch
    csr = small_csr(N=N, dt=DT, speeds=(2.0, ))
    rng = np.random.default_rng(12)
    G = rng.uniform(0.0, 0.8, P).astype(np.float32)
    x0 = MODELS['Generic2dOscillator']().initial_state(
        N, P, np.random.default_rng(13)).astype(np.float32)
    xi = rng.standard_normal((T, 2, N, P)).astype(np.float32)
    frames = _frames(csr, G, x0, xi)
    _, a = _device_logp(csr, G, frames, waves=4, node_block=4)
    sim, _ = _device_logp(csr, G, frames[:2], waves=4, node_block=4)
    sim.step = 0
    sim.set_state(frames[0])
    sim.init_history('zeros')
    b = sim.logp(torch.from_numpy(frames).cuda())
    np.testing.assert_array_equal(a, b)
    # the state can be copied into device memory as well
    out = torch.zeros((2, N, P), device='cuda')
    sim.state_into(out.data_ptr())
    np.testing.assert_array_equal(out.cpu().numpy(), frames[-1])

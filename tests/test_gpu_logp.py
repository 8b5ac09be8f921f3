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

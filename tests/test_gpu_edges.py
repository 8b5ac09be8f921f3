"""Edge cases of the device integrator against the CPU oracle: a network
with no edges, a star whose hub row has more than 128 edges beside empty
rows, one node with a delayed self-connection, a single instance, zero
steps, and the monitor's smallest FC."""

import numpy as np
import pytest

from harness import CFUNS, MODELS, oracle_for, scaled_err
from tvb_hpc import scheme as SC
from tvb_hpc.hip import Simulation, TvbxError
from tvb_hpc.network import CSR

pytestmark = pytest.mark.gpu

TOL = 1e-5


def _csr(N, edges):
    """CSR from (target, source, weight, delay) tuples (rows = targets)."""
    edges = sorted(edges)
    row = np.zeros(N + 1, np.int32)
    for t, _, _, _ in edges:
        row[t + 1] += 1
    row = np.cumsum(row).astype(np.int32)
    col = np.array([s for _, s, _, _ in edges], np.int32)
    w = np.array([w for _, _, w, _ in edges], np.float32)
    d = np.array([[d for _, _, _, d in edges]], np.int32).reshape(1, -1)
    return CSR(N, row, col, w, d)


def _pair(csr, P, launch, model='Generic2dOscillator', cfun='Sigmoidal',
          dt=0.1, a=0.5, steps=300, seed=3):
    m = MODELS[model]()
    c = CFUNS[cfun](m)
    sim = Simulation(m, c, SC.EulerMaryuyamaStep(dt), csr, P,
                     params={'c:a': np.full(P, a, np.float32),
                             'sigma': np.full(P, 0.01, np.float32)},
                     kinds={'sigma': 'inst'}, **launch)
    rng = np.random.default_rng(seed)
    x0 = m.initial_state(csr.nnode, P, rng)
    sim.set_state(x0)
    sim.init_history('zeros')
    ref = oracle_for(model, cfun, csr, P, sim.H, dt, 'em', cparams={'a': a},
                     sigma=0.01)
    ref.set_state(x0)
    ref.init_history('zeros')
    S = len(m.state_sym)
    noise = rng.standard_normal((steps, S, csr.nnode, P)).astype(np.float32)
    return sim, ref, noise


BLOCKED = dict(waves=4, node_block=4, steps_per_launch=4)


@pytest.mark.parametrize('launch', [dict(waves=4), BLOCKED])
def test_network_without_edges(launch):
    """No edges at all: every node integrates alone (the coupling sum is
    empty), as on the oracle."""
    csr = _csr(7, [])
    assert csr.nnz == 0
    sim, ref, noise = _pair(csr, 70, launch)
    sim.run(300, noise=noise)
    ref.run(300, noise=noise)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL
    assert not sim.flags().any()


@pytest.mark.parametrize('launch', [dict(waves=4), BLOCKED,
                                    dict(ipl=2, waves=1, node_block=1,
                                         steps_per_launch=4)])
def test_star_with_long_hub_row_and_empty_rows(launch):
    """Node 0 receives from all 149 others (a row past 128 edges), nodes
    1 .. 9 from node 0, the rest from nobody (empty rows)."""
    N = 150
    rng = np.random.default_rng(1)
    edges = [(0, j, float(rng.uniform(0.1, 1.0)), int(rng.integers(4, 30)))
             for j in range(1, N)]
    edges += [(i, 0, 0.5, 5 + i) for i in range(1, 10)]
    csr = _csr(N, edges)
    assert csr.row[1] - csr.row[0] == 149 and csr.row[-1] - csr.row[10] == 0
    sim, ref, noise = _pair(csr, 128, launch)
    sim.run(300, noise=noise)
    ref.run(300, noise=noise)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL


@pytest.mark.parametrize('launch', [dict(waves=1), dict(
    waves=1, node_block=1, steps_per_launch=3)])
def test_single_node_with_delayed_self_connection(launch):
    csr = _csr(1, [(0, 0, 0.7, 3)])
    sim, ref, noise = _pair(csr, 64, launch)
    sim.run(300, noise=noise)
    ref.run(300, noise=noise)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL


def test_single_instance_and_zero_steps():
    """P = 1 (63 padding lanes of one wave) matches the oracle; run(0)
    changes nothing."""
    rng = np.random.default_rng(4)
    N = 12
    edges = [(i, j, float(rng.uniform(0.1, 1.0)), int(rng.integers(1, 20)))
             for i in range(N) for j in range(N) if i != j and
             rng.uniform() < 0.4]
    csr = _csr(N, edges)
    sim, ref, noise = _pair(csr, 1, dict(waves=4))
    x = sim.get_state()
    sim.run(0)
    np.testing.assert_array_equal(sim.get_state(), x)
    assert sim.step == 0
    sim.run(300, noise=noise)
    ref.run(300, noise=noise)
    assert sim.get_state().shape == (2, N, 1)
    assert scaled_err(sim.get_state(), ref.get_state()) < TOL
    assert not sim.flags().any()


def test_fc_needs_two_samples():
    """One monitor sample gives no correlation (an error, not NaNs); two
    give the correlation of those two samples (every entry +-1)."""
    rng = np.random.default_rng(5)
    N = 6
    edges = [(i, (i + 1) % N, 0.5, 2) for i in range(N)]
    csr = _csr(N, edges)
    m = MODELS['Generic2dOscillator']()
    sim = Simulation(m, CFUNS['Sigmoidal'](m), SC.EulerMaryuyamaStep(0.1),
                     csr, 64, params={'c:a': 0.5, 'sigma': 0.01},
                     tavg_period=10, fc=True, sample_slots=4)
    sim.set_state(m.initial_state(N, 64, rng))
    sim.init_history('zeros')
    sim.set_noise(seed=1)
    sim.run(10)
    with pytest.raises(TvbxError, match='two samples'):
        sim.fc(0, 1)
    sim.run(10)
    fc = sim.fc(0, 64)
    assert sim.n_fc_samples == 2
    assert fc.shape == (64, N, N)
    np.testing.assert_allclose(np.abs(fc), 1.0, atol=1e-6)

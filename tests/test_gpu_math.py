"""The emitted sine, cosine and division (tvbx_math.h: a bounded-argument
reduction and polynomials, the library's functions beyond |x| = 100; a
reciprocal corrected once by its residual) against float64 numpy, through
models whose drift is (0, sin x, cos x) or (0, 0, x / y): one deterministic
Euler step of dt = 1 from zeros leaves the values exactly as the device
computed them."""

import numpy as np
import pytest

from harness import small_csr
from tvb_hpc import coupling, model, scheme
from tvb_hpc.hip import Simulation

pytestmark = pytest.mark.gpu


class _SinCos(model.BaseModel):
    "x constant; s' = sin x, c' = cos x."
    state = 'x s c'
    input = 'i'
    param = 'a'
    drift = '0', 'sin(x)', 'cos(x)'
    diffs = 0, 0, 0
    obsrv = 'x',
    const = {'a': 0.0}
    limit = (-1, 1), (-1, 1), (-1, 1)


def _ulps(got, ref):
    ref32 = np.abs(ref).astype(np.float32)
    ulp = np.spacing(np.maximum(ref32, np.float32(2.0 ** -126)))
    return np.abs(got.astype(np.float64) - ref) / ulp.astype(np.float64)


def test_emitted_sin_cos_within_two_ulp():
    N, P = 16, 4096
    csr = small_csr(N=N, density=0.3, seed=3, speeds=(2.0, ), dt=0.1)
    m = _SinCos()
    c = coupling.Linear(m)
    c.param['a'] = 0.0
    rng = np.random.default_rng(9)
    x = np.concatenate([
        rng.uniform(-7.0, 7.0, N * P // 2),      # wrapped phases
        rng.uniform(-100.0, 100.0, N * P // 4),  # the reduced range
        rng.uniform(-1e4, 1e4, N * P // 8),      # the library's range
    ]).astype(np.float32)
    special = np.array([0.0, -0.0, np.pi / 2, np.pi, 2 * np.pi, -np.pi / 4,
                        99.99999, 100.0, 100.00001, -100.0, 1e-30, 1e30,
                        3e38], np.float32)
    x = np.concatenate([special, x])
    x = np.resize(x, N * P).reshape(N, P)
    sim = Simulation(m, c, scheme.EulerStep(1.0), csr, P, waves=4,
                     node_block=4)
    x0 = np.stack([x, np.zeros_like(x), np.zeros_like(x)]).astype(np.float32)
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.run(1)
    out = sim.get_state()
    assert np.array_equal(out[0], x)  # x' = 0
    ref_s = np.sin(x.astype(np.float64))
    ref_c = np.cos(x.astype(np.float64))
    es, ec = _ulps(out[1], ref_s), _ulps(out[2], ref_c)
    small = np.abs(x) <= 100.0
    print('|x| <= 100: sin %.2f, cos %.2f ulp max; beyond: sin %.2f, cos %.2f'
          % (es[small].max(), ec[small].max(), es[~small].max(),
             ec[~small].max()))
    assert es[small].max() <= 2.0 and ec[small].max() <= 2.0
    assert es[~small].max() <= 2.0 and ec[~small].max() <= 2.0
    # signs of zero and exact values at zero
    assert out[1][0, 0] == 0.0 and out[2][0, 0] == 1.0


class _Div(model.BaseModel):
    "x, y constant; q' = x / y."
    state = 'x y q'
    input = 'i'
    param = 'a'
    drift = '0', '0', 'x / y'
    diffs = 0, 0, 0
    obsrv = 'x',
    const = {'a': 0.0}
    limit = (-1, 1), (-1, 1), (-1, 1)


def test_emitted_division_within_one_ulp():
    N, P = 16, 4096
    csr = small_csr(N=N, density=0.3, seed=3, speeds=(2.0, ), dt=0.1)
    m = _Div()
    c = coupling.Linear(m)
    c.param['a'] = 0.0
    rng = np.random.default_rng(11)
    n = N * P
    # magnitudes over many binades, both signs, and quotients near 1
    x = (rng.choice([-1.0, 1.0], n) *
         np.exp2(rng.uniform(-60, 60, n))).astype(np.float32)
    y = (rng.choice([-1.0, 1.0], n) *
         np.exp2(rng.uniform(-60, 60, n))).astype(np.float32)
    y[: n // 8] = (x[: n // 8] * rng.uniform(0.99, 1.01, n // 8)).astype(
        np.float32)
    # special operands: infinite and zero divisors and dividends
    sx = np.array([1.0, -3.0, 0.0, 5.0, np.inf, 0.0, 2.0], np.float32)
    sy = np.array([np.inf, -np.inf, 7.0, 0.0, 3.0, 0.0, 1e-30], np.float32)
    x[:sx.size], y[:sy.size] = sx, sy
    x, y = x.reshape(N, P), y.reshape(N, P)
    sim = Simulation(m, c, scheme.EulerStep(1.0), csr, P, waves=4,
                     node_block=4)
    sim.set_state(np.stack([x, y, np.zeros_like(x)]).astype(np.float32))
    sim.init_history('zeros')
    sim.run(1)
    out = sim.get_state()
    with np.errstate(divide='ignore', invalid='ignore'):
        ref = x.astype(np.float64) / y.astype(np.float64)
    q = out[2].ravel()
    # x / +-inf = +-0, 0 / 7 = 0, 5 / 0 = inf, inf / 3 = inf, 0 / 0 = NaN
    assert q[0] == 0.0 and q[1] == 0.0 and q[2] == 0.0
    assert q[3] == np.inf and q[4] == np.inf and np.isnan(q[5])
    ok = np.isfinite(ref)
    ok.ravel()[:sx.size] = False
    e = _ulps(out[2][ok], ref[ok])
    print('division: max %.3f ulp, correctly rounded %.4f' % (
        e.max(), np.mean(out[2][ok] == ref[ok].astype(np.float32))))
    assert e.max() <= 1.0


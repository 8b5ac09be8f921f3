"""The emitted sine and cosine (tvbx_math.h: a bounded-argument reduction
and polynomials, the library's functions beyond |x| = 100) against float64
numpy, through a model whose drift is (0, sin x, cos x): one deterministic
Euler step of dt = 1 from s = c = 0 leaves s = sin(x0) and c = cos(x0)
exactly as the device computed them."""

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

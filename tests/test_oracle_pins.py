"""Pins of the CPU oracle (CPU only).

The oracle restates the reference's algorithm; these tests tie it to things
that do not depend on that restatement: rocRAND's Philox, the published
Random123 known-answer vectors, closed forms, numpy / scipy library routines,
the reference's own expression strings evaluated independently, and the
reference's end-to-end known-answer check (examples/hackathon.py)."""

import itertools
import os

import numpy as np
import pytest
from scipy import sparse
from scipy.integrate import solve_ivp
from scipy.stats import kstest

import oracle as O
from harness import MODELS, small_csr
from kat_hackathon import initial_theta, trend_stats, workload
from tvb_hpc import expr as E
from tvb_hpc.bold import BalloonWindkessel

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


# ------------------------------------------------------------------ RNG

def _golden_rows():
    rows = []
    with open(os.path.join(GOLDEN, 'philox4x32_10_rocrand.txt')) as fd:
        for line in fd:
            rows.append([int(v, 16) for v in line.split()])
    return rows


def test_philox_matches_rocrand_vectors():
    rows = _golden_rows()
    assert len(rows) == 11
    for r in rows:
        assert O.philox4x32_10(r[:4], r[4:6]).tolist() == r[6:], r


def test_philox_random123_known_answers():
    "Random123's published kat_vectors for philox4x32-10"
    kat = [
        ([0] * 4, [0] * 2, [0x6627e8d5, 0xe169c58d, 0xbc57ac4c, 0x9b00dbd8]),
        ([0xffffffff] * 4, [0xffffffff] * 2,
         [0x408f276d, 0x41c83b0e, 0xa20bc7c6, 0x6d5451fd]),
        ([0x243f6a88, 0x85a308d3, 0x13198a2e, 0x03707344],
         [0xa4093822, 0x299f31d0],
         [0xd16cfe09, 0x94fdcceb, 0x5001e420, 0x24126ea1]),
    ]
    for ctr, key, out in kat:
        assert O.philox4x32_10(ctr, key).tolist() == out


def test_boxmuller_normality():
    "the reference's (skipped) criteria, tests.py:329-342"
    z = O.normals_f32(1 << 20, seed=42).astype(np.float64)
    assert abs(z.mean()) < 0.005
    assert abs(z.std() - 1) < 0.005
    d, _ = kstest(z, 'norm')
    assert d < 0.01


def _bm(a, b):
    u = ((a >> 8) + 1) / 16777216.0
    v = (b >> 8) / 16777216.0
    r = np.sqrt(-2 * np.log(u))
    return r * np.cos(2 * np.pi * v), r * np.sin(2 * np.pi * v)


def test_step_normal_counter_layout():
    """normal of (instance p, step n, node i, svar s) = Box-Muller of
    Philox((n, i | (s >> 2) << 24, p, n >> 32), seed)"""
    seed = 0x1234567890ab
    for (p, n, i, s) in [(0, 0, 0, 0), (5, 17, 3, 1), (63, 2 ** 33 + 5, 998, 5),
                         (1000, 123456, 77, 3)]:
        ctr = [n & 0xffffffff, i | ((s >> 2) << 24), p, n >> 32]
        o = O.philox4x32_10(ctr, [seed & 0xffffffff, seed >> 32]).astype(
            np.int64)
        q = s & 3
        z = _bm(o[0], o[1]) if q < 2 else _bm(o[2], o[3])
        want = z[q & 1]
        got = O.step_normal(seed, n, i, p, s)
        assert abs(got - want) < 1e-12


# ------------------------------------------------------------------ closed forms

def _sim(model, cfun, csr, P, dt, **kw):
    kw.setdefault('dtype', np.float64)
    return O.OracleSim(model, cfun, csr.nnode, csr.row, csr.col, csr.weight,
                       csr.delays, P, csr.max_delay + 2, dt, **kw)


def test_kuramoto_uncoupled_is_rotation():
    "G = 0, no noise: theta_n = theta_0 + n omega dt (mod 2 pi)"
    csr = small_csr(N=8)
    P = 5
    om = np.linspace(0.01, 0.2, 8)
    sim = _sim('Kuramoto', 'Kuramoto', csr, P, 0.1, params={'omega': om},
               cparams={'a': 0.0}, R=2)
    th0 = np.random.default_rng(0).uniform(0, 6.28, (1, 8, P))
    sim.set_state(th0)
    sim.init_history('zeros')
    sim.run(1000)
    want = th0 + 1000 * 0.1 * om[None, :, None]
    d = np.angle(np.exp(1j * (sim.get_state() - want)))
    assert np.abs(d).max() < 1e-10
    assert (sim.get_state() >= 0).all() and (sim.get_state() <= 2 * np.pi).all()


def test_linear_uncoupled_euler_closed_form():
    csr = small_csr(N=6)
    sim = _sim('Linear', 'Linear', csr, 3, 0.01, params={'lam': -2.0},
               cparams={'a': 0.0})
    x0 = np.random.default_rng(1).normal(size=(1, 6, 3))
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.run(300)
    np.testing.assert_allclose(sim.get_state(), x0 * (1 - 0.02) ** 300,
                               rtol=1e-12)


def _dense(csr):
    W = np.zeros((csr.nnode, csr.nnode))
    for i in range(csr.nnode):
        for e in range(csr.row[i], csr.row[i + 1]):
            W[i, csr.col[e]] += csr.weight[e]
    return W


def test_linear_network_zero_delay_matrix_power():
    """all d = 0: x(n+1) = (I + dt(lam I + a W)) x(n) + dt b, as a power of
    the augmented matrix (numpy.linalg.matrix_power)."""
    csr = small_csr(N=7, zero_delays=0)
    csr.delays[:] = 0
    N, dt, lam, a, b = 7, 0.05, -0.7, 0.3, 0.2
    W = _dense(csr)
    A = np.eye(N) + dt * (lam * np.eye(N) + a * W)
    M = np.eye(N + 1)
    M[:N, :N] = A
    M[:N, N] = dt * b
    x0 = np.random.default_rng(2).normal(size=N)
    want = (np.linalg.matrix_power(M, 200) @ np.r_[x0, 1.0])[:N]
    sim = O.OracleSim('Linear', 'Linear', N, csr.row, csr.col, csr.weight,
                      np.zeros(csr.nnz, int), 1, 2, dt,
                      params={'lam': lam}, cparams={'a': a, 'b': b})
    sim.set_state(x0[None, :, None])
    sim.init_history('zeros')
    sim.run(200)
    np.testing.assert_allclose(sim.get_state()[0, :, 0], want, rtol=1e-11)


def _companion(csr, delays, dt, lam, a, b, order):
    """Linear model + Linear coupling with delays as one linear map on
    z = (x(n), x(n-1), ..., x(n-D), 1); zero history before step 0.  With the
    reference order and no wrap, zero-delay terms read the never-written
    slot (0)."""
    N = csr.nnode
    D = int(delays.max())
    K = N * (D + 1) + 1
    M = np.zeros((K, K))
    for i in range(N):
        M[i, i] += 1 + dt * lam
        M[i, K - 1] += dt * b
        for e in range(csr.row[i], csr.row[i + 1]):
            d = int(delays[e])
            if d == 0 and order == 'reference':
                continue
            M[i, d * N + csr.col[e]] += dt * a * float(csr.weight[e])
    for lag in range(1, D + 1):
        for i in range(N):
            M[lag * N + i, (lag - 1) * N + i] = 1.0
    M[K - 1, K - 1] = 1.0
    return M


@pytest.mark.parametrize('order', ['cuda', 'reference'])
def test_linear_network_with_delays_companion_form(order):
    csr = small_csr(N=5, density=0.6, seed=3, speeds=(5.0, ), lmax=12,
                    zero_delays=2)
    d = csr.delays[0]
    assert (d == 0).any() and d.max() >= 2
    N, dt, lam, a, b = 5, 0.1, -0.5, 0.4, 0.1
    M = _companion(csr, d, dt, lam, a, b, order)
    x0 = np.random.default_rng(4).normal(size=N)
    z0 = np.zeros(M.shape[0])
    z0[:N] = x0
    z0[-1] = 1.0
    steps = 60
    want = (np.linalg.matrix_power(M, steps) @ z0)[:N]
    H = steps + int(d.max()) + 5  # no wrap: reference order reads zeros
    sim = O.OracleSim('Linear', 'Linear', N, csr.row, csr.col, csr.weight,
                      d, 1, H, dt, params={'lam': lam},
                      cparams={'a': a, 'b': b}, order=order)
    sim.set_state(x0[None, :, None])
    sim.init_history('zeros')
    sim.run(steps)
    np.testing.assert_allclose(sim.get_state()[0, :, 0], want, rtol=1e-10,
                               atol=1e-14)


@pytest.mark.parametrize('d', [0, 1, 3, 7])
def test_impulse_arrives_after_delay(d):
    """one edge 0 <- 1 with delay d, x_1(0) = 1, lam = 0: x_0 first leaves 0
    at step index d + 1 (reads y_1(n - d) at step n)."""
    row = np.array([0, 1, 1])
    col = np.array([1])
    sim = O.OracleSim('Linear', 'Linear', 2, row, col, [1.0], [d], 1, d + 3,
                      1.0, params={'lam': 0.0}, cparams={'a': 1.0, 'b': 0.0})
    sim.set_state(np.array([[[0.0], [1.0]]]))
    sim.init_history('zeros')
    tr = sim.run(d + 4, traj=True)['traj'][:, 0, 0, 0]  # x_0(k + 1)
    first = int(np.argmax(tr != 0)) + 1
    assert first == d + 1
    assert tr[d] == 1.0


def test_diff_coupling_conserves_sum():
    "symmetric W, Diff coupling, zero delays, lam = 0: sum_i x_i constant"
    csr = small_csr(N=12, density=0.5)
    csr.delays[:] = 0
    sim = O.OracleSim('Linear', 'Diff', 12, csr.row, csr.col, csr.weight,
                      csr.delays, 4, 2, 0.05, params={'lam': 0.0},
                      cparams={'a': 0.7, 'b': 0.0})
    x0 = np.random.default_rng(5).normal(size=(1, 12, 4))
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.run(500)
    x = sim.get_state()
    np.testing.assert_allclose(x.sum(axis=1), x0.sum(axis=1), rtol=1e-12,
                               atol=1e-12)
    assert np.abs(x - x0).max() > 1e-3  # it did move


def test_csr_convention_matches_scipy():
    """row = target, col = source (the reference's tests.py:211-244): one
    Euler step with dt = 1, lam = 0, a = 1 is x + W x with W the CSR
    matrix as scipy reads it."""
    rng = np.random.default_rng(6)
    mat = sparse.csr_matrix(np.ones((40, 40)) * (rng.random((40, 40)) < 0.1)
                            * rng.random((40, 40)))
    x0 = rng.random(40)
    sim = O.OracleSim('Linear', 'Linear', 40, mat.indptr, mat.indices,
                      mat.data, np.zeros(mat.nnz, int), 1, 2, 1.0,
                      params={'lam': 0.0}, cparams={'a': 1.0, 'b': 0.0})
    sim.set_state(x0[None, :, None])
    sim.init_history('zeros')
    sim.run(1)
    np.testing.assert_allclose(sim.get_state()[0, :, 0], x0 + mat @ x0,
                               rtol=1e-12)


# ------------------------------------------------------------------ models

ORACLE_MODELS = ['Kuramoto', 'G2DO', 'Generic2dOscillator', 'RWW',
                 'ReducedWongWang', 'JansenRit', 'MontbrioPazoRoxin',
                 'Linear']


@pytest.mark.parametrize('name', ORACLE_MODELS)
def test_oracle_dfun_matches_reference_strings(name):
    """The hand-written oracle models against numpy evaluation of the model
    classes' expression strings (copied verbatim from the reference's
    model.py): one Euler step from random states with a constant input."""
    m = MODELS[name]()
    S = len(m.state_sym)
    N, P, dt, b = 6, 4, 1e-3, 0.37
    row = np.arange(N + 1) * 0
    sim = O.OracleSim(name, 'Linear', N, row, np.zeros(0, int),
                      np.zeros(0), np.zeros(0, int), P, 2, dt,
                      cparams={'a': 0.0, 'b': b}, R=1)
    rng = np.random.default_rng(7)
    x0 = m.initial_state(N, P, rng, np.float64)
    if name in ('RWW', 'ReducedWongWang'):
        x0 = rng.uniform(0.05, 0.95, x0.shape)
    sim.set_state(x0)
    sim.init_history('zeros')
    sim.run(1)
    x1 = sim.get_state()
    env = {v.name: x0[k] for k, v in enumerate(m.state_sym)}
    env.update({k: v for k, v in m.const.items()})
    env[m.input_sym[0].name] = b
    for lhs, rhs in m.auxex_exprs():
        env[lhs] = E.evaluate(rhs, env)
    f = np.stack([np.broadcast_to(E.evaluate(e, env), x0[0].shape)
                  for e in m.drift_exprs()])
    want = x0 + dt * f
    if name == 'Kuramoto':
        want = np.where(want > 2 * np.pi, want - 2 * np.pi, want)
    np.testing.assert_allclose(x1, want, rtol=1e-12, atol=1e-14)
    # ring slot 1 holds the first observable of the new state
    env1 = {v.name: x1[k] for k, v in enumerate(m.state_sym)}
    env1.update(m.const)
    y = E.evaluate(m.obsrv_exprs()[0], env1)
    np.testing.assert_allclose(sim.get_ring()[1, 0], y, rtol=1e-12,
                               atol=1e-14)


COUPLING_CASES = [
    ('Linear', {'a': 0.7, 'b': 0.2}),
    ('Diff', {'a': -0.45, 'b': 0.15}),
    ('Sigmoidal', {'cmin': 0.0, 'cmax': 0.9, 'midpoint': 0.3, 'r': 2.5,
                   'a': 1.3}),
    ('Kuramoto', {'a': 1.7, 'b': 0.0}),
]


@pytest.mark.parametrize('cfun,cparams', COUPLING_CASES)
def test_oracle_coupling_matches_reference_strings(cfun, cparams):
    """The oracle's hand-written pre / post coupling functions against
    numpy evaluation of the coupling classes' expression strings (copied
    verbatim from the reference's coupling.py:103-121), summed over CSR
    rows with mean = sum / nnode (network.py:115-136): one Euler step of
    the Linear model with lam = 0 and dt = 1, so x1 - x0 is the coupling
    term itself, from a history filled with the initial observables and
    random delays."""
    from harness import CFUNS
    N, P = 9, 5
    rng = np.random.default_rng(17)
    W = rng.uniform(0.1, 1.0, (N, N)) * (rng.uniform(size=(N, N)) < 0.6)
    np.fill_diagonal(W, 0.0)
    A = sparse.csr_matrix(W)
    row, col, w = A.indptr, A.indices, A.data
    delays = rng.integers(1, 6, size=col.size)
    sim = O.OracleSim('Linear', cfun, N, row, col, w, delays, P, 8, 1.0,
                      params={'lam': 0.0}, cparams=cparams)
    x0 = rng.uniform(-1.5, 1.5, (1, N, P))
    sim.set_state(x0)
    sim.init_history('constant')
    sim.run(1)
    got = sim.get_state()[0] - x0[0]
    m = MODELS['Linear']()
    c = CFUNS[cfun](m)
    pre = E.as_expr(c.pre_sum_sym[0])
    post = E.as_expr(c.post_sum_sym[0])
    want = np.zeros((N, P))
    for i in range(N):
        acc = np.zeros(P)
        for e in range(row[i], row[i + 1]):
            env = dict(cparams, pre_syn=x0[0, col[e]], post_syn=x0[0, i])
            acc = acc + w[e] * np.broadcast_to(E.evaluate(pre, env), (P, ))
        env = dict(cparams, sum=acc, mean=acc / N)
        want[i] = np.broadcast_to(E.evaluate(post, env), (P, ))
    np.testing.assert_allclose(got, want, rtol=1e-12, atol=1e-13)
    # a wrong statistic (mean by the row's degree, or no division) fails
    if cfun == 'Kuramoto':
        deg = np.diff(row)[:, None]
        assert not np.allclose(got, want * N / np.maximum(deg, 1))


# ------------------------------------------------------------------ BOLD / FC

def test_bold_rest_is_fixed_point():
    out, b = O.bold_f64(np.zeros(100), 1e-3)
    assert np.abs(out).max() == 0.0
    np.testing.assert_array_equal(b, [0.0, 1.0, 1.0, 1.0])


def test_bold_euler_converges_to_solve_ivp():
    """Balloon-Windkessel driven by a smooth input: the oracle's Euler steps
    against scipy's adaptive RK on bold.py's own expression strings."""
    bw = BalloonWindkessel()
    drift = bw.drift_exprs()
    names = [v.name for v in bw.state_sym]

    def xin(t):
        return 0.5 * np.sin(2 * np.pi * t / 3.0) ** 2

    def rhs(t, z):
        env = dict(zip(names, z))
        env.update(bw.const)
        env['x'] = xin(t)
        return [float(E.evaluate(e, env)) for e in drift]

    T, dt = 6.0, 1e-3
    ts = np.arange(int(T / dt)) * dt
    out, _ = O.bold_f64(xin(ts), dt)
    sol = solve_ivp(rhs, (0, T + dt), [0, 1, 1, 1], t_eval=ts + dt,
                    rtol=1e-10, atol=1e-12)
    env = dict(zip(names, sol.y))
    env.update(bw.const)
    ref = E.evaluate(bw.obsrv_exprs()[0], env)
    amp = np.abs(ref).max()
    assert amp > 0.05
    assert np.abs(out - ref).max() < 0.01 * amp


def test_online_fc_matches_corrcoef():
    rng = np.random.default_rng(8)
    x = rng.normal(size=(500, 9, 3))
    x[:, 1] += 0.8 * x[:, 0]
    x[:, 5] -= 0.5 * x[:, 2]
    fc = O.fc_f64(x)
    for p in range(3):
        np.testing.assert_allclose(fc[p], np.corrcoef(x[:, :, p].T),
                                   atol=1e-12)


# ------------------------------------------------------------------ KAT

def _hackathon(order):
    conn, couplings, speeds = workload()
    N = conn.nnode
    items = list(itertools.product(speeds, couplings))
    P = len(items)
    delays = np.stack([(conn.lnz / s / 1.0).astype(np.uintc)
                       for s, _ in items]).astype(np.int32)
    H = int(delays.max()) + 3 + 4000  # hackathon.py:46
    a = np.array([c for _, c in items])
    sim = O.OracleSim('Kuramoto', 'Kuramoto', N, conn.row, conn.col,
                      conn.wnz, delays, P, H, 1.0,
                      params={'omega': 2 * np.pi * 10 / 1e3},
                      cparams={'a': a}, group_size=1, R=2, order=order,
                      tavg_period=10)
    sim.set_state(initial_theta(N, P)[None])
    sim.init_history('zeros')
    out = sim.run(4000)
    trace = (out['samples'] * 10).transpose(2, 0, 1)
    return trend_stats(trace, conn.weights, speeds, couplings)


def test_hackathon_known_answer_reference_order():
    """examples/hackathon.py:88-89 holds under the loopy kernel's order
    (post_syn / zero delays read slot i_time before the model writes it)."""
    ds, dc = _hackathon('reference')
    assert ds > 350.0
    assert dc < -500.0


def test_hackathon_cuda_order_fails_speed_trend():
    """With the dsl_cuda kernels' order (post_syn = current phase) the same
    sweep misses the speed assertion -- evidence that the reference's loopy
    path ran in the read-before-write order."""
    ds, dc = _hackathon('cuda')
    assert ds < 350.0
    assert dc < -500.0


def test_u8_history_quantiser():
    """8-bit fixed point of [0, 1] values: integers 0 .. 255 within half a
    step of 255 x, monotone, exact on the k / 255 grid, clamped outside."""
    x = np.linspace(-0.2, 1.2, 20001).astype(np.float32)
    q = O.quantize(x, 'u8')
    assert np.all(q == np.round(q)) and q.min() == 0 and q.max() == 255
    inside = (x >= 0) & (x <= 1)
    assert np.all(np.abs(q[inside] - 255 * x[inside].astype(np.float64))
                  <= 0.5 + 1e-4)
    assert np.all(np.diff(q) >= 0)
    k = np.arange(256, dtype=np.float32)
    assert np.array_equal(O.quantize(k / np.float32(255), 'u8'), k)
    assert np.all(q[x < 0] == 0) and np.all(q[x > 1] == 255)


def test_u8_history_error_bound():
    """One step of a Jansen-Rit network whose history holds the 8-bit fixed
    point of the scaled sigmoid: every coupling sum is within
    sum_e |w_e| * 0.5 / 255 / ring_scale of the exact one, so the state
    after one Euler step moves by at most dt * |df/dc| * a times that."""
    from tvb_hpc import coupling, model, network
    W, L = network.synth_connectome(30, 0.3, 5, lmin=1.0, lmax=40.0)
    csr = network.CSR.from_dense(W, L, [2.0], 0.05)
    P = 16
    cf = coupling.Sigmoidal(model.JansenRit())
    cmax = cf.param['cmax']
    rng = np.random.default_rng(0)
    m = model.JansenRit()
    x0 = m.initial_state(30, P, rng)
    out = []
    for rd in ('f32', 'u8'):
        ref = O.OracleSim('JansenRit', 'Sigmoidal', 30, csr.row, csr.col,
                          csr.weight, csr.delays, P, csr.max_delay + 2, 0.05,
                          scheme='euler', cparams={'a': 300.0},
                          dtype=np.float32, ring_dtype=rd, ring_pre=True,
                          ring_scale=1.0 / cmax)
        ref.set_state(x0)
        ref.init_history('constant')
        ref.run(1)
        out.append(ref.get_state())
    rowsum = np.add.reduceat(np.abs(csr.weight), csr.row[:-1])
    dc = 300.0 * rowsum * 0.5 / 255 * cmax        # bound on |c_u8 - c|
    # JR couples into dy4/dt = A a (mu + ... + c): |df4/dc| = A a = 0.325;
    # plus the float32 spacing of y4 for the final rounding
    bound = 0.05 * 3.25 * 0.1 * dc[:, None] * (1 + 1e-3) + \
        2 * np.spacing(np.abs(out[0][4]))
    d = np.abs(out[1] - out[0])
    assert np.all(d[4] <= bound), (d[4].max(), bound.max())
    assert d[4].max() > 10 * np.spacing(np.abs(out[0][4])).max()
    assert np.all(d[[0, 1, 2, 3, 5]] <= 1e-9)
    assert d[4].max() > 0


# ------------------------------------------------------------ region maps

def test_rmap_oracle_reference_known_answers():
    """tvb_hpc/tests.py:453-473: 16 vertices in 4 regions of 4; to_avg
    gives the row sums of node reshaped (4, 4), from_avg of roi = [0..3]
    gives rmap itself."""
    i = np.r_[:16].reshape((-1, 1))
    rmap = i // 4
    node = i.astype('f')
    roi = O.rmap_to_avg(rmap, node, 4)
    np.testing.assert_allclose(roi[:, 0], node.reshape((4, 4)).sum(axis=1))
    back = O.rmap_from_avg(rmap, np.r_[:4].reshape((4, 1)).astype('f'))
    np.testing.assert_allclose(rmap, back)


def test_rmap_oracle_order_and_invariants():
    """to_avg adds in ascending vertex order: equal, bit for bit, to the
    reference kernel's loop written out; the total over regions equals the
    total over vertices for integer-valued data; an empty region is 0;
    from_avg(to_avg) of a region-constant field scales by region size."""
    rng = np.random.default_rng(3)
    nnode, nroi, nvar, P = 300, 7, 2, 5
    rmap = rng.integers(0, nroi - 1, size=nnode)      # region nroi-1 empty
    node = rng.standard_normal((nnode, nvar, P)).astype(np.float32)
    roi = O.rmap_to_avg(rmap, node, nroi)
    loop = np.zeros((nroi, nvar, P), np.float32)
    for j in range(nvar):
        for k in range(nnode):
            loop[rmap[k], j] = loop[rmap[k], j] + node[k, j]
    assert np.array_equal(roi, loop)
    assert not roi[nroi - 1].any()
    ints = rng.integers(-50, 50, size=(nnode, nvar, P)).astype(np.float32)
    np.testing.assert_array_equal(O.rmap_to_avg(rmap, ints, nroi).sum(0),
                                  ints.sum(0))
    const = rng.standard_normal((nroi, 1, 1)).astype(np.float32)
    field = O.rmap_from_avg(rmap, const)
    size = np.bincount(rmap, minlength=nroi).astype(np.float64)
    np.testing.assert_allclose(O.rmap_to_avg(rmap, field.astype(np.float64),
                                             nroi)[:, 0, 0],
                               size * const[:, 0, 0], rtol=1e-12)


# ------------------------------------------------------ grouped weights

def test_oracle_grouped_weights_equal_separate_runs():
    """An instance of delay group g with weights per group steps exactly as
    the same instance in a run whose only weights and delays are group g's
    (the grouped layout only selects rows)."""
    from tvb_hpc.network import CSR
    rng = np.random.default_rng(8)
    csr = small_csr(N=14, speeds=(1.0, 3.0))
    w2 = np.stack([csr.weight, rng.uniform(0.1, 1.0, csr.nnz)]).astype(
        np.float32)
    P, gs = 4, 2
    G = np.linspace(0.1, 0.9, P)
    x0 = MODELS['Generic2dOscillator']().initial_state(
        14, P, np.random.default_rng(2))
    kw = dict(dt=0.1, scheme='euler', cparams={'a': G})
    grp = O.OracleSim('Generic2dOscillator', 'Sigmoidal', 14, csr.row,
                      csr.col, w2, csr.delays, P, csr.max_delay + 2,
                      group_size=gs, **kw)
    grp.set_state(x0)
    grp.init_history('zeros')
    grp.run(300)
    for g in range(2):
        one = O.OracleSim('Generic2dOscillator', 'Sigmoidal', 14, csr.row,
                          csr.col, w2[g], csr.delays[g], P,
                          csr.max_delay + 2, **kw)
        one.set_state(x0)
        one.init_history('zeros')
        one.run(300)
        sl = slice(g * gs, (g + 1) * gs)
        np.testing.assert_array_equal(grp.get_state()[..., sl],
                                      one.get_state()[..., sl])
    # and the groups do differ
    assert not np.array_equal(grp.get_state()[..., :gs],
                              grp.get_state()[..., gs:])


def test_group_sparse_csr_builder():
    """CSR.group_sparse keeps the edges present in every subject (off the
    diagonal), in row-major order, with each subject's own weights and
    truncated delays."""
    from tvb_hpc.network import CSR
    rng = np.random.default_rng(5)
    S, N = 4, 9
    W = rng.uniform(0.1, 1.0, (S, N, N)) * (rng.random((S, N, N)) < 0.8)
    L = rng.uniform(1.0, 50.0, (S, N, N))
    csr = CSR.group_sparse(W, L, [1.0, 2.0, 4.0, 8.0], 0.1)
    mask = (W != 0).all(axis=0) & ~np.eye(N, dtype=bool)
    assert csr.nnz == mask.sum()
    assert csr.weight.shape == (S, csr.nnz) == csr.delays.shape
    for s, v in enumerate([1.0, 2.0, 4.0, 8.0]):
        dense = csr.dense_weights(group=s)
        np.testing.assert_allclose(dense, np.where(mask, W[s], 0.0),
                                   rtol=1e-6)
        rows = np.repeat(np.arange(N), np.diff(csr.row))
        np.testing.assert_array_equal(
            csr.delays[s], np.floor(L[s][rows, csr.col] / v / 0.1))


# ---------------------------------------------- EM transition log density

def test_oracle_logp_of_own_trajectory_is_minus_half_noise_energy():
    """For a trajectory the float64 Euler-Maruyama recursion generated with
    injected normals xi, each transition's residual is sqrt(dt) g xi, so
    the log density of scheme.py:42-48 summed over the run is
    -sum xi^2 / 2 per instance (up to float64 rounding).  A dropped dt, a
    wrong sign, or a frame off by one breaks it."""
    csr = small_csr(N=10, speeds=(2.0, ))
    P, T, dt, sig = 3, 120, 0.1, 0.02
    m = MODELS['Generic2dOscillator']()
    x0 = m.initial_state(10, P, np.random.default_rng(1)).astype(np.float64)
    xi = np.random.default_rng(2).standard_normal((T, 2, 10, P))
    kw = dict(scheme='em', cparams={'a': np.full(P, 0.4)}, sigma=sig,
              dtype=np.float64)
    gen = O.OracleSim('Generic2dOscillator', 'Sigmoidal', 10, csr.row,
                      csr.col, csr.weight, csr.delays, P, csr.max_delay + 2,
                      dt, **kw)
    gen.set_state(x0)
    gen.init_history('zeros')
    tr = gen.run(T, noise=xi, traj=True)['traj']
    frames = np.concatenate([x0[None], tr])
    ev = O.OracleSim('Generic2dOscillator', 'Sigmoidal', 10, csr.row,
                     csr.col, csr.weight, csr.delays, P, csr.max_delay + 2,
                     dt, **kw)
    ev.set_state(x0)              # init_history writes h(x0) at slot 0
    ev.init_history('zeros')
    lp = ev.logp(frames)
    np.testing.assert_allclose(lp, -0.5 * (xi ** 2).sum(axis=(0, 1, 2)),
                               rtol=1e-9)
    np.testing.assert_array_equal(ev.get_state(), frames[-1])
    # a frame off by one is far from it
    lp2 = O.OracleSim('Generic2dOscillator', 'Sigmoidal', 10, csr.row,
                      csr.col, csr.weight, csr.delays, P, csr.max_delay + 2,
                      dt, **kw)
    lp2.set_state(x0)
    lp2.init_history('zeros')
    bad = lp2.logp(np.concatenate([frames[:1], frames[:-1]]))
    assert (bad < 1.5 * lp).all()

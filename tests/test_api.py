"""The reference's Python API surface (CPU only): the unit tests of
tvb_hpc/tests.py restated for this package, plus the expression layer and
the emitter."""

import os
import re

import numpy as np
import pytest
from scipy import sparse

from tvb_hpc import bold, coupling, emit, model, network, scheme, transforms
from tvb_hpc import expr as E
from tvb_hpc.coupling import PostSumStat
from tvb_hpc.utils import VarSubst, exprs, parse, vars

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def test_var_subst():
    "tests.py:66-71"
    subst = VarSubst(b=parse('b[i, j]'))
    ex = subst(parse('a + b * pre_syn[i, j]'))
    assert str(ex) == 'a + b[i, j]*pre_syn[i, j]'


def test_vars_and_exprs():
    v = vars('x y z')
    assert [s.name for s in v] == ['x', 'y', 'z']
    e = exprs(('x + 1', 0.5))
    assert str(e[0]) == 'x + 1' and e[1] == 0.5
    with pytest.raises(Exception, match=re.escape("'x +'")):
        exprs(('x +', ))


@pytest.mark.parametrize('M,C,stat', [
    (model.G2DO, coupling.Linear, PostSumStat.sum),
    (model.G2DO, coupling.Diff, PostSumStat.sum),
    (model.JansenRit, coupling.Sigmoidal, PostSumStat.sum),
    (model.Kuramoto, coupling.Kuramoto, PostSumStat.mean),
])
def test_post_stat(M, C, stat):
    "tests.py:345-365"
    assert C(M()).post_stat(0) == stat


def test_io_truncates_obsrv_to_inputs():
    "coupling.py:68-84: HMJE has 2 inputs and 4 observables"
    cf = coupling.Linear(model.HMJE())
    assert len(cf.io) == 2
    assert [str(t[0]) for t in cf.io] == ['x1', 'x2']
    assert [str(t[3]) for t in cf.io] == ['c1', 'c2']
    assert len(coupling.Kuramoto(model.Kuramoto()).io) == 1


def test_prep_arrays():
    "model.py:67-84"
    m = model.JansenRit()
    st, inp, par, dr, df, ob = m.prep_arrays(10)
    assert st.shape == (10, 6) and inp.shape == (10, 1) and par.shape == (10, 2)
    assert dr.shape == (10, 6) and df.shape == (10, 6) and ob.shape == (10, 1)
    np.testing.assert_array_equal(par[:, 0], np.float32(m.const['v0']))
    for k, (lo, hi) in enumerate(m.limit):
        assert (st[:, k] >= lo).all() and (st[:, k] <= hi).all()


def test_models_copy_class_dicts():
    a, b = model.Kuramoto(), model.Kuramoto()
    a.const['omega'] = 3.0
    assert b.const['omega'] == 1.0


def _fd(e, env, name, h=1e-6):
    e1, e2 = dict(env), dict(env)
    e1[name] += h
    e2[name] -= h
    return (E.evaluate(e, e1) - E.evaluate(e, e2)) / (2 * h)


def test_logp_partials_match_finite_differences():
    "tests.py:281-294, with the partials checked numerically"
    m = model._TestModel()
    logp = scheme.euler_maruyama_logp(m.state_sym, m.drift_sym,
                                      m.diffs_sym).sum()
    parts = m.partial(logp)
    assert len(parts) == len(m.indvars)
    rng = np.random.default_rng(0)
    env = {n: float(rng.uniform(0.5, 1.5)) for n in
           ['y1', 'y2', 'i1', 'a', 'b', 'c', 'y1_n', 'y2_n', 'dt', 'd', 'e']}
    env['c'] = 0.3
    env['y1_3'] = env['y1'] ** 3
    for var, part in zip(m.indvars, parts):
        if var.name == 'y1':
            continue  # y1_3 is an auxiliary of y1 in the model, a free name here
        np.testing.assert_allclose(E.evaluate(part, env),
                                   _fd(logp, env, var.name), rtol=1e-6)


@pytest.mark.parametrize('text', [
    'sin(x) * exp(-y / 2)', 'x**3 - 2*x*y + y**2', '(x - y) / (1 + x*x)',
    'sqrt(1 + x*x) * log(2 + y)', 'x**y', 'tanh(x - y) + abs(x)',
    'u / (-expm1(-d * u) + (u == 0)) + (u == 0) / d',
])
def test_diff_matches_finite_differences(text):
    ex = parse(text)
    env = {'x': 0.7, 'y': 1.3, 'u': 0.2, 'd': 3.0}
    for n in ('x', 'y', 'u'):
        d = E.diff(ex, n)
        np.testing.assert_allclose(E.evaluate(d, env), _fd(ex, env, n),
                                   rtol=1e-6, atol=1e-9)


def test_c_literal_is_exact():
    for v in (0.1, 1 / 3, 2 * np.pi, -12.23904e-2, 1e-30, 8706.0):
        lit = E.c_literal(v)
        assert lit.rstrip(')').lstrip('(').endswith('f')
        body = lit.strip('()').rstrip('f')
        assert float.fromhex(body) == float(np.float32(v))


def test_to_c_operation_order():
    assert E.to_c(parse('a - b - c')) == '(((a) - (b)) - (c))'
    assert E.to_c(parse('V**3')) == '((V)*(V)*(V))'
    assert E.to_c(parse('x < 0')) == '((float)((x) < 0x0.0p+0f))'
    assert 'powf' in E.to_c(parse('v**3.125'))


BUILTIN = [(model.Kuramoto, coupling.Kuramoto), (model.G2DO, coupling.Sigmoidal),
           (model.Generic2dOscillator, coupling.Sigmoidal),
           (model.RWW, coupling.Linear), (model.ReducedWongWang, coupling.Linear),
           (model.JansenRit, coupling.Sigmoidal), (model.HMJE, coupling.Linear),
           (model.Linear, coupling.Diff), (model._TestModel, coupling.Linear),
           (model.MontbrioPazoRoxin, coupling.Linear)]


@pytest.mark.parametrize('M,C', BUILTIN)
@pytest.mark.parametrize('sch', ['euler', 'em'])
def test_emitter_writes_every_builtin_model(M, C, sch):
    "the codegen smoke tests of tests.py:297-326, 368-385, 388-405"
    m = M()
    s = scheme.EulerStep(0.1) if sch == 'euler' else \
        scheme.EulerMaryuyamaStep(0.1)
    spec = emit.KernelSpec(m, C(m), s)
    src = spec.source()
    assert 'tvbx_step' in src and 'tvbx_dfun' in src
    assert '#define TVBX_S %d' % len(m.state_sym) in src
    assert '#define TVBX_EM %d' % (sch == 'em') in src


def test_kernel_options_reach_the_source():
    """unroll / min_waves / pipe / prio become the template's defines and
    give distinct code-object keys; progress priorities are on by default."""
    m = model.JansenRit()
    c = coupling.Sigmoidal(m)
    sch = scheme.EulerStep(0.05)
    base = emit.KernelSpec(m, c, sch)
    src = base.source()
    assert '#define TVBX_PRIO 1' in src and 'TVBX_PIPE 1' not in src
    assert '#define TVBX_UNROLL 8' in src and '#define TVBX_MIN_WAVES 1' in src
    keys = {base.key()}
    for kw, text in ((dict(prio=False), '#define TVBX_PRIO 0'),
                     (dict(pipe=True), '#define TVBX_PIPE 1'),
                     (dict(unroll=16), '#define TVBX_UNROLL 16'),
                     (dict(min_waves=4), '#define TVBX_MIN_WAVES 4')):
        sp = emit.KernelSpec(m, c, sch, **kw)
        assert text in sp.source()
        keys.add(sp.key())
    assert len(keys) == 5
    with pytest.raises(ValueError):
        emit.KernelSpec(m, c, sch, unroll=12)
    with pytest.raises(ValueError):
        emit.KernelSpec(m, c, sch, min_waves=7)


def test_emitter_rejects_undefined_names():
    class Bad(model.Linear):
        drift = 'lam * x + c + undefined_thing',
    m = Bad()
    with pytest.raises(ValueError, match='undefined'):
        emit.KernelSpec(m, coupling.Linear(m), scheme.EulerStep(0.1)).source()


def test_param_kinds_and_layout():
    m = model.JansenRit()
    c = coupling.Sigmoidal(m)
    spec = emit.KernelSpec(m, c, scheme.EulerStep(0.05),
                           kinds={'c:a': 'inst', 'm:mu': 'inst',
                                  'm:A': 'glob', 'm:J': 'ni'})
    lay = {k: (kd, i) for k, kd, i in spec.param_layout()}
    assert lay['c:a'][0] == 'inst' and lay['m:mu'][0] == 'inst'
    assert lay['m:A'] == ('glob', 0) and lay['m:J'] == ('ni', 0)
    assert lay['m:v0'][0] == 'node' and lay['m:r'][0] == 'node'
    assert spec.counts() == {'glob': 1, 'inst': 2, 'node': 2, 'ni': 1}
    with pytest.raises(ValueError):
        emit.KernelSpec(m, c, scheme.EulerStep(0.05), kinds={'c:a': 'node'})
    with pytest.raises(ValueError):
        emit.KernelSpec(m, c, scheme.EulerStep(0.05), kinds={'zz': 'inst'})


def test_ring_pre_only_without_post_syn():
    m = model.Kuramoto()
    spec = emit.KernelSpec(m, coupling.Kuramoto(m), scheme.EulerStep(0.1))
    assert not spec.ring_pre and spec.uses_post
    with pytest.raises(ValueError):
        emit.KernelSpec(m, coupling.Kuramoto(m), scheme.EulerStep(0.1),
                        ring_pre=True)
    m = model.JansenRit()
    assert emit.KernelSpec(m, coupling.Sigmoidal(m),
                           scheme.EulerStep(0.1)).ring_pre


def test_adjoint_unroll_reaches_the_adjoint_source_only():
    """adj_unroll sets the adjoint's gathers per group (0: the forward's
    unroll); the forward-only source does not depend on it."""
    m = model.MontbrioPazoRoxin()
    c = coupling.Linear(m)
    sc = scheme.EulerStep(0.1)

    def spec(**kw):
        return emit.KernelSpec(m, c, sc, unroll=16, **kw)
    assert '#define TVBX_ADJ_UNROLL 16' in spec(adjoint=True).source()
    assert '#define TVBX_ADJ_UNROLL 32' in spec(adjoint=True,
                                                adj_unroll=32).source()
    assert 'TVBX_ADJ_UNROLL' not in spec().source()
    assert spec().key() == spec(adj_unroll=32).key()
    with pytest.raises(ValueError, match='adj_unroll'):
        spec(adjoint=True, adj_unroll=12)


def test_scheme_dt_symbolic_needs_value():
    s = scheme.EulerStep(E.var('dt'))
    with pytest.raises(ValueError):
        s.dt_value()
    assert s.dt_value(0.5) == 0.5


def test_network_time_step_returns_executor():
    from tvb_hpc.hip import HipTarget, NetworkStepExecutor
    osc = model.Kuramoto()
    knl = transforms.network_time_step(osc, coupling.Kuramoto(osc),
                                       scheme.EulerStep(1.0),
                                       target=HipTarget())
    assert isinstance(knl, NetworkStepExecutor)
    assert knl.order == 'reference'


def test_connectivity_hcp0():
    "network.py:66-93 on the shipped connectome"
    c = network.Connectivity.from_npz(os.path.join(GOLDEN, 'hcp0.npz'))
    assert c.nnode == 84 and c.nnz == 2299
    sw = sparse.csr_matrix(c.weights)
    np.testing.assert_array_equal(c.row, sw.indptr)
    np.testing.assert_array_equal(c.col, sw.indices)
    np.testing.assert_array_equal(c.wnz, sw.data)
    d = c.delays(3.0, 1.0)
    np.testing.assert_array_equal(d, (c.lnz / 3.0 / 1.0).astype(np.uintc))


def test_synth_connectome():
    W, L = network.synth_connectome(84, 0.4, seed=1)
    assert np.allclose(W, W.T) and np.allclose(L, L.T)
    assert np.diag(W).max() == 0 and W.max() == 1.0
    dens = (W > 0).sum() / (84 * 83)
    assert 0.35 < dens < 0.45
    nz = L[W > 0]
    assert nz.min() >= 10 and nz.max() <= 200
    W2, _ = network.synth_connectome(84, 0.4, seed=1)
    np.testing.assert_array_equal(W, W2)
    _, Lb = network.synth_connectome(200, 0.1, seed=3, lengths='ball')
    assert Lb.max() <= 300.0 + 1e-3


def test_csr_validation():
    row = np.array([0, 1, 2])
    with pytest.raises(ValueError, match='nnz'):
        network.CSR(2, row, [1], [1.0], [0])
    with pytest.raises(ValueError, match='col'):
        network.CSR(2, row, [1, 2], [1.0, 1.0], [0, 0])
    with pytest.raises(ValueError, match='negative'):
        network.CSR(2, row, [1, 0], [1.0, 1.0], [0, -1])
    c = network.CSR(2, row, [1, 0], [1.0, 2.0], [[3, 4], [5, 6]])
    assert c.n_group == 2 and c.max_delay == 6
    assert network.horizon_for(c.max_delay) == 8


def test_bold_constants():
    "bold.py:24-57"
    k = bold._balloon_windkessel_constants()
    assert k['k1'] == pytest.approx(4.3 * 40.3 * 0.4 * 0.04)
    assert k['k2'] == pytest.approx(0.5 * 25.0 * 0.4 * 0.04)
    assert k['k3'] == pytest.approx(0.5)
    assert bold.BalloonWindkessel().limit[1] == (1, 1)


def test_configs_build():
    from tvb_hpc import configs
    c0 = configs.kuramoto76()
    assert c0.csr.nnode == 76 and c0.n_inst == 64
    c1 = configs.g2do84()
    assert c1.n_inst == 8192 and c1.csr.n_group == 128
    assert c1.group_size == 64
    assert c1.horizon == c1.csr.max_delay + 2 and 1900 < c1.horizon <= 2002
    c3 = configs.jansenrit998()
    assert c3.csr.nnode == 998 and 95000 < c3.csr.nnz < 105000
    assert c3.n_inst == 1024
    sl = c1.sliced(64)
    assert sl.csr.n_group == 1 and sl.n_inst == 64


def test_mpr_workload_starts_in_a_stable_basin():
    """Every instance of the gradient workload (MPR N=84, Euler-Maruyama at
    dt 0.1 ms, coupling 0 .. 1) stays finite over its 20000 steps from the
    model's initial-condition box (float32 oracle on a sample)."""
    import oracle as O
    from tvb_hpc import configs
    cfg = configs.ALL['mpr84']()
    idx = np.arange(0, cfg.n_inst, 512)
    csr = cfg.csr
    ref = O.OracleSim('MontbrioPazoRoxin', 'Linear', csr.nnode, csr.row,
                      csr.col, csr.weight, csr.delays, len(idx),
                      csr.max_delay + 2, cfg.scheme.dt, scheme='em',
                      params={}, cparams={'a': np.asarray(
                          cfg.params['c:a'])[idx]},
                      sigma=0.01, dtype=np.float32, n_threads=4)
    ref.set_state(cfg.initial_state()[..., idx])
    ref.init_history('zeros')
    ref.run(20000, seed=5)
    assert np.isfinite(ref.get_state()).all()


def test_region_map_rejects_bad_indices_before_the_device():
    from tvb_hpc.hip import RegionMap
    with pytest.raises(ValueError):
        RegionMap(np.array([0, 1, 5]), nroi=3)
    with pytest.raises(ValueError):
        RegionMap(np.array([0, -1, 2]), nroi=3)

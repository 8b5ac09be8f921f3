"""
The sweep workloads the integrator is measured and checked on.

Five configurations (Kuramoto N=76; G2DO N=84 with BOLD + FC; reduced
Wong-Wang with uneven N; Jansen-Rit N=998 with delay bins; Montbrio-Pazo-Roxin
for FC-fit gradients), each built from the seeded synthetic-connectome
generator in ``tvb_hpc.network``.  Choices the workload descriptions leave
open are fixed here and listed in the repository's design notes:

* natural frequencies are in rad/ms: omega ~ N(2 pi 10 / 1e3, 1e-3);
* "150 mm sphere" is a ball of radius 150 mm;
* the couplings not named for RWW and MPR are ``Linear``;
* delays are round(L / (v dt)) with v in m/s (= mm/ms) and dt in ms;
* sweep instances are ordered with the speed axis outer and the coupling
  axis inner (dsl/dsl_cuda/example/cuda_setup.py:88-94), so 64 consecutive
  instances (one wavefront) share one conduction speed and one delay row.
"""

import math
from dataclasses import dataclass, field
from typing import Dict, Optional

import numpy as np

from . import coupling as CF
from . import model as M
from . import scheme as SC
from .network import CSR, horizon_for, lengths_to_delays, synth_connectome
from .network import Connectivity


@dataclass
class SweepConfig:
    name: str
    model: object
    cfun: object
    scheme: object
    csr: CSR
    n_inst: int
    n_steps: int
    params: Dict[str, object] = field(default_factory=dict)
    kinds: Dict[str, str] = field(default_factory=dict)
    group_size: Optional[int] = None
    tavg_period: int = 0
    bold: bool = False
    fc: bool = False
    bold_obs: int = 0
    seed: int = 0
    description: str = ''
    ring_dtype: str = 'f32'
    ring_scale: float = 1.0
    delay_bins: Optional[int] = None
    # monitor samples buffered on the device between FC merges (each merge
    # reads and writes the FP64 co-moment triangle once)
    sample_slots: int = 256
    # coupling sums: 'gather' (CSR rows) or 'mfma' (delay-binned, matrix
    # cores)
    coupling: str = 'gather'

    @property
    def horizon(self):
        return horizon_for(self.csr.max_delay)

    def sliced(self, n_inst):
        """The first n_inst instances (n_inst <= group_size when the config
        has several delay groups, so that one group is kept)."""
        P = self.n_inst
        params = {}
        for k, v in self.params.items():
            a = np.asarray(v)
            if a.ndim >= 1 and a.shape[-1] == P:
                params[k] = a[..., :n_inst]
            else:
                params[k] = v
        csr = self.csr
        gs = self.group_size
        if gs is not None and n_inst <= gs:
            csr = CSR(csr.nnode, csr.row, csr.col,
                      csr.weight[:1] if csr.weight_groups else csr.weight,
                      csr.delays[:1])
            gs = None
        return SweepConfig(self.name, self.model, self.cfun, self.scheme,
                           csr, n_inst, self.n_steps, params,
                           dict(self.kinds), gs, self.tavg_period, self.bold,
                           self.fc, self.bold_obs, self.seed,
                           self.description, self.ring_dtype, self.ring_scale,
                           self.delay_bins, self.sample_slots, self.coupling)

    def initial_state(self, dtype=np.float32):
        rng = np.random.default_rng(1000 + self.seed)
        return self.model.initial_state(self.csr.nnode, self.n_inst, rng,
                                        dtype)

    def simulation(self, **kw):
        from .hip import Simulation
        opts = dict(params=self.params, kinds=self.kinds,
                    group_size=self.group_size, tavg_period=self.tavg_period,
                    bold=self.bold, fc=self.fc, bold_obs=self.bold_obs,
                    ring_dtype=self.ring_dtype, ring_scale=self.ring_scale,
                    sample_slots=self.sample_slots, coupling=self.coupling)
        opts.update(kw)
        return Simulation(self.model, self.cfun, self.scheme, self.csr,
                          self.n_inst, **opts)

    def spec(self, **kw):
        "Kernel spec of the device-resident path (for ahead-of-time builds)."
        from .hip import kind_of_value
        from .emit import KernelSpec
        from .hip import _resolve
        probe = KernelSpec(self.model, self.cfun, self.scheme)
        kinds = dict(self.kinds)
        N, P = self.csr.nnode, self.n_inst
        for name, val in self.params.items():
            key = 'sigma' if name == 'sigma' else _resolve(probe, name)
            kinds.setdefault(key, kind_of_value(val, N, P))
        sk = kinds.pop('sigma', None)
        if sk is None and getattr(self.scheme, 'sigma', None) is not None:
            sk = 'const'
        kw.setdefault('ring_dtype', self.ring_dtype)
        return KernelSpec(self.model, self.cfun, self.scheme, kinds=kinds,
                          sigma_kind=sk, **kw)


@dataclass
class BatchConfig:
    """A sweep over subjects with different connectomes, integrated in one
    launch (``tvb_hpc.hip.Batch``): subjects = [(csr, n_inst), ...]."""
    name: str
    model: object
    cfun: object
    scheme: object
    subjects: list
    n_steps: int
    params: Dict[str, object] = field(default_factory=dict)
    kinds: Dict[str, str] = field(default_factory=dict)
    seed: int = 0
    description: str = ''

    @property
    def n_inst(self):
        return sum(p for _, p in self.subjects)

    def initial_states(self, dtype=np.float32):
        rng = np.random.default_rng(1000 + self.seed)
        return [self.model.initial_state(csr.nnode, p, rng, dtype)
                for csr, p in self.subjects]

    def batch(self, **kw):
        from .hip import Batch
        return Batch(self.model, self.cfun, self.scheme, self.subjects,
                     params=self.params, kinds=self.kinds, **kw)

    def spec(self, **kw):
        "Kernel spec of the batch (for ahead-of-time builds)."
        from .emit import KernelSpec
        kinds = dict(self.kinds)
        sk = kinds.pop('sigma', None)
        return KernelSpec(self.model, self.cfun, self.scheme, kinds=kinds,
                          sigma_kind=sk, **kw)


def rww_uneven(n_subjects=32, inst_per_subject=64, n_steps=20000,
               sizes=(68, 84, 148, 246, 400), seed=2):
    """Batched sweep, uneven sizes in one launch: reduced Wong-Wang on
    n_subjects connectomes with N drawn uniformly from `sizes`, each built
    like the Kuramoto N=76 one (30 % density, lognormal weights, lengths
    U(10, 200) mm) with its own seed and a conduction speed U(1, 10) m/s;
    each subject has inst_per_subject instances (one wavefront tile) with
    G, J_NMDA in [0.1, 0.3] and sigma in [1e-3, 1e-2] drawn uniformly."""
    rng = np.random.default_rng(seed)
    dt = 0.1
    subjects = []
    for k in range(n_subjects):
        N = int(rng.choice(sizes))
        v = float(rng.uniform(1.0, 10.0))
        W, L = synth_connectome(N, 0.30, seed=100 * seed + k)
        subjects.append((csr_from(W, L, [v], dt), inst_per_subject))
    P = n_subjects * inst_per_subject
    G = rng.uniform(0.0, 0.5, P)
    J = rng.uniform(0.1, 0.3, P)
    sigma = rng.uniform(1e-3, 1e-2, P)
    osc = M.ReducedWongWang()
    return BatchConfig(
        'rww_uneven', osc, CF.Linear(osc), SC.EulerMaryuyamaStep(dt),
        subjects, n_steps,
        params={'c:a': G, 'm:J': J, 'sigma': sigma},
        kinds={'c:a': 'inst', 'm:J': 'inst', 'sigma': 'inst'}, seed=seed,
        description='RWW, %d subjects (N in %s), P=%d, 20000 steps'
        % (n_subjects, '/'.join(str(s) for s in sizes), P))


def csr_from(weights, lengths, speeds, dt, policy='round'):
    return CSR.from_dense(weights, lengths, speeds, dt, policy)


def kuramoto76(n_inst=64, n_steps=10000):
    "Plain method, small size: Kuramoto N=76 (CPU-runnable case)."
    N, dt, v = 76, 0.1, 3.0
    W, L = synth_connectome(N, 0.30, seed=0)
    csr = csr_from(W, L, [v], dt)
    rng = np.random.default_rng(100)
    omega = rng.normal(2 * math.pi * 10 / 1e3, 1e-3, size=N)
    G = np.linspace(0.0, 2.0, n_inst)
    osc = M.Kuramoto()
    return SweepConfig(
        'kuramoto76', osc, CF.Kuramoto(osc),
        SC.EulerMaryuyamaStep(dt, sigma=0.01), csr, n_inst, n_steps,
        params={'omega': omega, 'c:a': G}, kinds={'m:omega': 'node'},
        seed=0, description='Kuramoto N=76, P=64, G in [0, 2], sigma 0.01')


def g2do84(n_coupling=64, n_speed=128, n_steps=10000, bold=True):
    "Batched sweep, uniform size: G2DO N=84, BOLD + FC."
    N, dt = 84, 0.1
    W, L = synth_connectome(N, 0.40, seed=1)
    speeds = np.linspace(1.0, 10.0, n_speed)
    csr = csr_from(W, L, speeds, dt)
    G = np.linspace(0.0, 1.0, n_coupling)
    P = n_coupling * n_speed
    osc = M.Generic2dOscillator()
    cf = CF.Sigmoidal(osc)
    # sigmoid centred on the oscillator's range so that G matters
    cf.param.update({'cmax': 1.0, 'midpoint': 0.0, 'r': 1.0})
    return SweepConfig(
        'g2do84', osc, cf,
        SC.EulerMaryuyamaStep(dt, sigma=0.005), csr, P, n_steps,
        params={'c:a': np.tile(G, n_speed)}, group_size=n_coupling,
        tavg_period=int(round(2.0 / dt)), bold=bold, fc=True, seed=1,
        description='G2DO N=84, P=%d (%d G x %d speeds), BOLD 2 ms, FC'
        % (P, n_coupling, n_speed))


def jansenrit998(n_inst=1024, n_steps=10000, tavg_period=20):
    "Mixed precision at high resolution: Jansen-Rit N=998, 10 % density."
    N, dt, v = 998, 0.05, 5.0
    W, L = synth_connectome(N, 0.10, seed=3, lengths='ball', radius=150.0)
    csr = csr_from(W, L, [v], dt)
    ng = int(round(math.sqrt(n_inst)))
    nm = n_inst // ng
    G = np.linspace(0.0, 1.0, ng)
    mu = np.linspace(0.1, 0.3, nm)
    Gp = np.tile(G, nm)[:n_inst]
    mup = np.repeat(mu, ng)[:n_inst]
    osc = M.JansenRit()
    return SweepConfig(
        'jansenrit998', osc, CF.Sigmoidal(osc), SC.EulerStep(dt), csr,
        n_inst, n_steps, params={'c:a': Gp, 'm:mu': mup},
        tavg_period=tavg_period, fc=True, seed=3,
        description='Jansen-Rit N=998, P=%d (G x mu), dt 0.05 ms' % n_inst)


def jansenrit998_mixed(n_inst=1024, n_steps=10000, tavg_period=20,
                       ring_dtype='bf16', bins=32, coupling='gather'):
    """Jansen-Rit N=998 with delays quantised to `bins` values and a
    reduced-precision history ring (bf16 / f16, or 8-bit fixed point) holding
    1/cmax-scaled sigmoid outputs in [0, 1]; state, sums and FC stay
    float32 / float64.  coupling 'mfma': the sum over bins of W_d X(t - d)
    on the matrix cores (bf16 / f16 rings)."""
    cfg = jansenrit998(n_inst, n_steps, tavg_period)
    csr, _ = cfg.csr.binned(bins)
    cfg.csr = csr.sorted_by_delay()
    cfg.name = 'jansenrit998_%s_b%d' % (ring_dtype, bins)
    cfg.ring_dtype = ring_dtype
    cfg.ring_scale = 1.0 / float(cfg.cfun.param['cmax'])
    cfg.delay_bins = bins
    cfg.description += ', %d delay bins, %s ring' % (bins, ring_dtype)
    if coupling != 'gather':
        cfg.name += '_' + coupling
        cfg.coupling = coupling
        cfg.description += ', %s coupling' % coupling.upper()
    return cfg


def mpr84(n_inst=4096, n_steps=20000, G=None, tavg_period=10):
    "Differentiation for inversion: MPR N=84."
    N, dt, v = 84, 0.1, 1.0
    W, L = synth_connectome(N, 0.30, seed=4)
    csr = csr_from(W, L, [v], dt)
    osc = M.MontbrioPazoRoxin()
    if G is None:
        G = np.linspace(0.0, 1.0, n_inst)
    return SweepConfig(
        'mpr84', osc, CF.Linear(osc), SC.EulerMaryuyamaStep(dt, sigma=0.01),
        csr, n_inst, n_steps, params={'c:a': G},
        kinds={'m:eta': 'ni'}, tavg_period=tavg_period, fc=True, seed=4,
        description='MPR N=84, P=%d, sigma 0.01' % n_inst)


def hcp100_sparse(n_subj=100, per_subj=64, n_steps=10000, speed=2.0):
    """Subjects on one sparsity pattern (the group-sparse layout of
    examples/hcp100.py:22-53): G2DO, linear coupling, Euler, dt 0.1 ms, N=84;
    each subject has its own weights and lengths over the edges present in
    every subject (15 % of node pairs, as the reference reports for its 100
    subjects), and is one delay group of `per_subj` instances sweeping the
    coupling."""
    N, dt = 84, 0.1
    W0, L0 = synth_connectome(N, 0.15, seed=6)
    rng = np.random.default_rng(60)
    W = np.stack([W0 * rng.uniform(0.7, 1.3, (N, N)) for _ in range(n_subj)])
    L = np.stack([np.where(L0 > 0, L0 * rng.uniform(0.9, 1.1, (N, N)), 0.0)
                  for _ in range(n_subj)])
    csr = CSR.group_sparse(W, L, speed, dt)
    G = np.linspace(0.0, 0.5, per_subj)
    P = n_subj * per_subj
    osc = M.Generic2dOscillator()
    return SweepConfig(
        'hcp100_sparse', osc, CF.Linear(osc), SC.EulerStep(dt), csr, P,
        n_steps, params={'c:a': np.tile(G, n_subj)}, group_size=per_subj,
        seed=6, description='G2DO N=84, %d subjects on one pattern (own '
        'weights and delays) x %d couplings' % (n_subj, per_subj))


def hcp0_kuramoto(path):
    """The reference's hackathon workload (examples/hackathon.py:23-89):
    Kuramoto on data/hcp0.npz with osc.dt = 1.0, omega = 2 pi 10 / 1e3."""
    conn = Connectivity.from_npz(path)
    return conn


ALL = {'kuramoto76': kuramoto76, 'g2do84': g2do84,
       'hcp100_sparse': hcp100_sparse,
       'jansenrit998': jansenrit998, 'mpr84': mpr84,
       'jansenrit998_bf16': lambda **kw: jansenrit998_mixed(
           ring_dtype='bf16', **kw),
       'jansenrit998_f16': lambda **kw: jansenrit998_mixed(
           ring_dtype='f16', **kw),
       'jansenrit998_u8': lambda **kw: jansenrit998_mixed(
           ring_dtype='u8', **kw),
       'jansenrit998_bf16_mfma': lambda **kw: jansenrit998_mixed(
           ring_dtype='bf16', coupling='mfma', **kw),
       'jansenrit998_f16_mfma': lambda **kw: jansenrit998_mixed(
           ring_dtype='f16', coupling='mfma', **kw)}

# sweeps over subjects with different connectomes (tvb_hpc.hip.Batch)
BATCHES = {'rww_uneven': rww_uneven}


def all_specs():
    """Kernel specs the builtin configurations use, for every launch width
    the autotuner may pick (for ahead-of-time builds)."""
    from .autotune import default_launch, table_launches, valid_ipls
    specs, keys = [], set()
    for fn in ALL.values():
        cfg = fn()
        launches = [dict(ipl=ipl) for ipl in valid_ipls(cfg)]
        launches += table_launches(cfg) + [default_launch(cfg)]
        for la in launches:
            sp = cfg.spec(ipl=la['ipl'], unroll=la.get('unroll', 8),
                          min_waves=la.get('min_waves', 1),
                          pipe=la.get('pipe', False),
                          prio=la.get('prio', True),
                          emaj=la.get('emaj', 0))
            if sp.key() not in keys:
                keys.add(sp.key())
                specs.append(sp)
    for fn in BATCHES.values():
        cfg = fn()
        launches = [dict(ipl=1)] + table_launches(cfg)
        for la in launches:
            sp = cfg.spec(ipl=la['ipl'], unroll=la.get('unroll', 8),
                          min_waves=la.get('min_waves', 1),
                          pipe=la.get('pipe', False),
                          prio=la.get('prio', True),
                          emaj=la.get('emaj', 0))
            if sp.key() not in keys:
                keys.add(sp.key())
                specs.append(sp)
    return specs

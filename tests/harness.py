"""Shared helpers for the parity tests: small seeded problems run side by
side on the device (through the C ABI) and on the CPU oracle."""

import numpy as np

import oracle as O
from tvb_hpc import coupling as CF
from tvb_hpc import model as M
from tvb_hpc import scheme as SC
from tvb_hpc.network import CSR, synth_connectome

MODELS = {'Kuramoto': M.Kuramoto, 'G2DO': M.G2DO,
          'Generic2dOscillator': M.Generic2dOscillator, 'RWW': M.RWW,
          'ReducedWongWang': M.ReducedWongWang,
          'JansenRit': M.JansenRit, 'MontbrioPazoRoxin': M.MontbrioPazoRoxin,
          'Linear': M.Linear}
CFUNS = {'Linear': CF.Linear, 'Sigmoidal': CF.Sigmoidal, 'Diff': CF.Diff,
         'Kuramoto': CF.Kuramoto}

# (model, coupling, dt, coupling params) exercised by the parity tests
CASES = [
    ('Kuramoto', 'Kuramoto', 0.1, {'a': 1.5}),
    ('Generic2dOscillator', 'Sigmoidal', 0.1, {'a': 0.5}),
    ('ReducedWongWang', 'Linear', 0.1, {'a': 0.05}),
    ('JansenRit', 'Sigmoidal', 0.05, {'a': 0.3}),
    ('MontbrioPazoRoxin', 'Linear', 0.01, {'a': 0.2}),
    ('Linear', 'Diff', 0.1, {'a': 0.3}),
]


def small_csr(N=24, density=0.3, seed=7, speeds=(2.0, ), dt=0.1, lmax=60.0,
              zero_delays=0):
    """Seeded small connectome; ``zero_delays`` edges get delay 0 (every
    group), so that the d = 0 read path is exercised."""
    W, L = synth_connectome(N, density, seed, lmin=1.0, lmax=lmax)
    csr = CSR.from_dense(W, L, list(speeds), dt)
    if zero_delays:
        rng = np.random.default_rng(seed + 1)
        e = rng.choice(csr.nnz, size=zero_delays, replace=False)
        d = csr.delays.copy()
        d[:, e] = 0
        csr = CSR(csr.nnode, csr.row, csr.col, csr.weight, d)
    return csr


def state_err(model, a, b):
    """Parity error: scaled_err, with Kuramoto phases compared modulo 2 pi
    (relative to the 2 pi range)."""
    if model == 'Kuramoto':
        return phase_err(a, b) / (2 * np.pi)
    return scaled_err(a, b)


def scaled_err(a, b):
    """max |a - b| / (max |b| per state variable): the error measure of the
    parity tests (relative to each variable's scale)."""
    scale = np.abs(b).max(axis=tuple(range(1, b.ndim)), keepdims=True)
    return float((np.abs(a - b) / (scale + 1e-30)).max())


def phase_err(a, b):
    "max wrapped phase difference"
    return float(np.abs(np.angle(np.exp(1j * (a.astype(np.float64)
                                              - b.astype(np.float64))))).max())


def oracle_for(model, cfun, csr, P, H, dt, scheme, cparams, params=None,
               sigma=None, dtype=np.float32, group_size=None, **kw):
    return O.OracleSim(model, cfun, csr.nnode, csr.row, csr.col, csr.weight,
                       csr.delays, P, H, dt, scheme=scheme,
                       params=params or {}, cparams=cparams, sigma=sigma,
                       dtype=dtype, group_size=group_size, **kw)

"""The reference's end-to-end acceptance check (examples/hackathon.py:23-89)
restated: an 8 x 8 (speed x coupling) Kuramoto sweep on the hcp0 connectome
(osc.dt = 1.0, omega = 2 pi 10 / 1e3, Euler, no noise, 400 windows of 10
steps), FC of 2 s windows of the 10-step sums of theta against the SC, and the
trend statistics derr_speed / derr_coupl."""

import os

import numpy as np

from tvb_hpc.network import Connectivity

GOLDEN = os.path.join(os.path.dirname(os.path.abspath(__file__)), 'golden')


def workload():
    conn = Connectivity.from_npz(os.path.join(GOLDEN, 'hcp0.npz'))
    couplings = np.logspace(0, 1.0, 8)
    speeds = np.logspace(0.0, 2.0, 8)
    return conn, couplings, speeds


def trend_stats(trace, weights, speeds, couplings):
    """hackathon.py:64-87: trace (n_work_items, 400, nnode) of 10-step sums;
    returns (derr_speed, derr_coupl)."""
    nnode = weights.shape[0]
    n_work_items = trace.shape[0]
    r, c = np.triu_indices(nnode, 1)
    win_size = 200
    tavg = np.transpose(trace, (1, 2, 0))
    win_tavg = tavg.reshape((-1, win_size) + tavg.shape[1:])
    err = np.zeros((len(win_tavg), n_work_items))
    for i, tavg_ in enumerate(win_tavg):
        for j in range(n_work_items):
            fc = np.corrcoef(tavg_[:, :, j].T)
            err[i, j] = ((fc[r, c] - weights[r, c]) ** 2).sum()
    err_ = err[-1].reshape((speeds.size, couplings.size))
    derr_speed = np.diff(err_.mean(axis=1)).sum()
    derr_coupl = np.diff(err_.mean(axis=0)).sum()
    return derr_speed, derr_coupl


def initial_theta(nnode, n, seed=0):
    "prep_arrays' uniform(0, 2 pi) initial phases (model.py:75-80), seeded"
    rng = np.random.default_rng(seed)
    return rng.uniform(0, 2 * np.pi, size=(nnode, n))

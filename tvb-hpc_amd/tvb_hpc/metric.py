"""
Functional-connectivity statistics (mirror of tvb_hpc/metric.py).

The reference describes a one-pass co-moment update (``Covar`` /
``OnlineCov``, metric.py:18-108), a batch covariance (``BatchCov``, 71-84)
and a cov -> corr normalisation (``CovToCorr``, 111-139) as OpenCL / loopy
kernels; the CUDA side has the same pair in
dsl/dsl_cuda/CUDAmodels/refs/covar.c:11-109.  Here they are device
operations of libtvbx in FP64:

* ``OnlineCov`` accumulates samples (n_samples, N, P) block by block with
  Chan et al.'s pairwise update of co-moments (exactly Welford's update with
  the count n = k + 1 at the k'th sample -- the reference kernels divide by
  the 0-based sample index, an off-by-one this package does not repeat);
* ``BatchCov`` is the same statistic over one batch;
* ``CovToCorr`` returns cov_ij / sqrt(cov_ii cov_jj).

The integrator itself folds its monitor samples into the same FP64
co-moments on the device (``Simulation(fc=True)``), so long runs never keep
their sample series.
"""

import numpy as np

from .hip import Context


class OnlineCov:
    """Streaming FC of P instances: ``update(samples)`` with samples of shape
    (n, N, P) (n may be 1), ``corr()`` -> (P, N, N)."""

    def __init__(self, n_node, n_inst, context=None):
        self.N, self.P = int(n_node), int(n_inst)
        self.ctx = context or Context.default()
        self._blocks = []
        self.n = 0

    def update(self, samples):
        s = np.asarray(samples, np.float32)
        if s.ndim == 2:
            s = s[None]
        if s.shape[1:] != (self.N, self.P):
            raise ValueError('samples must be (n, N, P)')
        self._blocks.append(s)
        self.n += s.shape[0]

    def corr(self):
        if self.n < 2:
            raise ValueError('FC needs at least two samples')
        return self.ctx.corr(np.concatenate(self._blocks))


class BatchCov:
    "metric.py:71-84: covariance statistics of one batch of samples"

    def __init__(self, context=None):
        self.ctx = context or Context.default()

    def corr(self, samples):
        return self.ctx.corr(samples)


class CovToCorr:
    "metric.py:111-139: cov_ij / sqrt(cov_ii cov_jj) of covariance stacks"

    @staticmethod
    def apply(cov):
        cov = np.asarray(cov, np.float64)
        d = np.sqrt(np.einsum('...ii->...i', cov))
        return cov / (d[..., :, None] * d[..., None, :])


def fc_sc_error(fc, weights):
    """Sum over node pairs i < j of (FC_ij - SC_ij)^2, the acceptance
    statistic of examples/hackathon.py:72-81; fc (..., N, N)."""
    N = weights.shape[0]
    r, c = np.triu_indices(N, 1)
    return ((np.asarray(fc)[..., r, c] - weights[r, c]) ** 2).sum(axis=-1)


def windowed_fc(trace, win):
    """FC of consecutive windows of a (T, N, P) trace: (T // win, P, N, N),
    numpy.corrcoef per window (hackathon.py:64-70; cuda_setup.py:104-122)."""
    T, N, P = trace.shape
    nw = T // win
    out = np.zeros((nw, P, N, N))
    for w in range(nw):
        x = trace[w * win:(w + 1) * win]
        for p in range(P):
            out[w, p] = np.corrcoef(x[:, :, p].T)
    return out


def windowed_fc_sc_error(sim, sc, n_windows, window_steps):
    """The windowed acceptance statistic of examples/hackathon.py:64-81 on the
    device: for each of n_windows consecutive windows of window_steps steps,
    the FC of that window's monitor samples (FP64 co-moments, restarted per
    window) and sum over i < j of (FC_ij - SC_ij)^2 per instance.  Only the
    (n_windows, P) errors leave the device.  ``sim`` is a
    ``hip.Simulation(fc=True)`` whose tavg period divides window_steps."""
    period = sim.desc.tavg_period
    if period <= 0 or window_steps % period:
        raise ValueError('window_steps must be a multiple of the tavg period')
    sim.fc_reset()               # the first window starts here
    err = np.zeros((n_windows, sim.P))
    for w in range(n_windows):
        sim.run(window_steps)
        err[w] = sim.fc_sc_error(sc)
        sim.fc_reset()
    return err


def sweep_trends(err, n_speed, n_coupling):
    """hackathon.py:82-87: the change of the FC-SC error along the speed and
    coupling axes of a (speed outer, coupling inner) sweep grid."""
    e = np.asarray(err).reshape((n_speed, n_coupling))
    return np.diff(e.mean(axis=1)).sum(), np.diff(e.mean(axis=0)).sum()

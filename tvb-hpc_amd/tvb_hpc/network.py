"""
Sparse delayed connectivity (mirror of tvb_hpc/network.py).

``Connectivity`` keeps the reference's dense -> CSR conversion
(network.py:40-93): ``row`` (N+1) / ``col`` (nnz) from scipy's CSR of the
weights, ``wnz`` / ``lnz`` the non-zero weights and lengths in row-major
order.  Row = target node i, column = source node j (pinned by the
reference's tests.py:211-244).

``Network`` pairs a model with a coupling function and describes

    input[i, k] = post_k( sum_{e in row(i)} w_e * pre_k(y_k[t - d_e, col_e],
                                                     y_k[t, i]) )

(network.py:115-136); the device kernel that evaluates it is emitted by
``tvb_hpc.emit``.

This module also holds the seeded synthetic-connectome generator used by the
sweep configurations and the integer delay computation.
"""

import logging

import numpy as np
from scipy import sparse

from .model import BaseModel
from .coupling import BaseCoupling

LOG = logging.getLogger(__name__)


class Connectivity:
    "Data class managing sparse connectivity (network.py:40-93)."

    def __init__(self, nnode, nz, nnz, col, row, wnz, lnz):
        self.nnode = nnode
        self.nz = nz
        self.nnz = nnz
        self.col = col
        self.row = row
        self.wnz = wnz
        self.lnz = lnz

    @classmethod
    def from_dense(cls, weights, lengths):
        "network.py:66-78"
        weights = np.asarray(weights)
        lengths = np.asarray(lengths)
        nnode = weights.shape[0]
        nz = ~(weights == 0)
        nnz = int(nz.sum())
        wnz = weights[nz]
        lnz = lengths[nz]
        sw = sparse.csr_matrix(weights)
        col = sw.indices.astype(np.uintc)
        row = sw.indptr.astype(np.uintc)
        obj = cls(nnode, nz, nnz, col, row, wnz, lnz)
        obj.weights = weights
        obj.lengths = lengths
        return obj

    @classmethod
    def from_npz(cls, fname, weights_key='weights', lengths_key='lengths'):
        "network.py:81-89"
        npz = np.load(fname, allow_pickle=False)
        return cls.from_dense(npz[weights_key], npz[lengths_key])

    @classmethod
    def hcp0(cls, fname='data/hcp0.npz'):
        "network.py:92-93 (path relative to the working directory)"
        return cls.from_npz(fname)

    def delays(self, speed, dt, policy='trunc'):
        """Integer delays of the non-zero edges for conduction ``speed``
        (mm/ms = m/s) and step ``dt`` (ms).  ``trunc`` reproduces
        examples/hackathon.py:44 (``astype(uintc)``); ``round`` is used by
        the sweep configurations."""
        return lengths_to_delays(self.lnz, speed, dt, policy)


def lengths_to_delays(lengths, speed, dt, policy='trunc'):
    x = np.asarray(lengths, np.float64) / float(speed) / float(dt)
    if policy == 'trunc':
        d = np.floor(x)
    elif policy == 'round':
        d = np.round(x)
    else:
        raise ValueError('unknown delay policy %r' % policy)
    return d.astype(np.uintc)


def bin_delays(delays, n_bins):
    """Quantise integer delays to at most n_bins distinct values (equal-width
    bins over [min, max], each delay mapped to its bin centre rounded to a
    step).  Returns (quantised delays, bin centres)."""
    d = np.asarray(delays, np.int64)
    lo, hi = int(d.min()), int(d.max())
    if hi - lo + 1 <= n_bins:
        return d.astype(np.int32), np.unique(d).astype(np.int32)
    edges = np.linspace(lo, hi + 1, n_bins + 1)
    idx = np.clip(np.searchsorted(edges, d, side='right') - 1, 0, n_bins - 1)
    centres = np.clip(np.round((edges[:-1] + edges[1:] - 1) / 2), lo, hi)
    centres = centres.astype(np.int32)
    return centres[idx].astype(np.int32), centres


def horizon_for(max_delay: int, block: int = 1) -> int:
    """Ring length for a maximum delay: slot n+1 is written while slots
    n - d (0 <= d <= max_delay) are read, so the ring needs max_delay + 2
    slots for the write never to alias a pending read.  A launch that
    advances `block` steps without synchronisation writes slots n+1 ..
    n+block while reading back to n - max_delay: max_delay + block + 1."""
    return int(max_delay) + int(block) + 1


class Network:
    """Delayed CSR coupling of a model through a coupling function
    (network.py:96-171)."""

    def __init__(self, model: BaseModel, cfun: BaseCoupling):
        self.model = model
        self.cfun = cfun

    @property
    def n_coupled(self):
        return len(self.cfun.io)

    def __repr__(self):
        return 'Network(%r, %r)' % (self.model, self.cfun)


# ---------------------------------------------------------------------------
# synthetic connectomes

def synth_connectome(nnode, density, seed, lengths='uniform', lmin=10.0,
                     lmax=200.0, radius=150.0, dtype=np.float32):
    """Seeded synthetic connectome.

    Upper-triangle Bernoulli(density) mask mirrored to a symmetric pattern
    with zero diagonal; lognormal(0, 1) weights drawn on the upper triangle,
    mirrored and scaled to max 1; lengths either uniform in [lmin, lmax] mm
    (symmetric) or, with ``lengths='ball'``, Euclidean distances between
    uniform random points in a ball of the given radius (mm).

    Returns dense (weights, lengths) arrays of shape (nnode, nnode)."""
    rng = np.random.default_rng(seed)
    iu = np.triu_indices(nnode, 1)
    mask_u = rng.random(iu[0].size) < density
    w_u = rng.lognormal(0.0, 1.0, size=iu[0].size)
    W = np.zeros((nnode, nnode))
    W[iu] = np.where(mask_u, w_u, 0.0)
    W = W + W.T
    if W.max() > 0:
        W /= W.max()
    if lengths == 'uniform':
        l_u = rng.uniform(lmin, lmax, size=iu[0].size)
        L = np.zeros((nnode, nnode))
        L[iu] = l_u
        L = L + L.T
    elif lengths == 'ball':
        # uniform points in a ball: direction x radius * U^(1/3)
        v = rng.normal(size=(nnode, 3))
        v /= np.linalg.norm(v, axis=1, keepdims=True)
        r = radius * rng.random(nnode) ** (1.0 / 3.0)
        pos = v * r[:, None]
        L = np.linalg.norm(pos[:, None, :] - pos[None, :, :], axis=-1)
    else:
        raise ValueError('unknown lengths mode %r' % lengths)
    np.fill_diagonal(L, 0.0)
    return W.astype(dtype), L.astype(dtype)


class CSR:
    """CSR connectome with per-group integer delays, in the layout the device
    kernels consume: ``row`` int32 (N+1), ``col`` int32 (nnz), ``weight``
    float32 (nnz) shared by all instances or (n_group, nnz) -- one row per
    delay group, for subjects that share one sparsity pattern with their own
    weights and delays (the group-sparse layout of examples/hcp100.py:22-53)
    -- and ``delays`` (n_group, nnz) int32."""

    def __init__(self, nnode, row, col, weight, delays):
        self.nnode = int(nnode)
        self.row = np.ascontiguousarray(row, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.weight = np.ascontiguousarray(weight, np.float32)
        d = np.asarray(delays)
        if d.ndim == 1:
            d = d[None, :]
        self.delays = np.ascontiguousarray(d, np.int32)
        self.validate()

    @property
    def nnz(self):
        return int(self.col.size)

    @property
    def n_group(self):
        return int(self.delays.shape[0])

    @property
    def weight_groups(self):
        "True when each delay group has its own weights"
        return self.weight.ndim == 2

    @property
    def max_delay(self):
        return int(self.delays.max()) if self.delays.size else 0

    def validate(self):
        N = self.nnode
        if self.row.shape != (N + 1, ):
            raise ValueError('row must have nnode+1 entries')
        if self.row[0] != 0 or self.row[-1] != self.col.size:
            raise ValueError('nnz (%d) disagrees with row[-1] (%d)'
                             % (self.col.size, self.row[-1]))
        if np.any(np.diff(self.row) < 0):
            raise ValueError('row must be non-decreasing')
        if self.col.size and (self.col.min() < 0 or self.col.max() >= N):
            raise ValueError('col out of range')
        if self.weight.shape != self.col.shape and \
                self.weight.shape != (self.delays.shape[0], self.col.size):
            raise ValueError('weights must be (nnz,) or (n_group, nnz)')
        if self.delays.shape[1] != self.col.size:
            raise ValueError('delays must have nnz entries per group')
        if self.delays.size and self.delays.min() < 0:
            raise ValueError('negative delay')

    @classmethod
    def from_dense(cls, weights, lengths, speeds, dt, policy='round'):
        "One delay group per conduction speed."
        conn = Connectivity.from_dense(weights, lengths)
        speeds = np.atleast_1d(speeds)
        d = np.stack([conn.delays(s, dt, policy) for s in speeds])
        return cls(conn.nnode, conn.row, conn.col, conn.wnz, d)

    def binned(self, n_bins=32):
        """Delays quantised to ``n_bins`` values: equal-width bins over
        [min, max] of all groups, each delay replaced by its bin's centre
        (rounded to a step).  The error per edge is at most half a bin.
        Returns (CSR, bin centres)."""
        d, centres = bin_delays(self.delays, n_bins)
        return CSR(self.nnode, self.row, self.col, self.weight, d), centres

    def sorted_by_delay(self):
        """The same connectome with each row's edges ordered by (delay of
        group 0, source): consecutive gathers of a row then read nearby
        ring slots.  (Changes the summation order, not the sum's terms.)"""
        perm = np.arange(self.nnz)
        for i in range(self.nnode):
            a, b = self.row[i], self.row[i + 1]
            if b - a > 1:
                key = np.lexsort((self.col[a:b], self.delays[0, a:b]))
                perm[a:b] = a + key
        return CSR(self.nnode, self.row, self.col[perm],
                   self.weight[..., perm], self.delays[:, perm])

    def dense_weights(self, group=0):
        "dense (N, N) weights (of one group when weights are grouped)"
        w = self.weight[group] if self.weight_groups else self.weight
        W = np.zeros((self.nnode, self.nnode), np.float64)
        for i in range(self.nnode):
            for e in range(self.row[i], self.row[i + 1]):
                W[i, self.col[e]] += w[e]
        return W

    @classmethod
    def group_sparse(cls, weights, lengths, speeds, dt, policy='trunc'):
        """Subjects on one sparsity pattern (examples/hcp100.py:22-53):
        weights, lengths (n_subj, N, N); the pattern is the edges present in
        every subject; each subject becomes one delay group with its own
        weights and delays (lengths / speed / dt, one speed per subject or
        one for all)."""
        W = np.asarray(weights, np.float64)
        L = np.asarray(lengths, np.float64)
        mask = (W != 0).all(axis=0)
        np.fill_diagonal(mask, False)
        i, j = np.nonzero(mask)      # row-major: rows ascending, cols within
        N = W.shape[1]
        row = np.zeros(N + 1, np.int32)
        np.add.at(row, i + 1, 1)
        row = np.cumsum(row).astype(np.int32)
        sp = np.broadcast_to(np.asarray(speeds, np.float64), (W.shape[0], ))
        ws = np.stack([w[i, j] for w in W]).astype(np.float32)
        ds = np.stack([lengths_to_delays(l[i, j], v, dt, policy)
                       for l, v in zip(L, sp)]).astype(np.int32)
        return cls(N, row, j.astype(np.int32), ws, ds)


class _RMapKernel:
    """Callable with the reference kernel's keyword arguments (nroi, nvar,
    nnode, rmap, node, roi); mutates roi (to_avg) or node (from_avg) in
    place.  A trailing instance axis on node and roi batches instances."""

    def __init__(self, target, to_avg):
        self.target, self.to_avg = target, to_avg
        self._maps = {}

    def _map(self, rmap, nroi):
        from .hip import RegionMap
        rm = np.ascontiguousarray(np.asarray(rmap).reshape(-1), np.int32)
        key = (rm.tobytes(), int(nroi))
        if key not in self._maps:
            ctx = self.target.context() if hasattr(self.target, 'context') \
                else None
            self._maps[key] = RegionMap(rm, nroi, context=ctx)
        return self._maps[key]

    def __call__(self, nroi, nvar, nnode, rmap, node, roi, **_):
        rm = np.asarray(rmap).reshape(-1)
        if rm.size != nnode:
            raise ValueError('rmap has %d entries for nnode = %d'
                             % (rm.size, nnode))
        m = self._map(rm, nroi)
        if np.shape(node)[:2] != (nnode, nvar) or \
                np.shape(roi)[:2] != (nroi, nvar):
            raise ValueError('node must be (nnode, nvar[, n]), roi (nroi, '
                             'nvar[, n])')
        if self.to_avg:
            roi[...] = m.to_avg(np.asarray(node, np.float32))
        else:
            node[...] = m.from_avg(np.asarray(roi, np.float32))


class RMapToAvg:
    """Region sums of vertex values: roi[rmap[k], j] += node[k, j] over k
    ascending, after zeroing roi (network.py:174-190 of the reference; the
    kernel sums, its name notwithstanding)."""

    def kernel(self, target):
        return _RMapKernel(target, True)


class RMapFromAvg:
    """Vertex values from region values: node[k, j] = roi[rmap[k], j]
    (network.py:193-202 of the reference)."""

    def kernel(self, target):
        return _RMapKernel(target, False)

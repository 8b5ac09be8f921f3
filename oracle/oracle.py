"""
ctypes wrapper of the CPU oracle (oracle/tvb_oracle.cpp).

TEST INFRASTRUCTURE ONLY: imported by tests/, __graft_entry__.smoke() and the
CPU-baseline leg of bench.py, as the checker.  The product package never
imports this module.

Arrays cross this boundary in the device kernels' convention (instance axis
innermost): state (S, N, P), ring (H, R, N, P), injected noise
(n_steps, S, N, P), samples (n_samples, N, P).  The oracle itself works
instance-major and the wrapper transposes.
"""

import ctypes
import os
import subprocess
import threading

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_build', 'libtvboracle.so')

MODEL_IDS = {'Kuramoto': 0, 'G2DO': 1, 'RWW': 2, 'JansenRit': 3,
             'MontbrioPazoRoxin': 4, 'Linear': 5, 'Generic2dOscillator': 6,
             'ReducedWongWang': 7}
CFUN_IDS = {'Linear': 0, 'Sigmoidal': 1, 'Diff': 2, 'Kuramoto': 3}

# parameter order of each hand-written model = the model's const dict order
MODEL_PARAMS = {
    'Kuramoto': ['omega'],
    'G2DO': ['tau', 'I', 'a', 'b', 'c', 'd', 'e', 'f', 'g', 'alpha', 'beta',
             'gamma'],
    'RWW': ['a', 'b', 'd', 'g', 'ts', 'J', 'w', 'io'],
    'JansenRit': ['A', 'B', 'a', 'b', 'v0', 'nu_max', 'r', 'J', 'a_1', 'a_2',
                  'a_3', 'a_4', 'mu'],
    'MontbrioPazoRoxin': ['tau', 'Delta', 'J', 'I', 'eta'],
    'Linear': ['lam'],
}
MODEL_PARAMS['Generic2dOscillator'] = MODEL_PARAMS['G2DO']
MODEL_DEFAULTS = {
    'Kuramoto': {'omega': 1.0},
    'G2DO': {'tau': 1.0, 'I': 0.0, 'a': -2.0, 'b': -10.0, 'c': 0.0, 'd': 0.02,
             'e': 3.0, 'f': 1.0, 'g': 0.0, 'alpha': 1.0, 'beta': 1.0,
             'gamma': 1.0},
    'RWW': {'a': 0.270, 'b': 0.108, 'd': 154.0, 'g': 0.641, 'ts': 100.0,
            'J': 0.2609, 'w': 0.6, 'io': 0.33},
    'JansenRit': {'A': 3.25, 'B': 22.0, 'a': 0.1, 'b': 0.05, 'v0': 5.52,
                  'nu_max': 0.0025, 'r': 0.56, 'J': 135.0, 'a_1': 1.0,
                  'a_2': 0.8, 'a_3': 0.25, 'a_4': 0.25, 'mu': 0.22},
    'MontbrioPazoRoxin': {'tau': 1.0, 'Delta': 1.0, 'J': 15.0, 'I': 0.0,
                          'eta': -5.0},
    'Linear': {'lam': -1.0},
}
MODEL_DEFAULTS['Generic2dOscillator'] = MODEL_DEFAULTS['G2DO']
MODEL_PARAMS['ReducedWongWang'] = MODEL_PARAMS['RWW']
MODEL_DEFAULTS['ReducedWongWang'] = MODEL_DEFAULTS['RWW']
MODEL_NSVAR = {'Kuramoto': 1, 'G2DO': 2, 'RWW': 1, 'JansenRit': 6,
               'MontbrioPazoRoxin': 2, 'Linear': 1, 'Generic2dOscillator': 2,
               'ReducedWongWang': 1}
MODEL_NOBS = {'Kuramoto': 2, 'G2DO': 2, 'RWW': 1, 'JansenRit': 1,
              'MontbrioPazoRoxin': 2, 'Linear': 1, 'Generic2dOscillator': 2,
              'ReducedWongWang': 1}
CFUN_PARAMS = {
    'Linear': ['a', 'b'], 'Diff': ['a', 'b'], 'Kuramoto': ['a', 'b'],
    'Sigmoidal': ['cmin', 'cmax', 'midpoint', 'r', 'a'],
}
CFUN_DEFAULTS = {
    'Linear': {'a': 1e-3, 'b': 0.0}, 'Diff': {'a': 1e-3, 'b': 0.0},
    'Kuramoto': {'a': 1e-3, 'b': 0.0},
    'Sigmoidal': {'cmin': 0.0, 'cmax': 0.005, 'midpoint': 6.0, 'r': 1.0,
                  'a': 0.56},
}


class OracleSpec(ctypes.Structure):
    _fields_ = [
        ('model', ctypes.c_int), ('cfun', ctypes.c_int),
        ('scheme', ctypes.c_int),
        ('N', ctypes.c_int), ('nnz', ctypes.c_int), ('P', ctypes.c_int),
        ('H', ctypes.c_int), ('R', ctypes.c_int),
        ('n_par', ctypes.c_int), ('n_cpar', ctypes.c_int),
        ('step0', ctypes.c_longlong), ('n_steps', ctypes.c_int),
        ('group_size', ctypes.c_int), ('n_group', ctypes.c_int),
        ('row', ctypes.c_void_p), ('col', ctypes.c_void_p),
        ('delay', ctypes.c_void_p),
        ('weight', ctypes.c_void_p), ('par', ctypes.c_void_p),
        ('cpar', ctypes.c_void_p), ('sigma', ctypes.c_void_p),
        ('dt', ctypes.c_double),
        ('noise_mode', ctypes.c_int), ('seed_lo', ctypes.c_uint),
        ('seed_hi', ctypes.c_uint),
        ('noise', ctypes.c_void_p), ('noise_tstride', ctypes.c_longlong),
        ('tavg_period', ctypes.c_int), ('bold', ctypes.c_int),
        ('bold_obs', ctypes.c_int), ('bold_dt', ctypes.c_double),
        ('state', ctypes.c_void_p), ('ring', ctypes.c_void_p),
        ('traj', ctypes.c_void_p), ('tavg', ctypes.c_void_p),
        ('bold_state', ctypes.c_void_p), ('samples', ctypes.c_void_p),
        ('sample_cap', ctypes.c_longlong),
        ('samples_written', ctypes.c_longlong),
        ('n_threads', ctypes.c_int),
        ('order', ctypes.c_int), ('stale', ctypes.c_void_p),
        ('ring_q', ctypes.c_int), ('ring_pre', ctypes.c_int),
        ('ring_scale', ctypes.c_double),
        ('inst_off', ctypes.c_int),
        ('weight_groups', ctypes.c_int),
        ('force', ctypes.c_void_p), ('logp', ctypes.c_void_p),
    ]


def quantize(x, ring_dtype):
    """round-to-nearest-even to bfloat16 / float16, returned as float32;
    'u8': the integer round(255 clamp(x, 0, 1))"""
    x = np.asarray(x, np.float32)
    if ring_dtype == 'u8':
        c = np.minimum(np.maximum(x, np.float32(0)), np.float32(1))
        return np.rint(c * np.float32(255)).astype(np.float32)
    if ring_dtype == 'f16':
        return x.astype(np.float16).astype(np.float32)
    if ring_dtype == 'bf16':
        u = x.view(np.uint32).astype(np.uint64)
        u = (u + 0x7fff + ((u >> 16) & 1)) & 0xffff0000
        return u.astype(np.uint32).view(np.float32)
    return x


_lib = None
_native = {}


def build(force=False):
    if force or not os.path.exists(LIB_PATH):
        subprocess.check_call(['make', '-s', '-C', HERE])
    return LIB_PATH


def _host_key():
    """Short key of this host's CPU model and flags: a -march=native build is
    only valid on the kind of CPU it was compiled on."""
    import hashlib
    h = hashlib.sha256()
    try:
        with open('/proc/cpuinfo') as fd:
            for line in fd:
                if line.startswith(('model name', 'flags')):
                    h.update(line.encode())
                    if line.startswith('flags'):
                        break
    except OSError:
        pass
    return h.hexdigest()[:12]


def build_native(fast_math=False):
    """The oracle compiled for this host as the reference compiles its C
    path (examples/hcp100.py:65-72): -O3 -march=native -fopenmp, with or
    without -ffast-math.  Used only as the timed CPU baseline, never as a
    checker (contraction and fast-math change the float32 rounding)."""
    sub = 'native_%s%s' % (_host_key(), '_fm' if fast_math else '')
    out = os.path.join(HERE, '_build', sub, 'libtvboracle.so')
    with _build_lock(sub):
        if not os.path.exists(out):
            subprocess.check_call(['make', '-s', '-C', HERE, 'native',
                                   'NATIVE_OUT=_build/' + sub,
                                   'NATIVE_FLAGS=' + ('-ffast-math'
                                                      if fast_math else '')])
    return out


_locks = {}
_locks_guard = threading.Lock()


def _build_lock(name):
    "one build of each library at a time in this process"
    with _locks_guard:
        return _locks.setdefault(name, threading.Lock())


def load_library(path=None):
    """ctypes handle of an oracle build (default: the checker build)."""
    path = path or LIB_PATH
    if path in _native:
        return _native[path]
    if not os.path.exists(path):
        build()
    L = ctypes.CDLL(path)
    L.oracle_run_f32.argtypes = [ctypes.POINTER(OracleSpec)]
    L.oracle_run_f64.argtypes = [ctypes.POINTER(OracleSpec)]
    L.oracle_spec_size.restype = ctypes.c_int
    L.oracle_philox4x32_10.argtypes = [ctypes.c_void_p] * 3
    L.oracle_normals_f32.argtypes = [ctypes.c_longlong, ctypes.c_uint,
                                     ctypes.c_uint, ctypes.c_void_p]
    L.oracle_step_normal_f64.restype = ctypes.c_double
    L.oracle_step_normal_f64.argtypes = [
        ctypes.c_uint, ctypes.c_uint, ctypes.c_longlong, ctypes.c_int,
        ctypes.c_int, ctypes.c_int]
    L.oracle_step_normal_f32.restype = ctypes.c_float
    L.oracle_step_normal_f32.argtypes = L.oracle_step_normal_f64.argtypes
    L.oracle_bold_f64.argtypes = [ctypes.c_void_p, ctypes.c_longlong,
                                  ctypes.c_double, ctypes.c_void_p,
                                  ctypes.c_void_p]
    L.oracle_fc_f64.argtypes = [ctypes.c_void_p, ctypes.c_longlong,
                                ctypes.c_int, ctypes.c_int, ctypes.c_void_p]
    assert L.oracle_spec_size() == ctypes.sizeof(OracleSpec)
    _native[path] = L
    return L


def lib():
    global _lib
    if _lib is None:
        _lib = load_library(LIB_PATH)
    return _lib


def _ptr(a):
    return None if a is None else a.ctypes.data


def philox4x32_10(ctr, key):
    ctr = np.ascontiguousarray(ctr, np.uint32)
    key = np.ascontiguousarray(key, np.uint32)
    out = np.zeros(4, np.uint32)
    lib().oracle_philox4x32_10(_ptr(ctr), _ptr(key), _ptr(out))
    return out


def normals_f32(n, seed=0):
    out = np.zeros(n, np.float32)
    lib().oracle_normals_f32(n, seed & 0xffffffff, (seed >> 32) & 0xffffffff,
                             _ptr(out))
    return out


def step_normal(seed, n, i, p, s, dtype=np.float64):
    f = lib().oracle_step_normal_f64 if dtype == np.float64 else \
        lib().oracle_step_normal_f32
    return f(seed & 0xffffffff, (seed >> 32) & 0xffffffff, n, i, p, s)


def bold_f64(x, dt, state=None):
    x = np.ascontiguousarray(x, np.float64)
    b = np.array([0.0, 1.0, 1.0, 1.0]) if state is None else \
        np.ascontiguousarray(state, np.float64).copy()
    out = np.zeros_like(x)
    lib().oracle_bold_f64(_ptr(x), x.size, float(dt), _ptr(b), _ptr(out))
    return out, b


def fc_f64(samples):
    "samples (n, N, P) -> corr (P, N, N)"
    s = np.ascontiguousarray(samples, np.float64)
    n, N, P = s.shape
    out = np.zeros((P, N, N))
    lib().oracle_fc_f64(_ptr(s), n, N, P, _ptr(out))
    return out


def rmap_to_avg(rmap, node, nroi):
    """Region sums (RMapToAvg, tvb_hpc/network.py:174-190): roi zeroed, then
    roi[rmap[k]] += node[k] for k = 0, 1, ... in order, in node's dtype.
    node (nnode, nvar[, P]) -> roi (nroi, nvar[, P])."""
    node = np.asarray(node)
    roi = np.zeros((nroi, ) + node.shape[1:], node.dtype)
    # unbuffered: repeated indices accumulate one at a time, in index order
    np.add.at(roi, np.asarray(rmap).reshape(-1), node)
    return roi


def rmap_from_avg(rmap, roi):
    """RMapFromAvg (tvb_hpc/network.py:193-202): node[k] = roi[rmap[k]]."""
    return np.asarray(roi)[np.asarray(rmap).reshape(-1)]


def _expand_par(names, defaults, overrides, N, P, dtype, per_node=True):
    """Parameter block [len(names)][N][P] (or [len(names)][P])."""
    shape = (len(names), N, P) if per_node else (len(names), P)
    out = np.zeros(shape, dtype)
    for k, nm in enumerate(names):
        v = np.asarray(overrides.get(nm, defaults[nm]), np.float64)
        if per_node:
            if v.ndim == 0:
                out[k] = v
            elif v.shape == (P, ):
                out[k] = v[None, :]
            elif v.shape == (N, ):
                out[k] = v[:, None]
            elif v.shape == (N, P):
                out[k] = v
            else:
                raise ValueError('bad shape for %s: %s' % (nm, v.shape))
        else:
            if v.ndim == 0:
                out[k] = v
            elif v.shape == (P, ):
                out[k] = v
            else:
                raise ValueError('coupling param %s must be scalar or (P,)'
                                 % nm)
    return out


class OracleSim:
    """One batch of P instances of a (model, coupling) network, stepped on the
    CPU.  ``dtype`` selects the float32 parity build or the float64 accuracy
    build."""

    def __init__(self, model, cfun, N, row, col, weight, delays, P, H,
                 dt, scheme='euler', params=None, cparams=None, sigma=None,
                 group_size=None, R=1, dtype=np.float64, tavg_period=0,
                 bold=False, bold_obs=0, n_threads=0, order='cuda',
                 ring_dtype='f32', ring_pre=False, ring_scale=1.0,
                 inst_off=0, library=None):
        self.model, self.cfun = model, cfun
        # ctypes handle of the oracle build to run (default: the checker)
        self._lib = library
        self.order = {'cuda': 0, 'reference': 1}[order]
        # Philox counters of instance p use inst_off + p
        self.inst_off = int(inst_off)
        # emulation of a reduced-precision history (float32 oracle only)
        self.ring_dtype = ring_dtype
        self.ring_q = {'f32': 0, 'bf16': 1, 'f16': 2, 'u8': 3}[ring_dtype]
        if self.ring_q == 3 and not ring_pre:
            raise ValueError('an 8-bit history holds ring_scale * pre(y)')
        self.ring_pre = bool(ring_pre)
        self.ring_scale = float(ring_scale)
        self.N, self.P, self.H, self.R = int(N), int(P), int(H), int(R)
        self.S = MODEL_NSVAR[model]
        self.dtype = dtype
        self.row = np.ascontiguousarray(row, np.int32)
        self.col = np.ascontiguousarray(col, np.int32)
        self.nnz = int(self.col.size)
        d = np.asarray(delays, np.int32)
        if d.ndim == 1:
            d = d[None, :]
        self.delay = np.ascontiguousarray(d)
        self.group_size = int(group_size or P)
        if self.delay.shape[0] * self.group_size < P:
            raise ValueError('not enough delay groups for P instances')
        if self.delay.size and self.delay.max() > H - 2:
            raise ValueError('max delay %d needs a ring of at least %d slots'
                             % (self.delay.max(), self.delay.max() + 2))
        # weight (nnz,) shared, or (n_group, nnz): one row per delay group
        w = np.asarray(weight)
        self.weight_groups = int(w.ndim == 2)
        if self.weight_groups and w.shape != self.delay.shape:
            raise ValueError('grouped weights must match delays (n_group, '
                             'nnz)')
        self.weight = np.ascontiguousarray(w, dtype)
        params = params or {}
        cparams = cparams or {}
        self.par = _expand_par(MODEL_PARAMS[model], MODEL_DEFAULTS[model],
                               params, N, P, dtype)
        self.cpar = _expand_par(CFUN_PARAMS[cfun], CFUN_DEFAULTS[cfun],
                                cparams, N, P, dtype, per_node=False)
        self.sigma = None if sigma is None else np.ascontiguousarray(
            np.broadcast_to(np.asarray(sigma, np.float64), (P, )), dtype)
        self.dt = float(dt)
        self.scheme = {'euler': 0, 'em': 1}[scheme]
        self.state = np.zeros((P, self.S, N), dtype)
        self.ring = np.zeros((P, H, R, N), dtype)
        self.tavg = np.zeros((P, N), dtype)
        bs = np.zeros((P, N, 4), dtype)
        bs[..., 1:] = 1.0
        self.bold_state = bs
        self.tavg_period = int(tavg_period)
        self.bold = bool(bold)
        self.bold_obs = int(bold_obs)
        self.step = 0
        self.n_threads = n_threads
        # order 'reference': content of the slot being rewritten, by parity
        self.stale = np.zeros((P, 2, R, N), dtype)

    # -- state in kernel layout ------------------------------------------

    def set_state(self, x):
        "x: (S, N, P)"
        self.state[...] = np.transpose(np.asarray(x), (2, 0, 1))

    def get_state(self):
        return np.ascontiguousarray(np.transpose(self.state, (1, 2, 0)))

    def get_ring(self):
        """(H, R, N, P).  With order 'reference' the slot of the current step
        shows its not-yet-rewritten content, as the reference's obsrv array
        would after the same steps."""
        r = np.transpose(self.ring, (1, 2, 3, 0)).copy()
        if self.order == 1:
            r[self.step % self.H] = np.transpose(
                self.stale[:, self.step & 1], (1, 2, 0))
        return np.ascontiguousarray(r)

    def set_ring(self, r):
        "(H, R, N, P), e.g. the reference's obsrv array"
        self.ring[...] = np.transpose(np.asarray(r), (3, 0, 1, 2))
        self.stale[:, self.step & 1] = self.ring[:, self.step % self.H]

    def init_history(self, mode='zeros'):
        """Slot (step mod H) holds y(step) = h(x(step)); the other slots hold
        zeros (``zeros``, the reference's np.zeros obsrv buffer) or the same
        value (``constant``); ``current`` rewrites only that slot.  With
        order 'reference' the slot's previous content is kept as what the
        step reads for post_syn and zero delays."""
        y = self.observe()  # (R, N, P)
        if self.ring_q and not self.ring_pre:
            y = quantize(y, self.ring_dtype).astype(self.dtype)
        yy = np.transpose(y, (2, 0, 1))  # (P, R, N)
        t = self.step % self.H
        if mode == 'zeros':
            self.ring[...] = 0
        elif mode == 'constant':
            self.ring[...] = yy[:, None]
        elif mode != 'current':
            raise ValueError(mode)
        self.stale[:, self.step & 1] = self.ring[:, t]
        self.ring[:, t] = yy

    def observe(self):
        "observables of the current state, (R, N, P)"
        x = self.get_state().astype(np.float64)
        if self.model == 'Kuramoto':
            ys = [x[0], np.sin(x[0])]
        elif self.model in ('G2DO', 'MontbrioPazoRoxin',
                            'Generic2dOscillator'):
            ys = [x[0], x[1]]
        elif self.model == 'JansenRit':
            ys = [x[0] - x[1]]
        else:
            ys = [x[0]]
        return np.stack(ys[:self.R]).astype(self.dtype)

    # -- stepping ---------------------------------------------------------

    def logp(self, traj):
        """Euler-Maruyama log density (float64, per instance) of the
        transitions of traj (n + 1, S, N, P) under this network, with
        teacher forcing (scheme.py:42-48); the state ends at traj[-1]."""
        if self.scheme != 1:
            raise ValueError('the transition density needs scheme em')
        out = np.zeros(self.P)
        self.run(np.shape(traj)[0] - 1, force=traj, logp=out)
        return out

    def run(self, n_steps, noise=None, seed=None, traj=False,
            sample_cap=None, force=None, logp=None):
        """Advance n_steps.  ``noise``: injected (n_steps, S, N, P) array,
        or (S, N, P) reused every step; ``seed``: Philox stream.  Returns the
        trajectory (n_steps, S, N, P) if ``traj`` and the monitor samples
        (n, N, P) if a monitor is configured."""
        sp = OracleSpec()
        sp.model = MODEL_IDS[self.model]
        sp.cfun = CFUN_IDS[self.cfun]
        sp.scheme = self.scheme
        sp.N, sp.nnz, sp.P, sp.H, sp.R = self.N, self.nnz, self.P, self.H, \
            self.R
        sp.n_par, sp.n_cpar = self.par.shape[0], self.cpar.shape[0]
        sp.step0 = self.step
        sp.n_steps = int(n_steps)
        sp.group_size, sp.n_group = self.group_size, self.delay.shape[0]
        sp.row, sp.col, sp.delay = _ptr(self.row), _ptr(self.col), \
            _ptr(self.delay)
        sp.weight, sp.par, sp.cpar = _ptr(self.weight), _ptr(self.par), \
            _ptr(self.cpar)
        sp.sigma = _ptr(self.sigma)
        sp.dt = self.dt
        keep = []
        if noise is not None:
            nz = np.ascontiguousarray(noise, self.dtype)
            keep.append(nz)
            sp.noise_mode = 2
            sp.noise = _ptr(nz)
            sp.noise_tstride = 0 if nz.ndim == 3 else \
                self.S * self.N * self.P
        elif seed is not None:
            sp.noise_mode = 1
            sp.seed_lo = seed & 0xffffffff
            sp.seed_hi = (seed >> 32) & 0xffffffff
        else:
            sp.noise_mode = 0
        sp.tavg_period = self.tavg_period
        sp.bold = int(self.bold)
        sp.bold_obs = self.bold_obs
        sp.bold_dt = self.tavg_period * self.dt * 1e-3
        sp.state, sp.ring = _ptr(self.state), _ptr(self.ring)
        tr = None
        if traj:
            tr = np.zeros((n_steps, self.S, self.N, self.P), self.dtype)
            sp.traj = _ptr(tr)
        sp.tavg, sp.bold_state = _ptr(self.tavg), _ptr(self.bold_state)
        smp = None
        if self.tavg_period > 0:
            cap = sample_cap or (n_steps // self.tavg_period + 1)
            smp = np.zeros((cap, self.N, self.P), self.dtype)
            sp.samples = _ptr(smp)
            sp.sample_cap = cap
        sp.n_threads = self.n_threads
        sp.order = self.order
        sp.stale = _ptr(self.stale)
        sp.ring_q = self.ring_q
        sp.ring_pre = int(self.ring_pre)
        sp.ring_scale = self.ring_scale
        sp.inst_off = self.inst_off
        sp.weight_groups = self.weight_groups
        if force is not None:
            # forced steps: frames (n_steps + 1, S, N, P); logp (P,) float64
            fr = np.ascontiguousarray(force, self.dtype)
            if fr.shape != (n_steps + 1, self.S, self.N, self.P):
                raise ValueError('force must be (n_steps + 1, S, N, P)')
            keep.append(fr)
            sp.force = _ptr(fr)
            sp.logp = _ptr(logp)
        L = self._lib or lib()
        f = L.oracle_run_f64 if self.dtype == np.float64 else \
            L.oracle_run_f32
        rc = f(ctypes.byref(sp))
        if rc != 0:
            raise RuntimeError('oracle_run failed (%d)' % rc)
        self.step += int(n_steps)
        out = {}
        if traj:
            out['traj'] = tr
        if smp is not None:
            out['samples'] = smp[:sp.samples_written]
        return out


def for_config(cfg, idx, dtype=np.float32, H=None, n_threads=0,
               library=None, emulate_ring=True):
    """OracleSim for instances ``idx`` of a sweep configuration (an object
    with the attributes of ``tvb_hpc.configs.SweepConfig``): each instance
    gets its own delay row (and weight row, for grouped weights), its
    parameters, sigma and initial state; with a contiguous ``idx`` the
    Philox counters are those of the device's instance numbering.  With
    ``emulate_ring`` a reduced-precision history of the config is
    reproduced (float32 oracle only)."""
    idx = np.asarray(idx, np.int64)
    csr = cfg.csr
    gs = cfg.group_size or cfg.n_inst
    grp = idx // gs if csr.n_group > 1 else np.zeros_like(idx)
    delays = np.asarray(csr.delays)[grp]
    weight = np.asarray(csr.weight)[grp] if csr.weight_groups \
        else csr.weight
    mp, cp = {}, {}
    for k, v in cfg.params.items():
        if k == 'sigma':
            continue
        if ':' in k:
            scope, nm = k.split(':', 1)
        else:
            scope, nm = ('c' if k in cfg.cfun.param else 'm'), k
        a = np.asarray(v)
        val = a[..., idx] if a.ndim >= 1 and a.shape[-1] == cfg.n_inst \
            else a
        (mp if scope == 'm' else cp)[nm] = val
    cp.update({k: v for k, v in cfg.cfun.param.items()
               if k not in cp and not hasattr(v, 'name')})
    mp.update({k: v for k, v in cfg.model.const.items()
               if k not in mp and k in MODEL_PARAMS.get(
                   type(cfg.model).__name__, ()) and np.ndim(v) == 0
               and not hasattr(v, 'name')})
    sigma = getattr(cfg.scheme, 'sigma', None)
    if sigma is not None and 'sigma' in cfg.params:
        sigma = np.asarray(cfg.params['sigma'])[idx]
    ring = dict(ring_dtype='f32')
    if emulate_ring and getattr(cfg, 'ring_dtype', 'f32') != 'f32':
        ring = dict(ring_dtype=cfg.ring_dtype, ring_pre=True,
                    ring_scale=cfg.ring_scale)
    contiguous = idx.size and np.all(np.diff(idx) == 1)
    ref = OracleSim(type(cfg.model).__name__, type(cfg.cfun).__name__,
                    csr.nnode, csr.row, csr.col, weight, delays, idx.size,
                    H or csr.max_delay + 2, cfg.scheme.dt,
                    scheme='euler' if sigma is None else 'em', params=mp,
                    cparams=cp, sigma=sigma, group_size=1, dtype=dtype,
                    tavg_period=cfg.tavg_period, bold=cfg.bold,
                    bold_obs=cfg.bold_obs, n_threads=n_threads,
                    inst_off=int(idx[0]) if contiguous else 0,
                    library=library, **ring)
    ref.set_state(cfg.initial_state()[..., idx])
    ref.init_history('zeros')
    return ref

"""
HIP target: ctypes bindings of libtvbx.so (include/tvbx.h) and the two ways
of driving the device integrator.

``HipTarget``
    The drop-in for the reference's loopy targets (tvb_hpc/numba.py:10-50,
    compiler.py:228-251).  ``transforms.network_time_step(model, cfun,
    scheme, target=HipTarget())`` returns a ``NetworkStepExecutor`` callable
    with the reference's keyword arguments (examples/hackathon.py:49-53); host
    arrays are mutated in place, like the reference's kernels.

``Simulation``
    The device-resident sweep path: P instances of one (model, coupling,
    scheme) network, with state, history ring, parameters, monitor and FP64
    FC living in HBM between calls.

There is no CPU fallback: without libtvbx.so or without a GPU these classes
raise.
"""

import ctypes
import logging
import os
import sys
import threading

import numpy as np

from . import emit
from . import expr as E
from .network import CSR, horizon_for

LOG = logging.getLogger(__name__)
HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, '_lib', 'libtvbx.so')
KDIR = os.path.join(HERE, '_lib', 'kernels')


class TvbxError(RuntimeError):
    "Raised when a libtvbx call returns a non-zero status."


class SimDesc(ctypes.Structure):
    "mirror of tvbx_sim_desc (include/tvbx.h)"
    _fields_ = [(n, ctypes.c_int) for n in (
        'n_node', 'n_inst', 'n_svar', 'n_ring', 'n_coupled', 'horizon', 'nnz',
        'n_group', 'group_size', 'n_pinst', 'n_pnode', 'n_pni', 'n_pglob')] + \
        [('dt', ctypes.c_float)] + [(n, ctypes.c_int) for n in (
            'scheme', 'tavg_period', 'bold', 'bold_obs', 'fc', 'sample_slots',
            'waves', 'node_block', 'steps_per_launch', 'compat_outputs',
            'order', 'ipl', 'ring_type')] + [('ring_scale', ctypes.c_float),
                                             ('xcd_map', ctypes.c_int),
                                             ('weight_groups', ctypes.c_int)]


class BatchDesc(ctypes.Structure):
    "Mirror of tvbx_batch_desc (include/tvbx.h)."
    _fields_ = [(n, ctypes.c_int) for n in (
        'n_subject', 'n_svar', 'n_ring', 'n_coupled', 'n_pinst', 'n_pnode',
        'n_pni', 'n_pglob')] + [('dt', ctypes.c_float)] + [
        (n, ctypes.c_int) for n in ('scheme', 'waves', 'node_block',
                                    'steps_per_launch', 'ipl',
                                    'ring_type')] + [
        ('ring_scale', ctypes.c_float)]


class SubjectDesc(ctypes.Structure):
    "Mirror of tvbx_subject_desc (include/tvbx.h)."
    _fields_ = [(n, ctypes.c_int) for n in ('n_node', 'n_inst', 'nnz',
                                            'horizon')]


_P = ctypes.c_void_p
_I = ctypes.c_int
_LL = ctypes.c_longlong
_SZ = ctypes.c_size_t
_SIGS = {
    'tvbx_abi_version': [],
    'tvbx_last_error': [ctypes.c_char_p, _SZ],
    'tvbx_device_count': [_P],
    'tvbx_ctx_create': [_I, _P],
    'tvbx_ctx_destroy': [_P],
    'tvbx_ctx_sync': [_P],
    'tvbx_ctx_device_info': [_P, ctypes.c_char_p, _SZ, _P, _P],
    'tvbx_module_compile': [_P, ctypes.c_char_p, ctypes.c_char_p, _P,
                            ctypes.c_char_p, _SZ],
    'tvbx_module_load': [_P, _P, _SZ, _P],
    'tvbx_module_image': [_P, _P, _SZ, _P],
    'tvbx_module_destroy': [_P],
    'tvbx_sim_create': [_P, _P, ctypes.POINTER(SimDesc), _P],
    'tvbx_sim_destroy': [_P],
    'tvbx_sim_set_connectome': [_P, _P, _P, _P, _P],
    'tvbx_sim_set_params': [_P, _I, _P, _SZ],
    'tvbx_sim_set_state': [_P, _P],
    'tvbx_sim_get_state': [_P, _P],
    'tvbx_sim_set_ring': [_P, _P],
    'tvbx_sim_get_ring': [_P, _P],
    'tvbx_sim_set_step': [_P, _LL],
    'tvbx_sim_get_step': [_P, _P],
    'tvbx_sim_init_history': [_P, _I],
    'tvbx_sim_set_noise': [_P, _I, ctypes.c_ulonglong],
    'tvbx_sim_run': [_P, _LL, _P, _LL],
    'tvbx_sim_set_exec': [_P, _I, _I, _I],
    'tvbx_sim_set_coupling': [_P, _I],
    'tvbx_sim_set_pacing': [_P, _I],
    'tvbx_sim_set_checkpoint_reuse': [_P, _I],
    'tvbx_sim_checkpoint_run': [_P, _LL, _I],
    'tvbx_sim_adjoint': [_P, _P, _I, _LL, _LL, _I, _P, _I, _P, _I],
    'tvbx_sim_get_outputs': [_P, _P, _P, _P],
    'tvbx_sim_get_samples': [_P, _P, _P, _P],
    'tvbx_sim_fc': [_P, _I, _I, _P, _P],
    'tvbx_sim_fc_reset': [_P],
    'tvbx_sim_fc_sc_error': [_P, _P, _I, _I, _P, _P],
    'tvbx_sim_logp': [_P, _P, _LL, _I, _P, _I, _P, _I, _P],
    'tvbx_sim_get_flags': [_P, _P],
    'tvbx_sim_timing': [_P, _P, _P, _I],
    'tvbx_rng_normals': [_P, ctypes.c_ulonglong, _P, _SZ],
    'tvbx_corr': [_P, _P, _LL, _I, _I, _P],
    'tvbx_corr_noalloc': [_P, _P, _LL, _I, _I, _P],
    'tvbx_fc_loss': [_P, _P, _LL, _I, _I, _P, _I, _P, _P, _P],
    'tvbx_transpose': [_P, _P, _LL, _LL, _P, _I, _P],
    'tvbx_stream_rates': [_P, _SZ, _P, _P],
    'tvbx_rmap_create': [_P, _P, _I, _I, _P],
    'tvbx_rmap_destroy': [_P],
    'tvbx_rmap_to_avg': [_P, _I, _I, _P, _P],
    'tvbx_rmap_from_avg': [_P, _I, _I, _P, _P],
    'tvbx_batch_create': [_P, _P, ctypes.POINTER(BatchDesc),
                          ctypes.POINTER(SubjectDesc), _P],
    'tvbx_batch_destroy': [_P],
    'tvbx_batch_set_connectome': [_P, _I, _P, _P, _P, _P],
    'tvbx_batch_set_params': [_P, _I, _P, _SZ],
    'tvbx_batch_set_state': [_P, _I, _P],
    'tvbx_batch_get_state': [_P, _I, _P],
    'tvbx_batch_set_step': [_P, _LL],
    'tvbx_batch_get_step': [_P, _P],
    'tvbx_batch_set_noise': [_P, _I, ctypes.c_ulonglong],
    'tvbx_batch_init_history': [_P, _I],
    'tvbx_batch_set_exec': [_P, _I, _I, _I],
    'tvbx_batch_run': [_P, _LL],
    'tvbx_batch_sync': [_P],
    'tvbx_batch_get_flags': [_P, _P],
}

_lib = None
_lib_lock = threading.Lock()


def _one_hip_runtime():
    """PyTorch-ROCm ships its own HIP and HSA runtimes (same sonames as
    /opt/rocm's).  Loaded after libtvbx they would be a second runtime in the
    process, which then finds no GPU; loaded first, libtvbx binds to them.
    So load PyTorch's, when it is installed, before libtvbx."""
    try:
        import torch  # noqa: F401
    except ImportError:
        pass


def lib():
    """Load libtvbx.so (never builds; run ``tvb_hpc.build`` first)."""
    global _lib
    with _lib_lock:
        if _lib is None:
            if not os.path.exists(LIB_PATH):
                raise TvbxError('libtvbx.so not built (%s); run '
                                'python -m tvb_hpc.build' % LIB_PATH)
            _one_hip_runtime()
            L = ctypes.CDLL(LIB_PATH)
            for name, args in _SIGS.items():
                fn = getattr(L, name)
                fn.argtypes = args
                fn.restype = ctypes.c_int
            if L.tvbx_abi_version() != 3:
                raise TvbxError('libtvbx ABI version mismatch')
            _lib = L
    return _lib


def check(rc):
    if rc != 0:
        buf = ctypes.create_string_buffer(4096)
        lib().tvbx_last_error(buf, 4096)
        raise TvbxError('libtvbx error %d: %s' % (rc, buf.value.decode()))


def _ptr(a):
    return None if a is None else a.ctypes.data


def device_count():
    n = ctypes.c_int(0)
    check(lib().tvbx_device_count(ctypes.byref(n)))
    return n.value


class Context:
    "A device and its stream (one per host thread)."

    _default = {}

    def __init__(self, device=0):
        self.device = device
        self.h = ctypes.c_void_p()
        check(lib().tvbx_ctx_create(device, ctypes.byref(self.h)))
        self._modules = {}

    @classmethod
    def default(cls, device=0):
        key = (threading.get_ident(), device)
        if key not in cls._default:
            cls._default[key] = cls(device)
        return cls._default[key]

    def info(self):
        arch = ctypes.create_string_buffer(64)
        ncu, clk = ctypes.c_int(), ctypes.c_int()
        check(lib().tvbx_ctx_device_info(self.h, arch, 64, ctypes.byref(ncu),
                                         ctypes.byref(clk)))
        return {'arch': arch.value.decode(), 'n_cu': ncu.value,
                'clock_khz': clk.value}

    def sync(self):
        check(lib().tvbx_ctx_sync(self.h))

    def module(self, spec: 'emit.KernelSpec'):
        """Module for a kernel spec: a cached code object with the spec's
        source hash if one exists, else hipRTC compilation of the source."""
        key = spec.key()
        if key in self._modules:
            return self._modules[key]
        path = os.path.join(KDIR, key + '.hsaco')
        h = ctypes.c_void_p()
        if os.path.exists(path):
            with open(path, 'rb') as fd:
                img = fd.read()
            buf = ctypes.create_string_buffer(img, len(img))
            check(lib().tvbx_module_load(self.h, buf, len(img),
                                         ctypes.byref(h)))
            LOG.debug('loaded code object %s', path)
        else:
            opts = ' '.join(['--offload-arch=' + emit.ARCH] +
                            emit.DEVICE_FLAGS)
            log = ctypes.create_string_buffer(65536)
            check(lib().tvbx_module_compile(self.h, spec.source().encode(),
                                            opts.encode(), ctypes.byref(h),
                                            log, 65536))
            LOG.debug('hipRTC compiled %s', key)
            try:
                size = ctypes.c_size_t()
                check(lib().tvbx_module_image(h, None, 0, ctypes.byref(size)))
                img = ctypes.create_string_buffer(size.value)
                check(lib().tvbx_module_image(h, img, size.value, None))
                os.makedirs(KDIR, exist_ok=True)
                tmp = path + '.%d.tmp' % os.getpid()
                with open(tmp, 'wb') as fd:
                    fd.write(img.raw)
                os.replace(tmp, path)
            except OSError:
                pass
        self._modules[key] = h
        return h

    def rng_normals(self, n, seed=0):
        out = np.zeros(int(n), np.float32)
        check(lib().tvbx_rng_normals(self.h, seed, _ptr(out), out.size))
        return out

    def corr(self, samples):
        "samples (n, N, P) -> corr (P, N, N), FP64 accumulation on device"
        s = np.ascontiguousarray(samples, np.float32)
        n, N, P = s.shape
        out = np.zeros((P, N, N), np.float32)
        check(lib().tvbx_corr(self.h, _ptr(s), n, N, P, _ptr(out)))
        return out

    def corr_noalloc(self, samples_ptr, n, N, P, corr_ptr):
        """tvbx_corr on device memory with no allocation: samples
        (n, N, P) and corr (P, N, N) float32 at the device addresses given
        (e.g. torch tensors' data_ptr()); queued on this context's stream
        (``sync()`` waits for it)."""
        check(lib().tvbx_corr_noalloc(self.h, ctypes.c_void_p(samples_ptr),
                                      int(n), int(N), int(P),
                                      ctypes.c_void_p(corr_ptr)))

    def fc_loss(self, samples_ptr, n, N, P, target_ptr, per_instance,
                loss_ptr, s_ptr, mean_ptr):
        """tvbx_fc_loss on device memory (addresses, e.g. torch tensors'
        data_ptr()): loss (P,), the loss's symmetric derivative matrix s
        (N, N, P) and the means (N, P) of samples (n, N, P) against target
        (N, N), or (N, N, P) when ``per_instance``; queued on this context's
        stream."""
        check(lib().tvbx_fc_loss(self.h, ctypes.c_void_p(samples_ptr),
                                 int(n), int(N), int(P),
                                 ctypes.c_void_p(target_ptr),
                                 int(bool(per_instance)),
                                 ctypes.c_void_p(loss_ptr),
                                 ctypes.c_void_p(s_ptr),
                                 ctypes.c_void_p(mean_ptr)))

    def stream_rates(self, mib=1024):
        """(read GB/s, copy GB/s): the best of a grid-stride read and a
        grid-stride copy of ``mib`` MiB on this device (the copy counts its
        reads and writes)."""
        r, c = ctypes.c_double(), ctypes.c_double()
        check(lib().tvbx_stream_rates(self.h, int(mib) << 20, ctypes.byref(r),
                                      ctypes.byref(c)))
        return r.value, c.value

    def transpose(self, in_ptr, rows, cols, out_ptr, sub_ptr=None,
                  sub_rows=1):
        """out (cols, rows) = in (rows, cols) transposed, minus
        sub[r % sub_rows] (cols,) per row r when sub_ptr is given; float32
        device memory, queued on this context's stream."""
        check(lib().tvbx_transpose(self.h, ctypes.c_void_p(in_ptr),
                                   int(rows), int(cols),
                                   ctypes.c_void_p(sub_ptr or 0),
                                   int(sub_rows), ctypes.c_void_p(out_ptr)))


class RegionMap:
    """Surface vertices <-> regions on the device (the reference's RMapToAvg /
    RMapFromAvg, network.py:174-202).  ``rmap[k]`` is vertex k's region in
    [0, nroi).  Arrays are (rows, nvar) or (rows, nvar, n_inst): numpy arrays
    (staged through the device) or torch CUDA float32 tensors (used in
    place)."""

    def __init__(self, rmap, nroi=None, context=None):
        rm = np.ascontiguousarray(np.asarray(rmap).reshape(-1), np.int32)
        if nroi is None:
            nroi = int(rm.max()) + 1 if rm.size else 0
        if rm.size and (rm.min() < 0 or rm.max() >= nroi):
            raise ValueError('rmap entries must lie in [0, %d)' % nroi)
        self.nnode, self.nroi = int(rm.size), int(nroi)
        self.ctx = context or Context.default()
        self.h = ctypes.c_void_p()
        check(lib().tvbx_rmap_create(self.ctx.h, _ptr(rm), self.nnode,
                                     self.nroi, ctypes.byref(self.h)))

    def __del__(self):
        h = getattr(self, 'h', None)
        if h and _lib is not None:
            _lib.tvbx_rmap_destroy(h)
            self.h = None

    def _run(self, fn, rows_in, rows_out, src, dst):
        torch = sys.modules.get('torch')
        is_t = torch is not None and isinstance(src, torch.Tensor)
        shape = tuple(src.shape)
        if len(shape) not in (2, 3) or shape[0] != rows_in:
            raise ValueError('expected an array of shape (%d, nvar[, n_inst])'
                             % rows_in)
        nvar = shape[1]
        n_inst = shape[2] if len(shape) == 3 else 1
        oshape = (rows_out, ) + shape[1:]
        if is_t:
            if not src.is_cuda or src.dtype != torch.float32:
                raise ValueError('torch inputs must be float32 CUDA tensors')
            src = src.contiguous()
            if dst is None:
                dst = torch.empty(oshape, dtype=torch.float32,
                                  device=src.device)
            if tuple(dst.shape) != oshape or not dst.is_contiguous():
                raise ValueError('output must be contiguous, shape %s'
                                 % (oshape, ))
            torch.cuda.current_stream(src.device).synchronize()
            check(fn(self.h, nvar, n_inst, ctypes.c_void_p(src.data_ptr()),
                     ctypes.c_void_p(dst.data_ptr())))
            self.ctx.sync()
            return dst
        s = np.ascontiguousarray(src, np.float32)
        out = np.empty(oshape, np.float32) if dst is None else dst
        if out.shape != oshape or out.dtype != np.float32 or \
                not out.flags.c_contiguous:
            raise ValueError('output must be C-contiguous float32, shape %s'
                             % (oshape, ))
        check(fn(self.h, nvar, n_inst, _ptr(s), _ptr(out)))
        return out

    def to_avg(self, node, roi=None):
        "Region sums roi[r] = sum over rmap[k] == r of node[k]."
        return self._run(lib().tvbx_rmap_to_avg, self.nnode, self.nroi,
                         node, roi)

    def from_avg(self, roi, node=None):
        "node[k] = roi[rmap[k]]."
        return self._run(lib().tvbx_rmap_from_avg, self.nroi, self.nnode,
                         roi, node)


# ---------------------------------------------------------------------------

def _resolve(spec, name):
    """Map a user parameter name to a spec key ('m:x', 'c:x' or 'sigma')."""
    if name == 'sigma' or name.startswith(('m:', 'c:')):
        return name
    m, c = 'm:' + name, 'c:' + name
    inm, inc = m in spec.kinds, c in spec.kinds
    if inm and inc:
        raise ValueError('%r is both a model and a coupling parameter; use '
                         '"m:%s" or "c:%s"' % (name, name, name))
    if inm:
        return m
    if inc:
        return c
    raise KeyError('unknown parameter %r' % name)


def kind_of_value(value, N, P):
    v = np.asarray(value)
    if v.ndim == 0:
        return 'glob'
    if v.shape == (P, ) and P != N:
        return 'inst'
    if v.shape == (N, ) and P != N:
        return 'node'
    if v.shape == (N, P):
        return 'ni'
    raise ValueError('cannot infer a parameter kind from shape %s (N=%d, '
                     'P=%d); give kinds explicitly' % (v.shape, N, P))


class Simulation:
    """P sweep instances of one (model, coupling, scheme) on the device.

    csr: ``network.CSR`` with one delay row per group; instance p uses group
        p // group_size (group_size a multiple of 64, default: all in one).
    params: {name: value} -- scalar, (P,), (N,) or (N, P) arrays; the kind is
        inferred unless ``kinds`` says otherwise.  Unlisted constants keep
        their model / coupling defaults as literals.
    monitor: tavg_period (steps), bold (bool), fc (bool), sample_slots.
    launch: ipl, waves, node_block, steps_per_launch, unroll, min_waves,
        pipe, prio, xcd_map, emaj, pace (see tvb_hpc.autotune); adj_unroll:
        gathers in flight per group of the adjoint's transposed rows (0: the
        forward's unroll; 32 allows at most 8 waves); coupling 'gather' (CSR
        rows) or 'mfma' (delay-binned sums on the matrix cores; bfloat16 /
        float16 rings holding pre(y), blocked launches of at most 16 steps).
    """

    def __init__(self, model, cfun, scheme, csr: CSR, n_inst, horizon=None,
                 params=None, kinds=None, group_size=None, ring='coupled',
                 ring_pre=None, tavg_period=0, bold=False, bold_obs=0,
                 fc=False, sample_slots=64, waves=4, node_block=None,
                 steps_per_launch=1, compat_outputs=False, dt=None,
                 order='cuda', ipl=1, ring_dtype='f32', ring_scale=1.0,
                 adjoint=False, unroll=8, min_waves=1, pipe=False,
                 prio=True, xcd_map=False, coupling='gather', emaj=0,
                 pace=False, adj_unroll=0, context=None):
        self.model, self.cfun, self.scheme = model, cfun, scheme
        self.csr = csr
        N, P = csr.nnode, int(n_inst)
        self.N, self.P = N, P
        params = dict(params or {})
        kinds = dict(kinds or {})
        probe = emit.KernelSpec(model, cfun, scheme, ring=ring,
                                ring_pre=ring_pre,
                                sigma_kind=None)
        resolved = {}
        for name, val in params.items():
            key = _resolve(probe, name) if name != 'sigma' else 'sigma'
            resolved[key] = val
            if key not in kinds:
                kinds[key] = kind_of_value(val, N, P)
        sigma_kind = kinds.pop('sigma', None)
        if sigma_kind is None and getattr(scheme, 'sigma', None) is not None:
            sigma_kind = 'const'
        self.spec = emit.KernelSpec(model, cfun, scheme, kinds=kinds,
                                    sigma_kind=sigma_kind, ring=ring,
                                    ring_pre=ring_pre, order=order, ipl=ipl,
                                    ring_dtype=ring_dtype, adjoint=adjoint,
                                    unroll=unroll, min_waves=min_waves,
                                    pipe=pipe, prio=prio, emaj=emaj,
                                    adj_unroll=adj_unroll)
        if ring_scale != 1.0 and not self.spec.ring_pre:
            raise ValueError('ring_scale needs a ring that holds pre(y)')
        if ring_dtype == 'u8':
            lo, hi = cfun.pre_bounds()
            if lo < 0 or hi * ring_scale > 1.0 + 1e-6:
                raise ValueError('an 8-bit history holds values in [0, 1]: '
                                 'ring_scale * pre must stay within it '
                                 '(pre in [%g, %g], ring_scale %g)'
                                 % (lo, hi, ring_scale))
        self.values = resolved
        self.ctx = context or Context.default()
        self.module = self.ctx.module(self.spec)
        nb, spl = int(node_block or N), int(steps_per_launch)
        # temporal blocking: several steps per launch with the node axis
        # split over workgroups (one node per wave), legal while every delay
        # d >= 1 satisfies steps <= d + 1
        self.blocked = spl > 1 and nb < N
        if self.blocked:
            if order == 'reference':
                raise ValueError('temporal blocking needs the CUDA ordering')
            if nb != int(waves) or nb > 8:
                raise ValueError('temporal blocking takes one node per wave: '
                                 'node_block (%d) must equal waves (%d), at '
                                 'most 8' % (nb, waves))
            dmin = int(np.min(csr.delays)) if csr.nnz else 0
            if csr.nnz and (dmin < 1 or spl > dmin + 1):
                raise ValueError('%d steps per launch need every delay >= %d '
                                 '(minimum delay %d)' % (spl, spl - 1, dmin))
        H_min = horizon_for(csr.max_delay, spl if self.blocked else 1)
        H = int(horizon or H_min)
        if H < H_min:
            raise ValueError('horizon %d too short for max delay %d' %
                             (H, csr.max_delay))
        self.H = H
        tile = 64 * int(ipl)
        pitch = (P + tile - 1) // tile * tile
        self.pitch = pitch
        if group_size is None:
            if csr.n_group != 1:
                raise ValueError('several delay groups need group_size')
            group_size = pitch
        if group_size % tile:
            raise ValueError('group_size must be a multiple of 64 * ipl')
        if csr.n_group * group_size < pitch:
            raise ValueError('delay groups do not cover all instances')
        counts = self.spec.counts()
        dt_v = scheme.dt_value(dt)
        self.dt = dt_v
        d = SimDesc()
        d.n_node, d.n_inst, d.n_svar = N, P, self.spec.S
        d.n_ring, d.n_coupled, d.horizon = self.spec.R, self.spec.C, H
        d.nnz, d.n_group, d.group_size = csr.nnz, csr.n_group, group_size
        d.n_pinst, d.n_pnode = counts['inst'], counts['node']
        d.n_pni, d.n_pglob = counts['ni'], counts['glob']
        d.dt = dt_v
        d.scheme = 1 if self.spec.em else 0
        d.tavg_period = int(tavg_period)
        d.bold, d.bold_obs, d.fc = int(bool(bold)), int(bold_obs), int(bool(fc))
        d.sample_slots = int(sample_slots)
        d.waves = int(waves)
        d.node_block = nb
        d.steps_per_launch = spl
        d.compat_outputs = int(bool(compat_outputs))
        d.order = 1 if order == 'reference' else 0
        d.ipl = int(ipl)
        d.ring_type = emit.RING_TYPES[ring_dtype]
        d.ring_scale = float(ring_scale)
        d.xcd_map = int(bool(xcd_map))
        d.weight_groups = int(csr.weight_groups)
        self.desc = d
        self.h = ctypes.c_void_p()
        check(lib().tvbx_sim_create(self.ctx.h, self.module, ctypes.byref(d),
                                    ctypes.byref(self.h)))
        check(lib().tvbx_sim_set_connectome(
            self.h, _ptr(csr.row), _ptr(csr.col), _ptr(csr.weight),
            _ptr(csr.delays)))
        self.set_params(resolved)
        self.coupling = None
        self.set_coupling(coupling)
        if pace:
            check(lib().tvbx_sim_set_pacing(self.h, int(pace)))

    def set_coupling(self, mode):
        """'gather': CSR row sums; 'mfma': delay-binned sums on the matrix
        cores (tvbx_sim_set_coupling)."""
        code = {'gather': 0, 'mfma': 1}[mode]
        check(lib().tvbx_sim_set_coupling(self.h, code))
        self.coupling = mode

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                lib().tvbx_sim_destroy(h)
            except Exception:
                pass
            self.h = None

    # -- parameters ---------------------------------------------------------

    def _default(self, key):
        if key == 'sigma':
            return self.scheme.sigma
        scope, name = key.split(':', 1)
        return (self.model.const if scope == 'm' else self.cfun.param)[name]

    def set_params(self, values):
        """Upload run-time parameters; unspecified ones take defaults."""
        N, P = self.N, self.P
        for k, v in values.items():
            key = _resolve(self.spec, k) if k != 'sigma' else 'sigma'
            self.values[key] = v
        counts = self.spec.counts()
        arrs = {'glob': np.zeros(counts['glob'], np.float32),
                'inst': np.zeros((counts['inst'], P), np.float32),
                'node': np.zeros((counts['node'], N), np.float32),
                'ni': np.zeros((counts['ni'], N, P), np.float32)}
        for key, kind, idx in self.spec.param_layout():
            v = self.values.get(key, None)
            if v is None:
                v = self._default(key)
            if isinstance(v, E.Expr) or v is None:
                raise ValueError('parameter %s needs a value' % key)
            v = np.asarray(v, np.float64)
            if kind == 'glob':
                arrs['glob'][idx] = v
            elif kind == 'inst':
                arrs['inst'][idx] = np.broadcast_to(v, (P, ))
            elif kind == 'node':
                arrs['node'][idx] = np.broadcast_to(v, (N, ))
            else:
                if v.shape == (P, ):
                    v = v[None, :]
                elif v.shape == (N, ):
                    v = v[:, None]
                arrs['ni'][idx] = np.broadcast_to(v, (N, P))
        for kind, code in (('glob', 0), ('inst', 1), ('node', 2), ('ni', 3)):
            a = np.ascontiguousarray(arrs[kind])
            check(lib().tvbx_sim_set_params(self.h, code, _ptr(a), a.size))

    # -- state ------------------------------------------------------------------

    def set_state(self, x):
        "x: (S, N, P)"
        x = np.ascontiguousarray(x, np.float32)
        if x.shape != (self.spec.S, self.N, self.P):
            raise ValueError('state must be (S, N, P) = %s' %
                             ((self.spec.S, self.N, self.P), ))
        check(lib().tvbx_sim_set_state(self.h, _ptr(x)))

    def get_state(self):
        x = np.zeros((self.spec.S, self.N, self.P), np.float32)
        check(lib().tvbx_sim_get_state(self.h, _ptr(x)))
        return x

    def set_ring(self, r):
        r = np.ascontiguousarray(r, np.float32)
        if r.shape != (self.H, self.spec.R, self.N, self.P):
            raise ValueError('ring shape')
        check(lib().tvbx_sim_set_ring(self.h, _ptr(r)))

    def get_ring(self):
        r = np.zeros((self.H, self.spec.R, self.N, self.P), np.float32)
        check(lib().tvbx_sim_get_ring(self.h, _ptr(r)))
        return r

    @property
    def step(self):
        s = ctypes.c_longlong()
        check(lib().tvbx_sim_get_step(self.h, ctypes.byref(s)))
        return s.value

    @step.setter
    def step(self, n):
        check(lib().tvbx_sim_set_step(self.h, int(n)))

    def init_history(self, mode='zeros'):
        m = {'zeros': 0, 'constant': 1, 'current': 2}[mode]
        check(lib().tvbx_sim_init_history(self.h, m))

    def set_noise(self, seed=None):
        "Philox noise with the given seed; None switches noise off."
        if seed is None:
            check(lib().tvbx_sim_set_noise(self.h, 0, 0))
        else:
            check(lib().tvbx_sim_set_noise(self.h, 1, int(seed)))

    def set_exec(self, graph=True, window=100, profile=False):
        """graph: replay hipGraph windows of `window` steps; profile: time
        every launch (or window) with events for ``timing()``."""
        check(lib().tvbx_sim_set_exec(self.h, int(bool(graph)), int(window),
                                      int(bool(profile))))

    def run(self, n_steps, noise=None):
        """Advance n_steps.  noise: injected normals (n_steps, S, N, P) or
        (S, N, P) reused every step."""
        if noise is not None:
            nz = np.ascontiguousarray(noise, np.float32)
            ts = 0 if nz.ndim == 3 else self.spec.S * self.N * self.P
            check(lib().tvbx_sim_run(self.h, int(n_steps), _ptr(nz), ts))
        else:
            check(lib().tvbx_sim_run(self.h, int(n_steps), None, 0))

    def sync(self):
        self.ctx.sync()

    # -- gradients ------------------------------------------------------------

    def checkpoint_run(self, n_steps, window):
        """Advance n_steps (a multiple of window) keeping checkpoints for
        ``adjoint`` (requires ``adjoint=True``).  window = 0 records every
        step's state and coupling sums instead ((S + 1) * N * P * 4 bytes
        per step of device memory): the adjoint then needs no replay."""
        check(lib().tvbx_sim_checkpoint_run(self.h, int(n_steps),
                                            int(window)))

    def set_checkpoint_reuse(self, keep):
        """keep=True (the default): ``adjoint`` restores each window's
        checkpoint by a copy and leaves them for another adjoint of the same
        run.  keep=False: it replays each window inside its checkpoint's
        buffers (no copies) and consumes them, so a second ``adjoint`` needs
        a new ``checkpoint_run``.  The gradients are the same."""
        check(lib().tvbx_sim_set_checkpoint_reuse(self.h, int(bool(keep))))

    def adjoint(self, dsamples, sample0=0, n_samples=None):
        """Gradients of a loss from dL/dsample of the samples produced by
        the last ``checkpoint_run``: a (n, N, P) float32 numpy array, or the
        device address of one (an int, e.g. a torch tensor's data_ptr(),
        with ``n_samples``; the caller has finished writing it).  Returns
        {param key: gradient}: coupling parameters (P,), model parameters
        of kind 'ni' (N, P) and 'inst' (P,)."""
        gc, gm = self.spec.grad_params()
        if isinstance(dsamples, int):
            if n_samples is None:
                raise ValueError('a device address needs n_samples')
            ptr, on_dev, n, keep = dsamples, 1, int(n_samples), None
        else:
            keep = np.ascontiguousarray(dsamples, np.float32)
            if keep.shape[1:] != (self.N, self.P):
                raise ValueError('dsamples must be (n, N, P)')
            ptr, on_dev, n = _ptr(keep), 0, keep.shape[0]
        g_c = np.zeros((max(len(gc), 1), self.P), np.float32)
        g_m = np.zeros((max(len(gm), 1), self.N, self.P), np.float32)
        check(lib().tvbx_sim_adjoint(self.h, ctypes.c_void_p(ptr), on_dev,
                                     int(sample0), n, len(gc), _ptr(g_c),
                                     len(gm), _ptr(g_m), 0))
        out = {}
        for k, key in enumerate(gc):
            out[key] = g_c[k]
        for k, key in enumerate(gm):
            out[key] = g_m[k] if self.spec.kinds[key] == 'ni' else \
                g_m[k].sum(axis=0)
        return out

    def outputs(self):
        C, S, N, P = self.spec.C, self.spec.S, self.N, self.P
        i = np.zeros((C, N, P), np.float32)
        dr = np.zeros((S, N, P), np.float32)
        df = np.zeros((S, N, P), np.float32)
        check(lib().tvbx_sim_get_outputs(self.h, _ptr(i), _ptr(dr), _ptr(df)))
        return i, dr, df

    def state_into(self, ptr):
        """Copy the state (S, N, P) float32 to device (or host) memory at
        address ``ptr`` (e.g. a torch tensor's data_ptr())."""
        check(lib().tvbx_sim_get_state(self.h, ctypes.c_void_p(ptr)))

    def samples_into(self, ptr):
        """Copy the pending monitor samples to device (or host) memory at
        address ``ptr`` ((n, N, P) float32, e.g. a torch tensor's
        data_ptr()); returns (n, index of the first sample)."""
        cnt, first = ctypes.c_int(), ctypes.c_longlong()
        check(lib().tvbx_sim_get_samples(self.h, ctypes.c_void_p(ptr),
                                         ctypes.byref(first),
                                         ctypes.byref(cnt)))
        return cnt.value, first.value

    def pending_samples(self):
        "(number of pending monitor samples, index of the first)"
        cnt, first = ctypes.c_int(), ctypes.c_longlong()
        check(lib().tvbx_sim_get_samples(self.h, None, ctypes.byref(first),
                                         ctypes.byref(cnt)))
        return cnt.value, first.value

    def samples(self):
        "pending monitor samples (n, N, P) and the index of the first"
        cnt, first = ctypes.c_int(), ctypes.c_longlong()
        check(lib().tvbx_sim_get_samples(self.h, None, ctypes.byref(first),
                                         ctypes.byref(cnt)))
        out = np.zeros((cnt.value, self.N, self.P), np.float32)
        check(lib().tvbx_sim_get_samples(self.h, _ptr(out),
                                         ctypes.byref(first),
                                         ctypes.byref(cnt)))
        return out, first.value

    def fc(self, p0=0, np_=None):
        "correlation matrices (np, N, N) of instances [p0, p0 + np)"
        np_ = self.P - p0 if np_ is None else np_
        out = np.zeros((np_, self.N, self.N), np.float32)
        ns = ctypes.c_longlong()
        check(lib().tvbx_sim_fc(self.h, p0, np_, _ptr(out), ctypes.byref(ns)))
        self.n_fc_samples = ns.value
        return out

    def fc_reset(self):
        check(lib().tvbx_sim_fc_reset(self.h))

    def logp(self, traj, window=0, grad=False):
        """Euler-Maruyama log-likelihood per instance (FP64) of a trajectory
        traj (n + 1, S, N, P): the transitions x(k) -> x(k + 1) with the
        delayed coupling of the trajectory itself (teacher forcing; the
        history before traj[0] is the ring as it stands).  Leaves the
        state at traj[-1] and the step advanced by n (scheme.py:42-48).
        With ``grad`` (a simulation created with adjoint=True) returns
        (logp, {param key: d logp / d param}) for the parameters the
        adjoint differentiates: coupling (P,), model 'ni' (N, P) and
        'inst' (P,)."""
        torch = sys.modules.get('torch')
        if torch is not None and isinstance(traj, torch.Tensor):
            if not traj.is_cuda or traj.dtype != torch.float32 or \
                    not traj.is_contiguous():
                raise ValueError('a torch trajectory must be a contiguous '
                                 'float32 CUDA tensor')
            torch.cuda.current_stream(traj.device).synchronize()
            tr, tptr = traj, ctypes.c_void_p(traj.data_ptr())
        else:
            tr = np.ascontiguousarray(traj, np.float32)
            tptr = _ptr(tr)
        if tr.ndim != 4 or tuple(tr.shape[1:]) != (self.spec.S, self.N,
                                                   self.P) \
                or tr.shape[0] < 2:
            raise ValueError('traj must be (n + 1, %d, %d, %d), n >= 1'
                             % (self.spec.S, self.N, self.P))
        out = np.zeros(self.P, np.float64)
        gc, gm = self.spec.grad_params() if grad else ([], [])
        if grad and not (gc or gm):
            raise ValueError('no parameter of kind inst / ni to '
                             'differentiate')
        g_c = np.zeros((max(len(gc), 1), self.P), np.float32)
        g_m = np.zeros((max(len(gm), 1), self.N, self.P), np.float32)
        check(lib().tvbx_sim_logp(self.h, tptr, tr.shape[0] - 1,
                                  int(window), _ptr(out), len(gc), _ptr(g_c),
                                  len(gm), _ptr(g_m)))
        if not grad:
            return out
        g = {}
        for k, key in enumerate(gc):
            g[key] = g_c[k]
        for k, key in enumerate(gm):
            g[key] = g_m[k] if self.spec.kinds[key] == 'ni' else \
                g_m[k].sum(axis=0)
        return out, g

    def fc_sc_error(self, sc, p0=0, np_=None):
        """sum over node pairs i < j of (FC_ij - SC_ij)^2 for instances
        [p0, p0 + np) (FP64, computed on the device from the FC accumulated
        since the last ``fc_reset``); sc (N, N)"""
        np_ = self.P - p0 if np_ is None else np_
        sc = np.ascontiguousarray(sc, np.float32)
        if sc.shape != (self.N, self.N):
            raise ValueError('sc must be (%d, %d)' % (self.N, self.N))
        out = np.zeros(np_, np.float64)
        ns = ctypes.c_longlong()
        check(lib().tvbx_sim_fc_sc_error(self.h, _ptr(sc), p0, np_, _ptr(out),
                                         ctypes.byref(ns)))
        self.n_fc_samples = ns.value
        return out

    def flags(self):
        f = np.zeros(self.P, np.uint32)
        check(lib().tvbx_sim_get_flags(self.h, _ptr(f)))
        return f

    def timing(self, reset=True):
        "(ms of step kernels, number of step launches) since the last reset"
        ms, n = ctypes.c_double(), ctypes.c_longlong()
        check(lib().tvbx_sim_timing(self.h, ctypes.byref(ms), ctypes.byref(n),
                                    int(reset)))
        return ms.value, n.value


# ---------------------------------------------------------------------------
# reference-compatible executor

class Batch:
    """Subjects with different connectomes, integrated in one launch.

    subjects: [(csr, n_inst), ...] -- each ``network.CSR`` with one delay
        row; subject k's instances are numbered after those of subjects
        0 .. k-1 (P_total instances in all).
    params: {name: value}; instance parameters as (P_total,) arrays, node
        parameters as a list of (N_k,) arrays, scalars for globals.
    launch: ipl, waves (= nodes per workgroup), steps_per_launch, min_waves,
        unroll, pipe, prio, queue (persistent work-queue workgroups; see
        tvb_hpc.autotune).
    Philox noise only (``set_noise``); no monitor.
    """

    def __init__(self, model, cfun, scheme, subjects, params=None,
                 kinds=None, ring_dtype='f32', ring_scale=1.0, waves=1,
                 steps_per_launch=1, ipl=1, min_waves=1, unroll=8, pipe=False,
                 prio=True, queue=False, emaj=0, dt=None,
                 context=None):
        self.model, self.cfun, self.scheme = model, cfun, scheme
        self.subjects = [(csr, int(p)) for csr, p in subjects]
        self.P_total = sum(p for _, p in self.subjects)
        params = dict(params or {})
        kinds = dict(kinds or {})
        probe = emit.KernelSpec(model, cfun, scheme, sigma_kind=None)
        resolved = {}
        for name, val in params.items():
            key = _resolve(probe, name) if name != 'sigma' else 'sigma'
            resolved[key] = val
            if key not in kinds:
                a = np.asarray(val)
                kinds[key] = 'inst' if a.shape == (self.P_total, ) else \
                    'const' if a.ndim == 0 else 'node'
        sigma_kind = kinds.pop('sigma', None)
        if sigma_kind is None and getattr(scheme, 'sigma', None) is not None:
            sigma_kind = 'const'
        self.spec = emit.KernelSpec(model, cfun, scheme, kinds=kinds,
                                    sigma_kind=sigma_kind, ipl=ipl,
                                    ring_dtype=ring_dtype, unroll=unroll,
                                    min_waves=min_waves, pipe=pipe,
                                    prio=prio, emaj=emaj)
        if ring_dtype == 'u8':
            lo, hi = cfun.pre_bounds()
            if lo < 0 or hi * ring_scale > 1.0 + 1e-6:
                raise ValueError('an 8-bit history holds values in [0, 1]')
        self.values = resolved
        self.ctx = context or Context.default()
        self.module = self.ctx.module(self.spec)
        L = int(steps_per_launch)
        sd = (SubjectDesc * len(self.subjects))()
        for k, (csr, p) in enumerate(self.subjects):
            if csr.n_group != 1:
                raise ValueError('a subject has one delay row')
            if L > 1 and csr.nnz and (int(np.min(csr.delays)) < max(L - 1, 1)):
                raise ValueError('subject %d: %d steps per launch need every '
                                 'delay >= %d' % (k, L, L - 1))
            sd[k].n_node, sd[k].n_inst, sd[k].nnz = csr.nnode, p, csr.nnz
            sd[k].horizon = horizon_for(csr.max_delay, L)
        self.H = [sd[k].horizon for k in range(len(self.subjects))]
        counts = self.spec.counts()
        d = BatchDesc()
        d.n_subject = len(self.subjects)
        d.n_svar, d.n_ring, d.n_coupled = self.spec.S, self.spec.R, self.spec.C
        d.n_pinst, d.n_pnode = counts['inst'], counts['node']
        d.n_pni, d.n_pglob = counts['ni'], counts['glob']
        d.dt = scheme.dt_value(dt)
        d.scheme = 1 if self.spec.em else 0
        d.waves = d.node_block = int(waves)
        d.steps_per_launch = L
        d.ipl = int(ipl)
        d.ring_type = emit.RING_TYPES[ring_dtype]
        d.ring_scale = float(ring_scale)
        self.desc = d
        self.h = ctypes.c_void_p()
        check(lib().tvbx_batch_create(self.ctx.h, self.module, ctypes.byref(d),
                                      sd, ctypes.byref(self.h)))
        for k, (csr, _) in enumerate(self.subjects):
            check(lib().tvbx_batch_set_connectome(
                self.h, k, _ptr(csr.row), _ptr(csr.col), _ptr(csr.weight),
                _ptr(np.ascontiguousarray(csr.delays[0], np.int32))))
        self.set_params(resolved)
        self.queue = bool(queue)
        self.set_exec(graph=False, queue=self.queue)

    def __del__(self):
        h = getattr(self, 'h', None)
        if h is not None and h.value:
            try:
                lib().tvbx_batch_destroy(h)
            except Exception:
                pass
            self.h = None

    def _default(self, key):
        if key == 'sigma':
            return self.scheme.sigma
        scope, name = key.split(':', 1)
        return (self.model.const if scope == 'm' else self.cfun.param)[name]

    def set_params(self, values):
        for k, v in values.items():
            key = _resolve(self.spec, k) if k != 'sigma' else 'sigma'
            self.values[key] = v
        counts = self.spec.counts()
        Ns = [csr.nnode for csr, _ in self.subjects]
        arrs = {'glob': np.zeros(counts['glob'], np.float32),
                'inst': np.zeros((counts['inst'], self.P_total), np.float32),
                'node': np.zeros((counts['node'], sum(Ns)), np.float32)}
        if counts['ni']:
            raise ValueError('batches take no node-and-instance parameters')
        for key, kind, idx in self.spec.param_layout():
            v = self.values.get(key, None)
            if v is None:
                v = self._default(key)
            if isinstance(v, E.Expr) or v is None:
                raise ValueError('parameter %s needs a value' % key)
            if kind == 'glob':
                arrs['glob'][idx] = v
            elif kind == 'inst':
                arrs['inst'][idx] = np.broadcast_to(np.asarray(v, np.float64),
                                                    (self.P_total, ))
            else:
                if isinstance(v, (list, tuple)):
                    v = np.concatenate([np.broadcast_to(np.asarray(
                        x, np.float64), (n, )) for x, n in zip(v, Ns)])
                arrs['node'][idx] = np.broadcast_to(np.asarray(v), (sum(Ns), ))
        for kind, code in (('glob', 0), ('inst', 1), ('node', 2)):
            a = np.ascontiguousarray(arrs[kind])
            check(lib().tvbx_batch_set_params(self.h, code, _ptr(a), a.size))

    def set_state(self, k, x):
        "x: (S, N_k, P_k) of subject k"
        csr, p = self.subjects[k]
        x = np.ascontiguousarray(x, np.float32)
        if x.shape != (self.spec.S, csr.nnode, p):
            raise ValueError('state of subject %d must be %s' %
                             (k, (self.spec.S, csr.nnode, p)))
        check(lib().tvbx_batch_set_state(self.h, k, _ptr(x)))

    def get_state(self, k):
        csr, p = self.subjects[k]
        x = np.zeros((self.spec.S, csr.nnode, p), np.float32)
        check(lib().tvbx_batch_get_state(self.h, k, _ptr(x)))
        return x

    @property
    def step(self):
        v = ctypes.c_longlong()
        check(lib().tvbx_batch_get_step(self.h, ctypes.byref(v)))
        return v.value

    @step.setter
    def step(self, n):
        check(lib().tvbx_batch_set_step(self.h, int(n)))

    def init_history(self, mode='zeros'):
        check(lib().tvbx_batch_init_history(
            self.h, {'zeros': 0, 'constant': 1, 'current': 2}[mode]))

    def set_noise(self, seed=None):
        if seed is None:
            check(lib().tvbx_batch_set_noise(self.h, 0, 0))
        else:
            check(lib().tvbx_batch_set_noise(self.h, 1, int(seed)))

    def set_exec(self, graph=True, window=100, queue=None):
        """graph: replay captured windows of `window` steps; queue:
        persistent work-queue workgroups (default: as constructed)."""
        if queue is not None:
            self.queue = bool(queue)
        check(lib().tvbx_batch_set_exec(self.h, int(self.queue),
                                        int(bool(graph)), int(window)))

    def run(self, n_steps):
        check(lib().tvbx_batch_run(self.h, int(n_steps)))

    def sync(self):
        check(lib().tvbx_batch_sync(self.h))

    def flags(self):
        f = np.zeros(self.P_total, np.uint32)
        check(lib().tvbx_batch_get_flags(self.h, _ptr(f)))
        return f


class NetworkStepExecutor:
    """Callable with the arguments of the reference's fused kernel
    (transforms.py:23-49; called as in examples/hackathon.py:49-53):

        knl(nstep, nnode, ntime, state, input, param, drift, diffs, obsrv,
            nnz, delays, row, col, weights, i_step_0, [cfun params], [dWt],
            [dt])

    Arrays are the reference's: state (nnode, nsvar), input (nnode, ninput),
    param (nnode, nparam), drift / diffs (nnode, nsvar), obsrv
    (ntime, nnode, nobs), CSR delays / row / col / weights; an extra trailing
    instance axis (P) is accepted on state / input / param / drift / diffs /
    obsrv.  They are updated in place.  ``dWt`` (EM only) is the injected
    normal increment: (nnode, nsvar[, P]) reused for every step, as in the
    reference, or (nstep, nnode, nsvar[, P]) one per step.

    Ordering within a step follows the reference's fused loopy kernel: step
    n (slot t = n % ntime) reads y(n - d) from slot (n - d) % ntime for
    d >= 1 and reads slot t itself (post_syn, zero delays) *before* writing
    obsrv[t] = h(state); after the call obsrv holds exactly what the
    reference's would.  (``order='cuda'`` selects the dsl_cuda kernels'
    order instead, in which slot t already holds h(x(n)).)
    """

    def __init__(self, target, model, cfun, scheme, order='reference'):
        self.target = target
        self.model, self.cfun, self.scheme = model, cfun, scheme
        self.order = order
        self._sims = {}

    def __call__(self, nstep, nnode, ntime, state, input, param, drift, diffs,
                 obsrv, nnz, delays, row, col, weights, i_step_0=0, dWt=None,
                 dt=None, **cfun_params):
        model, cfun = self.model, self.cfun
        S = len(model.state_sym)
        nobs = len(model.obsrv_sym)
        batched = np.ndim(state) == 3
        P = state.shape[2] if batched else 1
        N = int(nnode)
        if int(nnz) != len(col):
            raise ValueError('nnz disagrees with col')
        csr = CSR(N, np.asarray(row, np.int64), np.asarray(col, np.int64),
                  np.asarray(weights, np.float32),
                  np.asarray(delays, np.int64))
        # parameter kinds: model params from `param` (per node, or per node and
        # instance when batched); cfun params from keywords
        _, _, mparam = model.names
        kinds, values = {}, {}
        for k, name in enumerate(mparam):
            pv = np.asarray(param)[:, k]
            kinds['m:' + name] = 'ni' if batched else 'node'
            values['m:' + name] = pv
        for name, val in cfun_params.items():
            if name not in cfun.param:
                raise TypeError('unexpected argument %r' % name)
            kinds['c:' + name] = 'glob' if np.ndim(val) == 0 else 'inst'
            values['c:' + name] = val
        for name, val in cfun.param.items():
            if isinstance(val, E.Expr) and name not in cfun_params:
                raise TypeError('missing coupling argument %r' % name)
        key = (N, int(ntime), P, csr.nnz, tuple(sorted(kinds.items())),
               dt if dt is not None else None)
        sim = self._sims.get(key)
        if sim is None:
            sim = Simulation(model, cfun, self.scheme, csr, P,
                             horizon=int(ntime), params=values, kinds=kinds,
                             ring='all',
                             ring_pre=False, compat_outputs=True, dt=dt,
                             order=self.order,
                             context=self.target.context())
            self._sims[key] = sim
        else:
            check(lib().tvbx_sim_set_connectome(
                sim.h, _ptr(csr.row), _ptr(csr.col), _ptr(csr.weight),
                _ptr(csr.delays)))
        sim.set_params(values)
        x = np.asarray(state, np.float32)
        x = x[..., None] if not batched else x
        sim.set_state(np.transpose(x, (1, 0, 2)))
        ob = np.asarray(obsrv, np.float32)
        ob = ob[..., None] if not batched else ob
        sim.set_ring(np.transpose(ob, (0, 2, 1, 3)))
        sim.step = int(i_step_0)
        sim.init_history('current')
        noise = None
        if dWt is not None:
            if not self.scheme.stochastic:
                raise TypeError('dWt given to a deterministic scheme')
            w = np.asarray(dWt, np.float32)
            if not batched:
                w = w[..., None]
            noise = np.transpose(w, (1, 0, 2)) if w.ndim == 3 else \
                np.transpose(w, (0, 2, 1, 3))
        sim.run(int(nstep), noise=noise)
        xs = sim.get_state()
        r = sim.get_ring()
        inp, dr, df = sim.outputs()
        state[...] = _unbatch(np.transpose(xs, (1, 0, 2)), batched)
        obsrv[...] = _unbatch(np.transpose(r, (0, 2, 1, 3)), batched)
        if input is not None and np.size(input):
            input[:, :inp.shape[0]] = _unbatch(np.transpose(inp, (1, 0, 2)),
                                                 batched)
        if drift is not None and np.size(drift):
            drift[...] = _unbatch(np.transpose(dr, (1, 0, 2)), batched)
        if diffs is not None and np.size(diffs):
            diffs[...] = _unbatch(np.transpose(df, (1, 0, 2)), batched)


def _unbatch(a, batched):
    return a if batched else a[..., 0]


class HipTarget:
    """Target plugin for ``transforms.network_time_step`` (replaces the
    reference's NumbaTarget / OpenMPCTarget; numba.py:10-32,
    compiler.py:228-251)."""

    def __init__(self, device=0):
        self.device = device

    def context(self):
        return Context.default(self.device)

    def get_kernel_executor_cache_key(self, *args, **kwargs):
        return 'default'

    def get_kernel_executor(self, model, cfun, scheme, order='reference'):
        return NetworkStepExecutor(self, model, cfun, scheme, order=order)

"""The C-ABI library (CPU only: loading, exports, layouts, argument
checking that needs no device) and hipRTC acceptance of the emitted
sources."""

import ctypes
import os
import re

import numpy as np
import pytest

from tvb_hpc import configs, emit, hip

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, 'include', 'tvbx.h')


def _declared():
    src = open(HEADER).read()
    return sorted(set(re.findall(r'^int\s+(tvbx_\w+)\s*\(', src, re.M)))


def test_library_exports_every_declared_function():
    names = _declared()
    assert len(names) >= 30
    L = ctypes.CDLL(hip.LIB_PATH)
    for n in names:
        assert hasattr(L, n), n
    # and the Python binding covers each of them
    assert set(names) <= set(hip._SIGS)


def test_abi_version_and_errors():
    L = hip.lib()
    assert L.tvbx_abi_version() == 3
    assert L.tvbx_last_error(None, 0) == -1
    n = ctypes.c_int(-1)
    assert L.tvbx_device_count(ctypes.byref(n)) == 0
    assert n.value >= 0
    assert L.tvbx_sim_run(None, 1, None, 0) == -1
    buf = ctypes.create_string_buffer(256)
    L.tvbx_last_error(buf, 256)
    assert b'bad argument' in buf.value


def test_sim_desc_mirror_matches_header():
    src = open(HEADER).read()
    body = src[src.index('typedef struct tvbx_sim_desc'):]
    body = body[:body.index('} tvbx_sim_desc;')]
    fields = re.findall(r'^\s*(int|float)\s+([\w, ]+);', body, re.M)
    names = []
    for ty, decl in fields:
        names += [(ty, n.strip()) for n in decl.split(',')]
    mirror = [(('float' if t is ctypes.c_float else 'int'), n)
              for n, t in hip.SimDesc._fields_]
    assert mirror == names


@pytest.mark.parametrize('cname,mirror', [('tvbx_batch_desc', 'BatchDesc'),
                                           ('tvbx_subject_desc',
                                            'SubjectDesc')])
def test_batch_desc_mirrors_match_header(cname, mirror):
    src = open(HEADER).read()
    body = src[src.index('typedef struct %s {' % cname):]
    body = body[:body.index('} %s;' % cname)]
    fields = re.findall(r'^\s*(int|float)\s+([\w, ]+);', body, re.M)
    names = []
    for ty, decl in fields:
        names += [(ty, n.strip()) for n in decl.split(',')]
    got = [(('float' if t is ctypes.c_float else 'int'), n)
           for n, t in getattr(hip, mirror)._fields_]
    assert got == names


def test_args_header_is_plain_c():
    src = open(os.path.join(emit.CSRC, 'tvbx_args.h')).read()
    for bad in ('std::', 'torch', 'template'):
        assert bad not in src


def test_no_gpu_context_fails_cleanly():
    if hip.device_count() > 0:
        pytest.skip('a device is visible')
    h = ctypes.c_void_p()
    rc = hip.lib().tvbx_ctx_create(0, ctypes.byref(h))
    assert rc != 0
    with pytest.raises(hip.TvbxError):
        hip.check(rc)


def _hiprtc_compile(src):
    L = ctypes.CDLL('/opt/rocm/lib/libhiprtc.so')
    prog = ctypes.c_void_p()
    assert L.hiprtcCreateProgram(ctypes.byref(prog), src.encode(), b'k.hip',
                                 0, None, None) == 0
    opts = ['--offload-arch=' + emit.ARCH] + emit.DEVICE_FLAGS
    arr = (ctypes.c_char_p * len(opts))(*[o.encode() for o in opts])
    rc = L.hiprtcCompileProgram(prog, len(opts), arr)
    n = ctypes.c_size_t()
    L.hiprtcGetProgramLogSize(prog, ctypes.byref(n))
    log = ctypes.create_string_buffer(n.value + 1)
    L.hiprtcGetProgramLog(prog, log)
    L.hiprtcGetCodeSize(prog, ctypes.byref(n))
    L.hiprtcDestroyProgram(ctypes.byref(prog))
    return rc, log.value.decode(), n.value


@pytest.mark.parametrize('name', sorted(configs.ALL))
def test_hiprtc_compiles_config_kernels(name):
    "the run-time compile path accepts the emitted source (no device needed)"
    spec = configs.ALL[name](n_steps=10).spec()
    rc, log, size = _hiprtc_compile(spec.source())
    assert rc == 0, log
    assert size > 1000


def test_aot_code_objects_match_keys():
    """build() writes code objects named by the source hash that the run-time
    lookup computes."""
    kdir = os.path.join(os.path.dirname(hip.LIB_PATH), 'kernels')
    for name, fn in configs.ALL.items():
        key = fn(n_steps=10).spec().key()
        assert os.path.exists(os.path.join(kdir, key + '.hsaco')), name

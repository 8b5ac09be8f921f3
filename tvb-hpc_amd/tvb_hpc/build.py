"""
Ahead-of-time build of the native parts, in-tree.

* ``_lib/libtvbx.so``: the C-ABI runtime (csrc/tvbx_runtime.hip), hipcc for
  gfx950, linked against hipRTC.
* ``_lib/kernels/<key>.hsaco``: code objects of the emitted step kernels for
  the builtin sweep configurations, compiled by hipcc from exactly the source
  the run-time hipRTC path would compile (same flags), so a fresh process
  finds them by source hash and skips run-time compilation.

Run as ``python -m tvb_hpc.build`` or through ``__graft_entry__.build()``.
"""

import os
import subprocess
import sys
import tempfile

from . import emit

HERE = os.path.dirname(os.path.abspath(__file__))
LIBDIR = os.path.join(HERE, '_lib')
KDIR = os.path.join(LIBDIR, 'kernels')
LIB = os.path.join(LIBDIR, 'libtvbx.so')
REPO_INCLUDE = os.path.normpath(os.path.join(HERE, '..', '..', 'include'))
HIPCC = os.environ.get('HIPCC', '/opt/rocm/bin/hipcc')


def _run(cmd):
    r = subprocess.run(cmd, stdout=subprocess.PIPE, stderr=subprocess.STDOUT,
                       text=True)
    if r.returncode != 0:
        raise RuntimeError('build failed: %s\n%s' % (' '.join(cmd), r.stdout))
    return r.stdout


def build_runtime():
    os.makedirs(LIBDIR, exist_ok=True)
    src = os.path.join(emit.CSRC, 'tvbx_runtime.hip')
    tmp = LIB + '.tmp'
    cmd = [HIPCC, '--offload-arch=' + emit.ARCH, '-shared', '-fPIC',
           '-O3', '-std=c++17', '-ffp-contract=off',
           '-Xclang', '-target-feature', '-Xclang', '-packed-fp32-ops',
           '-I' + emit.CSRC, '-I' + REPO_INCLUDE, '-o', tmp, src,
           '-lhiprtc', '-Wno-unused-command-line-argument']
    _run(cmd)
    os.replace(tmp, LIB)
    return LIB


def build_kernel(spec: 'emit.KernelSpec', force=False):
    """hipcc --genco of the spec's source into _lib/kernels/<key>.hsaco."""
    os.makedirs(KDIR, exist_ok=True)
    key = spec.key()
    out = os.path.join(KDIR, key + '.hsaco')
    if os.path.exists(out) and not force:
        return out
    with tempfile.TemporaryDirectory() as td:
        src = os.path.join(td, key + '.hip')
        with open(src, 'w') as fd:
            fd.write('#include <hip/hip_runtime.h>\n' + spec.source())
        tmp = os.path.join(td, key + '.hsaco')
        cmd = [HIPCC, '--genco', '--offload-arch=' + emit.ARCH] + \
            emit.DEVICE_FLAGS + ['-o', tmp, src]
        _run(cmd)
        os.replace(tmp, out)
    return out


def builtin_specs():
    "Kernel specs used by the sweep configurations, tests and benchmark."
    from . import configs
    return configs.all_specs()


def build_all(jobs=8, kernels=True):
    build_runtime()
    if not kernels:
        return
    specs = builtin_specs()
    # drop code objects of sources that no longer exist
    keep = {s.key() + '.hsaco' for s in specs}
    if os.path.isdir(KDIR):
        for f in os.listdir(KDIR):
            if f.endswith('.hsaco') and f not in keep:
                os.remove(os.path.join(KDIR, f))
    from concurrent.futures import ThreadPoolExecutor
    with ThreadPoolExecutor(max_workers=jobs) as ex:
        list(ex.map(build_kernel, specs))


if __name__ == '__main__':
    build_all(kernels='--no-kernels' not in sys.argv)
    print(LIB)

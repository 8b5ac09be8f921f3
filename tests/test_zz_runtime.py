"""libtvbx and PyTorch share one HIP runtime in a process (checked in a
fresh interpreter, where libtvbx is loaded before torch is imported)."""

import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_libtvbx_binds_the_runtime_torch_loads():
    code = '''
import os, sys
sys.path.insert(0, os.path.join(%r, 'tvb-hpc_amd'))
from tvb_hpc import hip
hip.lib()
maps = open('/proc/self/maps').read()
libs = {l.split()[-1] for l in maps.splitlines()
        if 'libamdhip64' in l or 'libhsa-runtime64' in l}
print(len([l for l in libs if 'libamdhip64' in l]),
      len([l for l in libs if 'libhsa-runtime64' in l]))
''' % ROOT
    out = subprocess.run([sys.executable, '-c', code], capture_output=True,
                         text=True, timeout=600)
    assert out.returncode == 0, out.stderr[-2000:]
    assert out.stdout.split() == ['1', '1'], out.stdout

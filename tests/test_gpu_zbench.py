"""The benchmark's default command line, run as its own process on the
device: it must print one valid JSON line with every leg measured."""

import json
import os
import subprocess
import sys
import time

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.timeout(900)
def test_bench_command_prints_a_complete_line():
    cmd = [sys.executable, 'bench.py', '--gpus', '1', '--steps', '20',
           '--warmup', '5']
    t0 = time.perf_counter()
    r = subprocess.run(cmd, cwd=ROOT, stdout=subprocess.PIPE,
                       stderr=subprocess.PIPE, text=True, timeout=850)
    wall = time.perf_counter() - t0
    print(r.stderr[-4000:])
    assert r.returncode == 0, r.stderr[-4000:]
    lines = [ln for ln in r.stdout.splitlines() if ln.startswith('{')]
    assert len(lines) == 1, r.stdout
    d = json.loads(lines[0])
    print(json.dumps(d, indent=1)[:6000])
    assert 'errors' not in d, d.get('errors')
    assert d['n_gpus'] == 1 and d['steps'] == 20 and d['warmup'] == 5
    assert d['value'] > 0 and d['ms_per_step'] > 0
    assert d['ms_per_step'] * d['steps'] / 1e3 < wall
    assert abs(d['value'] - d['config']['n_inst'] /
               (d['ms_per_step'] * 1e-3)) < 1e-6 * d['value']
    rf = d['roofline']
    for k in ('bound', 'achieved', 'peak', 'unit', 'frac'):
        assert rf[k] is not None, k
    assert 0 < rf['compulsory']['frac'] <= 1.0
    cb = d['cpu_baseline']
    assert cb['value'] > 0 and cb['cores'] >= 1 and cb['kind'] == 'port'
    assert cb['cpu_model'] and cb['nproc'] >= 1
    fc = d['fc_err_vs_fp64']
    assert fc['steps'] == 10000
    for key in ('jansenrit998_bf16_b32', 'plain', 'jansenrit998_f16',
                'jansenrit998_u8'):
        assert fc[key] <= 1e-3, (key, fc[key])
    assert d['vs_plain'] > 1
    assert d['grad']['evals_per_s'] > 0
    assert 'checkpoint every 1000' in d['grad']['workload']

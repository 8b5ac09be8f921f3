"""The benchmark's default command line, run as its own process on the
device: it must print one valid JSON line with every leg measured."""

import json
import os
import subprocess
import sys
import tempfile
import time

import pytest

pytestmark = pytest.mark.gpu

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


@pytest.mark.timeout(900)
def test_bench_command_prints_a_complete_line(capsys):
    cmd = [sys.executable, 'bench.py', '--gpus', '1', '--steps', '20',
           '--warmup', '5']
    t0 = time.perf_counter()
    with tempfile.TemporaryFile(mode='w+') as fout:
        proc = subprocess.Popen(cmd, cwd=ROOT, stdout=fout,
                                stderr=subprocess.PIPE, text=True)
        # the benchmark's progress lines, shown as they come (past pytest's
        # output capture) while the command runs
        err = []
        for line in proc.stderr:
            err.append(line)
            with capsys.disabled():
                sys.stdout.write('  bench: ' + line)
                sys.stdout.flush()
        rc = proc.wait(timeout=60)
        wall = time.perf_counter() - t0
        fout.seek(0)
        out = fout.read()
    err = ''.join(err)
    assert rc == 0, err[-4000:]
    lines = [ln for ln in out.splitlines() if ln.startswith('{')]
    assert len(lines) == 1, out
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

"""bench.py's contract pieces that need no GPU: more than one GPU is refused
with a message, and the HBM traffic per launch it reports is read from the
committed rocprof summary for both forms and all three kernel classes."""
import os
import subprocess
import sys

import bench

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=300, cwd=ROOT)
    assert r.returncode == 2
    assert 'one GPU' in r.stderr
    assert r.stdout == ''


def test_traffic_summary_covers_every_kernel_class():
    for precision in ('fp32-bulk', 'fp64'):
        for k in bench.LAUNCHES_PER_STEP:
            v = bench.hbm_traffic(precision, k)
            assert v is not None and 1e6 < v < 1e10, (precision, k, v)


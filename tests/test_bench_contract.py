"""bench.py's command-line contract on the CPU: more than one GPU is
refused with a message (one-GPU build), and the traffic table it reads from
profiles/ parses."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_is_refused():
    r = subprocess.run([sys.executable, os.path.join(ROOT, 'bench.py'), '--gpus', '2'],
                       capture_output=True, text=True, timeout=120)
    assert r.returncode != 0
    assert 'one GPU' in r.stderr
    assert r.stdout.strip() == ''


def test_measured_traffic_parses():
    sys.path.insert(0, ROOT)
    import bench
    t = bench.measured_traffic('C2 level-0 Jacobi')
    assert t is None or 1e7 < t < 1e9
    assert bench.measured_traffic('no such kernel') is None

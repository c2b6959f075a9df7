"""bench.py's command-line contract on the host: a run asked for more than one GPU exits with
status 2 and says why, before touching a GPU; the algorithmic bytes of one VI sweep (the
roofline's numerator) follow the layout: S and T per entry, R per (s, a), V read, V' and pi
written per state."""
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_more_than_one_gpu_exits_with_a_message():
    for n in (2, 8):
        p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", str(n)],
                           capture_output=True, text=True, timeout=120)
        assert p.returncode == 2
        assert "one GPU" in p.stderr and p.stdout == ""


def test_vi_bytes_per_sweep():
    sys.path.insert(0, ROOT)
    import bench
    # n = 1e6, m = 8, ns = 16: 1.024e9 (S, T) + 3.2e7 (R) + 1.2e7 (V, V', pi)
    assert bench.vi_bytes_per_sweep(1_000_000, 8, 16) == 1_068_000_000
    assert bench.vi_bytes_per_sweep(1, 1, 1) == 8 + 4 + 12

"""bench.py pieces that run without a GPU: the one-GPU guard, the traffic
figures it reads from the committed rocprofv3 summary, and the per-iteration
J+K timing wrapper of the SCF."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)
import bench  # noqa: E402


def test_more_than_one_gpu_is_refused_with_a_message():
    r = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True,
                       text=True, timeout=120)
    assert r.returncode == 2
    assert "one GPU" in r.stderr
    assert r.stdout == ""


def test_traffic_figures_come_from_the_committed_profile():
    f64 = bench.measured_traffic("traffic_bytes_per_fp64_build")
    mixed = bench.measured_traffic("traffic_bytes_per_mixed_build")
    assert f64 is not None and mixed is not None
    # above the algorithmic bytes (D in, J and K out: 61.5 MB) and below 1 TB
    assert 6.15e7 < mixed < 1e12 and 6.15e7 < f64 < 1e12
    assert bench.measured_traffic("no_such_key") is None


def test_timed_engine_records_every_device_build():
    class Stats:
        def __init__(self, t):
            self.jk_seconds = t

    class FakeEngine:
        nbf = 7

        def __init__(self):
            self.calls = 0

        def fock_2e_device(self, dP, dG):
            self.calls += 1
            return Stats(0.01 * self.calls)

        def set_density_screening(self, thr):
            return thr

    e = FakeEngine()
    t = bench.JKTimedEngine(e)
    assert t.nbf == 7 and t.set_density_screening(3.0) == 3.0
    for _ in range(3):
        assert t.fock_2e_device(None, None).jk_seconds == pytest.approx(0.01 * e.calls)
    assert t.jk == pytest.approx([0.01, 0.02, 0.03])

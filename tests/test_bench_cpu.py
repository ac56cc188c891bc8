"""bench.py's contract on the CPU: more than one GPU is refused with a
message, the traffic figure comes from the committed profile summary, and
the detectors are the 8 annuli at 2-10 cm."""
import os
import subprocess
import sys

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import bench  # noqa: E402


def test_more_than_one_gpu_is_refused():
    p = subprocess.run([sys.executable, os.path.join(ROOT, "bench.py"), "--gpus", "2"], capture_output=True, text=True,
                       timeout=120)
    assert p.returncode == 2 and "one GPU" in p.stderr and p.stdout == ""


def test_traffic_from_the_committed_profile():
    t = bench.traffic_per_photon()
    assert t is not None and 1e3 < t < 1e6
    with open(bench.PROFILE) as f:
        assert "traffic_bytes_per_photon: %s" % ("%.1f" % t) in f.read()


def test_detectors():
    lo, hi = bench.detectors()
    c = 0.5 * (lo + hi)
    assert len(c) == 8 and np.isclose(c[0], 2.0) and np.isclose(c[-1], 10.0)
    assert np.all(hi - lo == 0.5) and np.all(np.diff(c) > 0.5)

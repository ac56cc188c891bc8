"""The torch binding of the adjoint on the GPU: torch.autograd through the
worker process against central differences in SpO2, and the Adam SpO2 fit.
Each case runs in a child process (tests/torch_adjoint_gpu_case.py) that
holds torch; that process starts the worker that holds liblp.so."""
import os
import subprocess
import sys
import threading

import pytest

pytestmark = pytest.mark.gpu
HERE = os.path.dirname(os.path.abspath(__file__))


def _run(case, timeout):
    """Runs one case, passing its output on line by line as it comes."""
    print("torch adjoint case", case, flush=True)
    p = subprocess.Popen([sys.executable, "-u", os.path.join(HERE, "torch_adjoint_gpu_case.py"), case],
                         stdout=subprocess.PIPE, stderr=subprocess.STDOUT, text=True)
    timer = threading.Timer(timeout, p.kill)
    timer.start()
    lines = []
    try:
        for line in p.stdout:
            lines.append(line)
            print(line, end="", flush=True)
        rc = p.wait()
    finally:
        timer.cancel()
    assert rc == 0 and any("case passed" in x for x in lines), "exit status %s" % rc


def test_autograd_against_central_differences():
    _run("autograd", 300)


def test_adam_spo2_fit():
    _run("fit", 480)

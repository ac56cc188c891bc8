"""Cases of tests/test_gpu_torch_adjoint.py, each run in a process of its
own: this process holds torch (on the CPU), and the transport runs in the
worker process that lightpropagation_amd.torch_adjoint starts (liblp.so on
the GPU).  A pytest process that has loaded liblp.so for other GPU tests
must not load torch as well, hence the separate process.

    python tests/torch_adjoint_gpu_case.py autograd|fit
"""
import os
import sys
import time

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402
import torch  # noqa: E402

from lightpropagation_amd import torch_adjoint as TA  # noqa: E402


def detectors():
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    return c - 0.25, c + 0.25


def case_autograd():
    """d/dSpO2 of sum_d log S_d(660 nm) through SpO2Optics, DetectorSignals
    and the worker, against central differences in SpO2 with common random
    numbers (forward launches only), 4 batches of 2e6 photons."""
    lo, hi = detectors()
    with TA.Simulator(("maternal_fetal", {}), seed=0x5EED0004, det_lo=lo, det_hi=hi) as sim:
        o = TA.SpO2Optics(sim, 660.0)
        x0 = np.array([0.97, 0.50])
        B, N, h = 4, 2_000_000, np.array([0.01, 0.05])
        dets = slice(0, 6)
        ad = np.zeros((B, 2))
        fd = np.zeros((B, 2))
        for b in range(B):
            x = torch.tensor(x0, dtype=torch.float64, requires_grad=True)
            S = TA.detector_signals(o.mua(x), o.mus.clone(), sim, b * N, N)
            torch.log(S[dets]).sum().backward()
            ad[b] = x.grad.numpy()
            for k in range(2):
                L = []
                for sgn in (-1, 1):
                    xs = x0.copy()
                    xs[k] += sgn * h[k]
                    with torch.no_grad():
                        nm = np.array(o.nmuasg, np.float32)
                        nm[:, 1] = o.mua(torch.tensor(xs)).numpy()
                    L.append(np.log(sim.forward(nm, b * N, N)["signals"][dets]).sum())
                fd[b, k] = (L[1] - L[0]) / (2 * h[k])
            print("batch %d: autograd %s  finite differences %s" % (b, ad[b], fd[b]), flush=True)
        a, f = ad.mean(0), fd.mean(0)
        se = np.sqrt(ad.var(0, ddof=1) / B + fd.var(0, ddof=1) / B)
        print("autograd", a, "finite differences", f, "se", se, flush=True)
        assert a[0] > 0 and f[0] > 0  # oxygenated blood absorbs less at 660 nm: the signals rise
        assert np.all(np.abs(a - f) <= 4 * se + 0.05 * np.abs(a)), (a, f, se)


def case_fit():
    """Adam fits of one saturation with the other known, to two-wavelength
    detector data made by the same photons at SpO2 (0.97, 0.50): the fit
    must return the saturation that made the data.

    Common random numbers do not make the loss smooth in the fetal
    saturation: changing a tissue's mua by 1e-5 moves float32 step lengths
    by an ulp, and the paths of the photons that cross that tissue diverge,
    so the loss carries noise of the size of those photons' Monte Carlo
    error.  On the 2.5 cm-deep head of the bench mesh the far detectors
    collect about 5 units of weight per 2e6 photons and that noise is as
    large as the fetal signal itself; the fetal fit therefore runs on the
    head 1.5 cm deep under thinner maternal layers (about 100 units at the
    far detectors), the maternal fit on the bench mesh.  (Fitting both at
    once needs the adjoint's fetal column to be accurate to better than its
    small angle to the maternal column -- J^T J has condition numbers of
    1e4-1e5 here -- which takes more photons per evaluation than a test
    runs.)"""
    lo, hi = detectors()
    cases = (("fetal", dict(t_fat=0.3, t_uterus=0.5, fetal_depth=1.5, freq_maternal=24, freq_fetal=24),
              4_000_000, (0.97, 0.75), (False, True), 1, 0.05, [np.ones(8), np.array([1, 1, 1, 1, 0, 0, 0, 0.0])]),
             ("maternal", {}, 2_000_000, (0.90, 0.50), (True, False), 0, 0.005,
              [np.ones(8), np.array([1, 1, 1, 0, 0, 0, 0, 0.0])]))
    for name, kw, count, init, free, k, tol, W in cases:
        with TA.Simulator(("maternal_fetal", kw), seed=0x5EED0005, det_lo=lo, det_hi=hi) as sim:
            opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
            meas = [sim.forward(sim.optics(o.wavelength, 0.97, 0.50), 0, count)["signals"] for o in opts]
            t0 = time.time()
            x, hist = TA.fit_spo2(sim, opts, meas, init=init, steps=40, count=count, weights=W, free=free,
                                  callback=lambda it, h: print("%s it %d loss %.4g SpO2 %s" % (name, it, h[0], h[1]),
                                                               flush=True))
            print(name, "fit", x, "%.1f s" % (time.time() - t0), flush=True)
            assert abs(x[k] - (0.97, 0.50)[k]) < tol and x[1 - k] == init[1 - k], x


if __name__ == "__main__":
    {"autograd": case_autograd, "fit": case_fit}[sys.argv[1]]()
    print("case passed", flush=True)

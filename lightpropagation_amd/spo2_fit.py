"""Fetal SpO2 from synthetic two-wavelength detector data on the
maternal-fetal mesh: the adjoint's torch binding and the whitened Adam fit
(lightpropagation_amd.torch_adjoint) end to end.

    python -m lightpropagation_amd.spo2_fit [--count 10000000] [--measured 400000000]
        [--final-count 100000000 --final-steps 8]

The "measured" signals are a forward run at SpO2 (0.97 maternal, 0.50 fetal)
with photons of their own (indices from 2^40); the fit starts the fetal
saturation at 0.75, keeps the maternal one at its known value (a maternal
pulse oximeter), and evaluates every step on photons [0, count) (common
random numbers); a refinement stage then continues from the result with
final_steps steps on photons [count, count + final_count) (the loss's
Monte Carlo noise, not the optimiser, limits the first stage: the fetal head
changes log S by about 0.2 per unit SpO2 at the far detectors).  At 940 nm
only the three nearest detectors collect enough weight to fit.  Prints one JSON line: the recovered saturation, its error,
the loss at the start and at the end, and the time per Adam step.  Runs the
transport in a worker process; this process holds torch.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import torch_adjoint as TA


def detectors():
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    return c - 0.25, c + 0.25


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=10_000_000, help="photons per loss evaluation")
    ap.add_argument("--measured", type=int, default=400_000_000, help="photons of the synthetic data")
    ap.add_argument("--steps", type=int, default=40)
    ap.add_argument("--final-count", type=int, default=100_000_000, help="photons per evaluation, refinement")
    ap.add_argument("--final-steps", type=int, default=8, help="refinement steps (0: none)")
    ap.add_argument("--init", type=float, default=0.75)
    args = ap.parse_args(argv)
    lo, hi = detectors()
    truth = (0.97, 0.50)
    with TA.Simulator(("maternal_fetal", {}), seed=0x5EED0006, det_lo=lo, det_hi=hi) as sim:
        opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
        t0 = time.perf_counter()
        meas = []
        for o in opts:
            meas.append(sim.forward(sim.optics(o.wavelength, *truth), 1 << 40, args.measured)["signals"])
            print("data at %g nm: %d photons" % (o.wavelength, args.measured), file=sys.stderr, flush=True)
        t_meas = time.perf_counter() - t0
        W = [np.ones(8), np.array([1, 1, 1, 0, 0, 0, 0, 0.0])]
        t0 = time.perf_counter()
        x, hist = TA.fit_spo2(sim, opts, meas, init=(truth[0], args.init), steps=args.steps, count=args.count,
                              weights=W, free=(False, True),
                              callback=lambda it, h: print("step %d loss %.4g fetal SpO2 %.4f" % (it, h[0], h[1][1]),
                                                           file=sys.stderr, flush=True))
        t_fit = time.perf_counter() - t0
        x1 = x.copy()
        t1 = time.perf_counter()
        hist2 = []
        if args.final_steps > 0:
            x, hist2 = TA.fit_spo2(sim, opts, meas, init=(truth[0], float(x[1])), steps=args.final_steps, lr=0.1,
                                   lr_end=0.005, first=args.count, count=args.final_count, weights=W,
                                   free=(False, True),
                                   callback=lambda it, h: print("refine %d loss %.4g fetal SpO2 %.4f" %
                                                                (it, h[0], h[1][1]), file=sys.stderr, flush=True))
        t_ref = time.perf_counter() - t1
    out = {"workload": "maternal-fetal mesh, 660 + 940 nm, 8 detectors; fetal SpO2 fitted from %.2f with the "
                       "maternal value known; %d photons per evaluation, synthetic data from %d independent photons"
                       % (args.init, args.count, args.measured),
           "fetal_spo2_true": truth[1], "fetal_spo2_fit": float(x[1]), "error": float(x[1] - truth[1]),
           "first_stage_fit": float(x1[1]), "loss_first": hist[0][0], "loss_first_stage_last": hist[-1][0],
           "loss_last": (hist2 or hist)[-1][0],
           "fetal_spo2_path": [round(float(h[1][1]), 4) for h in hist[::5]] + [round(float(h[1][1]), 4) for h in hist2],
           "steps": args.steps, "seconds_per_step": t_fit / args.steps, "final_steps": args.final_steps,
           "final_count": args.final_count, "refine_seconds": t_ref, "data_seconds": t_meas}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()

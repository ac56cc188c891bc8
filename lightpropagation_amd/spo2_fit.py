"""Maternal and fetal SpO2 fitted jointly to synthetic two-wavelength
detector data on the maternal-fetal mesh (fetal head 2.5 cm deep): the
adjoint's path records, their reweighting in torch and the whitened Adam
fit (lightpropagation_amd.torch_adjoint.fit_spo2_paths) end to end.

    python -m lightpropagation_amd.spo2_fit [--count 100000000] [--measured 400000000]
        [--rounds 3] [--seeds 1] [--init 0.90,0.70]

For each seed the "measured" signals are a forward run at SpO2 (0.97
maternal, 0.50 fetal) with photons of their own (indices from 2^40, context
seed 0x5EED0100 + seed); the fit starts both saturations at --init and runs
--rounds rounds: each runs photons [0, count) at the current estimate with
path records at 660 and 940 nm and fits on them (the earlier rounds with
count / 4 photons: they only bring the estimate near the answer, where the
last round's reweighting is nearly the identity).  Each detector's residual
is scaled by its Monte Carlo standard error, so the loss at the optimum is a
chi-square and the Gauss-Newton covariance predicts the spread of the
recovered saturations over seeds.  Prints one JSON line: per seed the
recovered SpO2, the predicted standard deviations, chi-square and times; over
the seeds the mean, the spread (sample standard deviation) and the mean
predicted standard deviation.  Runs the transport in a worker process; this
process holds torch.
"""
import argparse
import json
import sys
import time

import numpy as np

from . import torch_adjoint as TA

TRUTH = (0.97, 0.50)


def detectors():
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    return c - 0.25, c + 0.25


def fit_one(seed, count, measured, rounds, init, free=(True, True), scene=("maternal_fetal", {}), log=None):
    """One synthetic data set and one joint fit; returns a dict of results."""
    lo, hi = detectors()
    with TA.Simulator(scene, seed=0x5EED0100 + seed, det_lo=lo, det_hi=hi) as sim:
        opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
        t0 = time.perf_counter()
        meas = [sim.forward(sim.optics(o.wavelength, *TRUTH), 1 << 40, measured)["signals"] for o in opts]
        t_data = time.perf_counter() - t0
        if log:
            log("seed %d: data from %d photons per wavelength in %.1f s" % (seed, measured, t_data))
        counts = [max(count // 4, 1)] * (rounds - 1) + [count]
        t0 = time.perf_counter()
        cb = None
        if log:
            def cb(it, h):
                if it % 50 == 0:
                    log("  step %d loss %.4g SpO2 %.4f %.4f" % (it, h[0], h[1][0], h[1][1]))
        x, info = TA.fit_spo2_paths(sim, opts, meas, init=init, count=counts, rounds=rounds, measured_count=measured,
                                    free=free, callback=cb)
        t_fit = time.perf_counter() - t0
    sd = np.sqrt(np.diag(info["cov"]))
    out = dict(seed=seed, spo2=[float(v) for v in x], error=[float(x[k] - TRUTH[k]) for k in range(2)],
               predicted_sd=[float(v) for v in sd], chi2=info["chi2"], dof=info["dof"],
               rounds=[dict(reference=[round(v, 4) for v in r["reference"]],
                            result=[round(v, 4) for v in r["result"]], photons=r["photons"],
                            transport_s=round(r["transport_s"], 2)) for r in info["rounds"]],
               data_s=t_data, fit_s=t_fit)
    if log:
        log("seed %d: SpO2 %.4f %.4f (predicted sd %.4f %.4f), chi2 %.1f / %d, %.1f s" %
            (seed, x[0], x[1], sd[0], sd[1], info["chi2"], info["dof"], t_fit))
    return out


def main(argv=None):
    ap = argparse.ArgumentParser()
    ap.add_argument("--count", type=int, default=100_000_000, help="photons per wavelength, last round")
    ap.add_argument("--measured", type=int, default=400_000_000, help="photons per wavelength of the data")
    ap.add_argument("--rounds", type=int, default=3)
    ap.add_argument("--seeds", type=int, default=1)
    ap.add_argument("--init", default="0.90,0.70")
    args = ap.parse_args(argv)
    init = tuple(float(v) for v in args.init.split(","))

    def log(msg):
        print(msg, file=sys.stderr, flush=True)

    res = [fit_one(s, args.count, args.measured, args.rounds, init, log=log) for s in range(args.seeds)]
    xs = np.array([r["spo2"] for r in res])
    out = {"workload": "maternal-fetal mesh (fetal head 2.5 cm deep), 660 + 940 nm, 8 detectors; maternal and fetal "
                       "SpO2 fitted jointly from %s on path records of %d photons per wavelength (last round of %d); "
                       "synthetic data at %s from %d independent photons per wavelength"
                       % (args.init, args.count, args.rounds, TRUTH, args.measured),
           "truth": list(TRUTH), "seeds": args.seeds, "fits": res,
           "mean": xs.mean(0).tolist(),
           "spread": xs.std(0, ddof=1).tolist() if len(xs) > 1 else None,
           "mean_predicted_sd": np.mean([r["predicted_sd"] for r in res], 0).tolist(),
           "seconds_per_fit": float(np.mean([r["data_s"] + r["fit_s"] for r in res]))}
    print(json.dumps(out), flush=True)
    return out


if __name__ == "__main__":
    main()

"""Benchmark: photons/s of the triangle-mesh photon transport on one MI355X.

Workload: the maternal-fetal abdomen mesh (7 concentric geodesic shells,
202 160 faces: maternal skin 0.2 cm, fat 1.0 cm, uterine wall 0.8 cm, amniotic
fluid 0.5 cm over a fetal head of radius 4.5 cm with scalp 0.1 cm and skull
0.3 cm at 2.5 cm depth), 660 nm, SpO2 0.97 maternal / 0.50 fetal, float32,
roulette 1e-4 / 0.1, pencil beam at the origin.  Tallies: 8 annular
detectors (2-10 cm), a 256^3 int64 fixed-point absorption grid (0.05 cm
voxels), per-region absorption, 200 radial bins, the SFT and RTd images.
One step = one lp_run of --photons photons with inputs resident on the GPU.

Prints one JSON line: the metric, the FP32-VALU roofline of the transport
kernel (algorithmic work per photon step over the launch time measured with
HIP events), its memory traffic per launch (from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary of this workload, scaled per photon), and
the CPU baseline (the oracle on a bounded sample of the same workload).
"""
import argparse
import json
import math
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

METRIC = "photons/sec on fetal mesh (1 GPU); adjoint vs finite-diff gradient time"

# Algorithmic fp32 work per photon step (unit = one loop iteration), counting
# add/sub/mul/min/max/compare-select as 1, FMA as 2, and div/sqrt/log/cos/acos
# as 1 each: interaction 200 + 25, boundary 110 + 15; ideal search: one
# triangle test (52 + 3) plus ceil(log2 T) box tests of 20.
F_INT, F_BND, F_TRI, F_BOX = 225.0, 125.0, 55.0, 20.0
FP32_PEAK_TFLOPS = 157.3
PROFILE = os.path.join(ROOT, "profiles", "transport_kernel_maternal_fetal_202k.md")


def traffic_per_photon():
    """Memory traffic of the transport kernel per photon, from the committed
    FETCH_SIZE / WRITE_SIZE summary of this workload (calibrated against a
    copy kernel there); None when the summary is absent."""
    try:
        with open(PROFILE) as f:
            for line in f:
                if line.startswith("traffic_bytes_per_photon:"):
                    return float(line.split(":", 1)[1])
    except OSError:
        pass
    return None


def detectors():
    centres = 2.0 + 8.0 * np.arange(8) / 7.0
    return centres - 0.25, centres + 0.25


def setup(ctx, sc, seed):
    ctx.set_scene(sc, seed=seed)
    lo, hi = detectors()
    ctx.set_detectors(lo, hi)
    ctx.set_voxel_grid([-6.4, -6.4, 0.0], [0.05, 0.05, 0.05], [256, 256, 256])
    ctx.set_radial(0.05, 200)
    mask = sum(1 << k for k in (0, 1, 4, 5, 6, 7, 8))  # sft rtd region vox rad det misc
    ctx.set_tallies(mask, record_exit=False)


def cpu_baseline(sc, seconds=20.0):
    """Oracle (the reference algorithm: brute-force O(T) face search per
    step, one photon per OpenMP thread) on a bounded sample of the same
    workload, on this host's cores."""
    sys.path.insert(0, os.path.join(ROOT, "tests"))
    import _oracle as O
    try:
        cores = len(os.sched_getaffinity(0))
    except AttributeError:
        cores = os.cpu_count() or 1
    cores = max(1, min(cores, 16))
    run = O.Run(sc, seed=2024)
    done, steps, t0 = 0, 0, time.perf_counter()
    batch = cores
    while time.perf_counter() - t0 < seconds and done < 4096:
        run.run(1000000 + done, batch, nthreads=cores)
        done += batch
    dt = time.perf_counter() - t0
    steps = run.tallies()["misc"]["steps"]
    return {"value": done / dt, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": "%d photons of the same workload (%d steps) in %.1f s; photon-steps/s %.3g" %
                      (done, steps, dt, steps / dt)}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--photons", type=int, default=16_000_000, help="photons per step")
    ap.add_argument("--wavelength", type=float, default=660.0)
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--lds-nodes", type=int, default=-1)
    ap.add_argument("--safety-cells", type=int, default=-1)
    ap.add_argument("--opt", action="append", default=[], help="tuning option key=value (lp_set_option)")
    ap.add_argument("--calibrate-copy", type=int, default=0,
                    help="also run a plain copy of this many MiB (reference for profiler byte counters)")
    args = ap.parse_args()
    if args.gpus != 1:
        print("bench.py measures one GPU: this library runs on a single MI355X; sharding photons over "
              "several GPUs (disjoint photon ranges, one int64 all-reduce of the tallies) is not "
              "implemented", file=sys.stderr)
        sys.exit(2)

    import lightpropagation_amd as lp
    from lightpropagation_amd import scenes

    sc = scenes.maternal_fetal(wavelength_nm=args.wavelength)
    T = len(sc["tri"])
    ctx = lp.Context(0)
    setup(ctx, sc, seed=0x5EED0001)
    ctx.set_launch(args.grid, args.lds_nodes, args.safety_cells)
    for kv in args.opt:
        k, v = kv.split("=")
        ctx.set_option(k, float(v))
    # lp_run blocks on a HIP event recorded after the kernel on the context's
    # stream, so the timed region is bracketed by completed device work.  torch
    # is not initialised on the GPU here: it ships its own HIP/HSA runtime, and
    # a second runtime on the device in one process is not supported.
    print("setup done: %d faces" % T, file=sys.stderr, flush=True)

    first = 0
    for _ in range(args.warmup):
        ctx.run(first, args.photons)
        first += args.photons
        print("warmup step done", file=sys.stderr, flush=True)
    ctx.reset()
    t0 = time.perf_counter()
    kms = []
    for i in range(args.steps):
        ctx.run(first, args.photons)
        first += args.photons
        kms.append(ctx.stats().kernel_ms)
        print("step %d: %.1f ms kernel" % (i, kms[-1]), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    st = ctx.stats()
    t = ctx.tallies()
    m = t["misc"]
    nph = args.steps * args.photons
    steps = m["steps"]
    kernel_s = sum(kms) / 1e3
    search = F_TRI + F_BOX * math.ceil(math.log2(max(T, 2)))
    flops = m["interact"] * (F_INT + search) + m["boundary"] * (F_BND + search)
    achieved = flops / kernel_s / 1e12
    tpp = traffic_per_photon()
    R = t["sft"].reshape(-1, 3).sum(0) / lp.FIXED / nph
    det = t["det"].reshape(-1, 2)[:, 0] / lp.FIXED / nph
    out = {
        "metric": METRIC,
        "value": nph / wall,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "maternal-fetal mesh, %d faces, %g nm, SpO2 0.97/0.50, %d photons per step, "
                        "8 detectors + 256^3 int64 grid + region/radial/SFT/RTd tallies" % (T, args.wavelength,
                                                                                         args.photons),
            "faces": T, "photons_per_step": args.photons, "rng": "philox4x32-10",
        },
        "photon_steps_per_s": steps / wall,
        "photons_per_min": nph / wall * 60,
        "kernel_ms_per_step": sum(kms) / len(kms),
        "steps_per_photon": steps / nph,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": (tpp * args.photons) if tpp is not None else None,
                     "traffic_unit": "bytes per launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)",
                     "work": "algorithmic fp32 per step: interaction %g + boundary %g, search %g (T=%d)" %
                             (F_INT, F_BND, search, T)},
        "counters": {k: m[k] for k in ("interact", "boundary", "escaped", "capped", "nan", "node_visits",
                                       "tri_tests", "skipped", "dist_queries", "dist_visits", "list_tests",
                                       "wave_iters", "search_rounds", "bvh_rounds")},
        "options": args.opt,
        "physics": {"R_total": float(R[0]), "F_total": float(R[1]), "T_total": float(R[2]),
                    "detector_R": [float(x) for x in det]},
        "grid_blocks": st.grid_blocks, "bvh_nodes": st.bvh_nodes, "bvh_depth": st.bvh_depth,
    }
    if args.calibrate_copy:
        ms = ctx.debug_copy(args.calibrate_copy << 20)
        out["copy_calibration"] = {"bytes": args.calibrate_copy << 20, "ms": ms,
                                   "GB_per_s": 2 * (args.calibrate_copy << 20) / ms / 1e6}
    ctx.close()
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

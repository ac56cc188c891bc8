"""Benchmark: photons/s of the triangle-mesh photon transport on one MI355X.

Workload (the headline run): the ten-region maternal-fetal abdomen mesh
(196 420 faces: maternal skin 0.2 cm, fat 1.0 cm, uterine wall 0.8 cm,
amniotic fluid 0.5 cm over a fetal head of radius 4.5 cm -- scalp 0.1 cm,
skull 0.3 cm, CSF 0.2 cm, grey matter 0.4 cm, white matter -- at 2.5 cm
depth, and a fetal torso below it), 10^8 photons per wavelength at 660 and
940 nm, SpO2 0.97 maternal / 0.50 fetal, float32, roulette 1e-4 / 0.1,
pencil beam at the origin.  Tallies: 8 annular detectors (2-10 cm), a 256^3
int64 fixed-point absorption grid (0.05 cm voxels), per-region absorption,
200 radial bins, the SFT and RTd images.  One step = one lp_run of
--photons photons of one wavelength with inputs resident on the GPU; the
steps alternate between 660 and 940 nm (warm-up likewise).

Prints one JSON line: the metric, the FP32-VALU roofline of the transport
kernel (algorithmic work per photon step over the launch time measured with
HIP events), its memory traffic per launch (from the committed rocprofv3
FETCH_SIZE / WRITE_SIZE summary of this workload, scaled per photon), the
adjoint against central finite differences at equal gradient error (the
metric's second half, on 20 region parameters and the voxel Jacobians), the
batched 128-instance and the 2e6-face workloads at reduced photon counts
(labelled), and the CPU baseline (the oracle on a bounded sample of the
same workload).
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
PROFILE = os.path.join(ROOT, "profiles", "transport_kernel_maternal_fetal_196k.md")


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
    avail = cores
    cores = max(1, min(cores, 16))  # at most 16 threads, whatever the host has
    run = O.Run(sc, seed=2024)
    done, steps, t0 = 0, 0, time.perf_counter()
    batch = cores
    while time.perf_counter() - t0 < seconds and done < 4096:
        run.run(1000000 + done, batch, nthreads=cores)
        done += batch
    dt = time.perf_counter() - t0
    steps = run.tallies()["misc"]["steps"]
    model, mhz = None, []
    try:
        with open("/proc/cpuinfo") as f:
            for ln in f:
                if ln.startswith("model name") and model is None:
                    model = ln.split(":", 1)[1].strip()
                elif ln.startswith("cpu MHz"):
                    mhz.append(float(ln.split(":", 1)[1]))
    except OSError:
        pass
    max_mhz = None
    try:
        with open("/sys/devices/system/cpu/cpu0/cpufreq/cpuinfo_max_freq") as f:
            max_mhz = int(f.read().strip()) / 1000.0
    except (OSError, ValueError):
        pass
    lscpu = {}
    try:  # lscpu's view of the host: model, clock, sockets and cores
        import subprocess
        out = subprocess.run(["lscpu"], stdout=subprocess.PIPE, stderr=subprocess.DEVNULL, text=True,
                             timeout=10).stdout
        want = ("Model name", "CPU(s)", "Thread(s) per core", "Core(s) per socket", "Socket(s)", "CPU max MHz",
                "CPU MHz", "BogoMIPS", "L3 cache", "Hypervisor vendor")
        for ln in out.splitlines():
            k, _, v = ln.partition(":")
            if k.strip() in want:
                lscpu[k.strip()] = v.strip()
    except (OSError, ValueError, subprocess.SubprocessError):
        pass
    return {"value": done / dt, "unit": "photons/s", "cores": cores, "kind": "port",
            "sample": "%d photons of the same workload (660 nm; %d steps) in %.1f s, one photon per OpenMP "
                      "thread; photon-steps/s %.3g" % (done, steps, dt, steps / dt),
            "host": {"cpu_model": model, "cpu_mhz_now_mean": (sum(mhz) / len(mhz)) if mhz else None,
                     "cpu_mhz_max": max_mhz, "cores_available": avail, "logical_cpus": os.cpu_count(),
                     "threads_used": cores, "lscpu": lscpu}}


def adjoint_vs_fd(sc, photons=10_000_000, batches=16, rels=(1e-3, 2e-2), jac_n=64):
    """Gradient of the 8 detector signals with respect to every tissue's mua
    and mus (10 regions x 2 = 20 parameters) and to mua and mus in each voxel
    of a jac_n^3 grid: the adjoint (lp_adjoint_voxels: forward launch +
    replay of the detected photons) against central finite differences with
    common random numbers (step rel x the value, two forward passes per
    parameter; the voxel Jacobians would take 4 jac_n^3 of them), for each
    step in rels.

    Times: one adjoint call over all `photons` against one forward pass
    (lp_run) of the same photons; finite differences cost 2 forward passes
    per parameter.  Standard errors: both methods over the same photons in
    `batches` batches (the spread of the batch means).  The time a method
    needs for a given error scales as time x SE^2, so the speedup at equal
    error is (T_fd SE_fd^2) / (T_adj SE_adj^2), over the entries the adjoint
    resolves (|mean| > 3 SE)."""
    import lightpropagation_amd as lp
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=0x5EED0003)
    lo, hi = detectors()
    ctx.set_detectors(lo, hi)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
    # emitter frame: x, y across the abdomen, z into it; 64^3 voxels over 20.5 x 20.5 x 7.7 cm
    ctx.set_jacobian_grid([-10.24, -10.24, 0.0], [0.32, 0.32, 0.12], [jac_n] * 3)
    nm0 = np.asarray(sc["nmuasg"], np.float32)
    params = [(t, c) for t in range(1, len(nm0)) for c in (1, 2)]
    nd, P, B, n = len(lo), len(params), batches, photons // batches
    # one forward pass and one adjoint over all photons (warm first)
    ctx.run(1 << 40, 100_000)
    ctx.reset()
    t0 = time.perf_counter()
    ctx.run(0, B * n)
    t_fwd = time.perf_counter() - t0
    fwd_kernel_ms = ctx.stats().kernel_ms
    ctx.reset()
    t0 = time.perf_counter()
    full = ctx.adjoint(0, B * n, jac=True, jac_mus=True)
    t_adj = time.perf_counter() - t0
    st = ctx.stats()
    if st.replay_mismatch:
        raise RuntimeError("adjoint replay mismatch")
    full_fwd_ms, full_rep_ms, pairs = st.kernel_ms, st.replay_ms, full["pairs"]
    jac_nz = int(np.count_nonzero(full["jac"]))
    jac_mus_nz = int(np.count_nonzero(full["jac_mus"]))
    print("adjoint: %d photons in %.2f s (forward %.0f ms + replay %.0f ms of %d pairs); forward pass %.2f s"
          % (B * n, t_adj, full_fwd_ms, full_rep_ms, pairs, t_fwd), file=sys.stderr, flush=True)
    # batch estimates for the standard errors
    adj = np.zeros((B, nd, P))
    t_batches = 0.0
    for b in range(B):
        ctx.reset()
        t0 = time.perf_counter()
        r = ctx.adjoint(b * n, n)
        t_batches += time.perf_counter() - t0
        if ctx.stats().replay_mismatch:
            raise RuntimeError("adjoint replay mismatch")
        g = r["grad"] / lp.FIXED / n
        for i, (t, c) in enumerate(params):
            adj[b, :, i] = g[:, t, c - 1]
    ma = adj.mean(0)
    sa = adj.std(0, ddof=1) / math.sqrt(B)
    res = np.abs(ma) > 3 * sa
    kinds = np.array([c for _, c in params])[None, :].repeat(nd, 0)
    out = {
        "workload": "maternal-fetal mesh (10 regions), %d photons, 8 detectors x %d parameters (%d regions x "
                    "mua, mus) + %d^3 voxel mua and mus Jacobians; central differences with common random numbers; "
                    "standard errors from %d batches" % (B * n, P, P // 2, jac_n, B),
        "parameters": P, "photons": B * n,
        "adjoint_s": t_adj, "forward_pass_s": t_fwd, "adjoint_over_forward": t_adj / t_fwd,
        "forward_kernel_ms": fwd_kernel_ms, "adjoint_forward_ms": full_fwd_ms, "adjoint_replay_ms": full_rep_ms,
        "replayed_pairs": pairs, "jacobian_nonzero_mua": jac_nz, "jacobian_nonzero_mus": jac_mus_nz,
        "adjoint_batches_s": t_batches,
        "adjoint_resolved_entries": int(res.sum()), "entries": nd * P,
        "fd_forward_passes": 2 * P, "fd_forward_passes_with_jacobians": 2 * P + 4 * jac_n ** 3,
        "fd_s": 2 * P * t_fwd, "fd_over_adjoint_time": 2 * P * t_fwd / t_adj, "fd": [],
    }
    t_fd = 2 * P * t_fwd  # 2 forward passes of all photons per parameter
    # The differences themselves are evaluated in batched launches (every
    # perturbed optics set of a photon batch in one launch: instances share
    # the per-photon random streams, so each equals its own lp_run -- common
    # random numbers); their cost above is that of separate forward passes.
    tab = []
    for rel in rels:
        for t, c in params:
            h = float(nm0[t, c]) * rel
            for sgn in (-1, 1):
                nm = nm0.copy()
                nm[t, c] = np.float32(nm0[t, c] + sgn * h)
                tab.append(nm)
    tab = np.stack(tab)
    t0 = time.perf_counter()
    Sb = np.zeros((B, len(tab), nd))
    for b in range(B):
        Sb[b] = ctx.run_batch(tab, b * n, n)["det"][:, :, 0] / lp.FIXED / n
    print("finite differences: %d batched launches of %d optics sets x %d photons, %.1f s" %
          (B, len(tab), n, time.perf_counter() - t0), file=sys.stderr, flush=True)
    for ir, rel in enumerate(rels):
        fd = np.zeros((B, nd, P))
        for i, (t, c) in enumerate(params):
            h = float(nm0[t, c]) * rel
            dh = float(np.float32(nm0[t, c] + h)) - float(np.float32(nm0[t, c] - h))
            k = 2 * (ir * P + i)
            fd[:, :, i] = (Sb[:, k + 1] - Sb[:, k]) / dh
        mf = fd.mean(0)
        sf = fd.std(0, ddof=1) / math.sqrt(B)
        ok = res & (sf > 0)
        speed = (t_fd * sf ** 2) / (t_adj * np.maximum(sa, 1e-300) ** 2)
        agree = np.abs(ma - mf) <= 4 * np.sqrt(sa ** 2 + sf ** 2) + 0.05 * np.abs(ma)

        def med(mask):
            return float(np.median(speed[mask])) if mask.any() else None
        out["fd"].append({
            "step_rel": rel,
            "fd_resolved_entries": int((np.abs(mf) > 3 * sf).sum()),
            "agree_within_4se_5pct": int((agree & res).sum()), "compared_entries": int(ok.sum()),
            "speedup_equal_error_median": med(ok), "speedup_equal_error_median_mua": med(ok & (kinds == 1)),
            "speedup_equal_error_median_mus": med(ok & (kinds == 2)),
            "speedup_equal_error_min": float(speed[ok].min()) if ok.any() else None,
        })
    ctx.close()
    return out


# Batched workload: 4 geometries (fetal head depth 1.5 / 2.5 / 3.5 / 4.5 cm;
# maternal / fetal icosphere frequencies for 56 480 / 196 420 / 272 400 /
# 377 120 faces; the 1.5 cm geometry has a 0.3 cm fat layer and a 0.5 cm
# uterine wall so that the head fits under them) x 2 wavelengths x 16 fetal
# SpO2 in [0.30, 0.75] (maternal 0.97) = 128 instances: all in one batched
# launch over the merged geometries, and for comparison one launch of 32
# instances per geometry.
C2_GEOMS = [(1.5, 19, 14, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, 38, 27, {}), (3.5, 45, 32, {}), (4.5, 53, 38, {})]


def batch_c2(photons=1_000_000, tune=False, per_geometry=False):
    """Photons/s of the batched workload: all 128 instances in one launch
    (batch_c2_one_launch); with per_geometry, also as 4 launches of 32
    instances (lp_run_batch through its captured HIP graph; detector, region
    and counter tallies), each autotuned for its geometry."""
    merged = batch_c2_one_launch(photons, tune)
    out = {"workload": "4 geometries x 2 wavelengths x 16 fetal SpO2 = 128 instances, %d photons each%s"
                       % (photons, "" if photons >= 10_000_000 else " (reduced from 10^7 to keep the run short)"),
           "one_launch": merged}
    if per_geometry:
        out["per_geometry_launches"] = batch_c2_per_geometry(photons)
    return out


def batch_c2_per_geometry(photons, tune=True):
    import lightpropagation_amd as lp
    from lightpropagation_amd import scenes
    spo2 = np.linspace(0.30, 0.75, 16)
    lo, hi = detectors()
    out, total_ph, total_s = [], 0, 0.0
    for depth, fm, ff, kw in C2_GEOMS:
        sc = scenes.maternal_fetal(fetal_depth=depth, freq_maternal=fm, freq_fetal=ff, **kw)
        tab = np.stack([scenes.mf_optics(wl, 0.97, sf) for wl in (660.0, 940.0) for sf in spo2])
        ctx = lp.Context(0)
        ctx.set_scene(sc, seed=0x5EED0002)
        ctx.set_detectors(lo, hi)
        ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
        tuned = None
        if tune:
            r = ctx.autotune(sample=2_000_000, passes=2)
            tuned = {"trials": r.trials, "ms_before": r.ms_before, "ms_after": r.ms_after, "waves": r.waves,
                     "blocks_per_cu": r.blocks_per_cu, "free_iters": r.free_iters, "refill_min": r.refill_min,
                     "list_min": r.list_min, "search_min": r.search_min}
        ctx.run_batch(tab, 1 << 40, max(photons // 100, 1000))  # capture + warm-up
        t0 = time.perf_counter()
        b = ctx.run_batch(tab, 0, photons)
        wall = time.perf_counter() - t0
        kms = ctx.stats().kernel_ms
        ctx.close()
        n = len(tab) * photons
        total_ph += n
        total_s += wall
        R = b["det"][:, :, 0].sum(1) / lp.FIXED / photons
        out.append({"fetal_depth": depth, "faces": len(sc["tri"]), "instances": len(tab), "photons": n,
                    "kernel_ms": kms, "wall_s": wall, "photons_per_s": n / wall, "autotune": tuned,
                    "detected_R_660_range": [float(R[:16].min()), float(R[:16].max())]})
        print("batch: depth %.1f, %d faces, %d x %d photons, %.0f ms" % (depth, len(sc["tri"]), len(tab), photons, kms),
              file=sys.stderr, flush=True)
    return {"photons_per_s": total_ph / total_s, "photons": total_ph, "seconds": total_s, "geometries": out}


def batch_c2_one_launch(photons=1_000_000, tune=False):
    """The batched workload as ONE persistent, graph-captured launch: the 4
    geometries merged into one mesh (lp_set_geometries: each keeps its own
    BVH and near-face grid), the 128 instances each naming its geometry
    (lp_run_batch_geom).  Photons/s over the whole launch."""
    import lightpropagation_amd as lp
    from lightpropagation_amd import scenes
    spo2 = np.linspace(0.30, 0.75, 16)
    lo, hi = detectors()
    scs = [scenes.maternal_fetal(fetal_depth=d, freq_maternal=fm, freq_fetal=ff, **kw) for d, fm, ff, kw in C2_GEOMS]
    merged, counts = scenes.geometries(scs)
    one = np.stack([scenes.mf_optics(wl, 0.97, sf) for wl in (660.0, 940.0) for sf in spo2])
    tab = np.concatenate([one] * len(scs))
    geom = np.repeat(np.arange(len(scs), dtype=np.int32), len(one))
    ctx = lp.Context(0)
    t0 = time.perf_counter()
    ctx.set_scene(merged, seed=0x5EED0002)
    ctx.set_geometries(counts)
    ctx.set_detectors(lo, hi)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    ctx.run_batch(tab[:1], 1 << 40, 1000)  # builds the 4 BVHs and near grids, uploads
    setup_s = time.perf_counter() - t0
    tuned = None
    if tune:  # trial launches transport in geometry 0 (default: the library's default options)
        r = ctx.autotune(sample=2_000_000, passes=2)
        tuned = {"trials": r.trials, "ms_before": r.ms_before, "ms_after": r.ms_after, "waves": r.waves,
                 "blocks_per_cu": r.blocks_per_cu, "free_iters": r.free_iters, "refill_min": r.refill_min,
                 "list_min": r.list_min, "search_min": r.search_min}
    ctx.run_batch(tab, 1 << 40, max(photons // 100, 1000), geom=geom)  # capture + warm-up
    t0 = time.perf_counter()
    b = ctx.run_batch(tab, 0, photons, geom=geom)
    wall = time.perf_counter() - t0
    kms = ctx.stats().kernel_ms
    ctx.close()
    n = len(tab) * photons
    R = b["det"][:, :, 0].sum(1) / lp.FIXED / photons
    print("batch, one launch: %d faces in %d geometries, %d x %d photons, %.0f ms" %
          (sum(counts), len(counts), len(tab), photons, kms), file=sys.stderr, flush=True)
    return {"faces": counts, "instances": len(tab), "photons": n, "kernel_ms": kms, "wall_s": wall,
            "photons_per_s": n / wall, "setup_s": setup_s, "autotune": tuned,
            "detected_R_660_range_per_geometry": [[float(R[k * 32:k * 32 + 16].min()),
                                                   float(R[k * 32:k * 32 + 16].max())] for k in range(len(counts))]}


def large_c4(photons=100_000_000):
    """The large workload: frequency-118 maternal and frequency-92 fetal
    head shells plus a frequency-40 fetal torso ellipsoid (1 992 320 faces,
    10 regions), 940 nm, a 512^3 int64 absorption grid (1 GiB) and 8
    detectors; photons/s of the forward pass, then the adjoint (forward +
    replay, 10 regions x (mua, mus)) of the same photons."""
    import lightpropagation_amd as lp
    from lightpropagation_amd import scenes
    sc = scenes.maternal_fetal(wavelength_nm=940.0, freq_maternal=118, freq_fetal=92, freq_torso=40)
    ctx = lp.Context(0)
    t0 = time.perf_counter()
    ctx.set_scene(sc, seed=0x5EED0004)
    lo, hi = detectors()
    ctx.set_detectors(lo, hi)
    ctx.set_voxel_grid([-12.8, -12.8, 0.0], [0.05, 0.05, 0.05], [512, 512, 512])
    ctx.set_tallies(sum(1 << lp.TALLY[k] for k in ("sft", "region", "vox", "det", "misc")))
    ctx.run(1 << 40, 100_000)  # builds the BVH and near-face lists, uploads
    setup_s = time.perf_counter() - t0
    ctx.reset()
    t0 = time.perf_counter()
    ctx.run(0, photons)
    fwd_s = time.perf_counter() - t0
    st = ctx.stats()
    m = ctx.tallies()["misc"]
    print("large: %d faces, %d photons in %.1f s" % (len(sc["tri"]), photons, fwd_s), file=sys.stderr, flush=True)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
    ctx.reset()
    t0 = time.perf_counter()
    r = ctx.adjoint(0, photons)
    adj_s = time.perf_counter() - t0
    sa = ctx.stats()
    ctx.close()
    return {"workload": "1 992 320 faces (shells + fetal torso), 940 nm, 512^3 int64 grid + 8 detectors, %d photons%s"
                        % (photons, "" if photons >= 1_000_000_000 else " (reduced from 10^9 to keep the run short)"),
            "setup_s": setup_s, "photons_per_s": photons / fwd_s, "kernel_ms": st.kernel_ms,
            "steps_per_photon": m["steps"] / photons, "device_bytes": st.device_bytes,
            "adjoint_s": adj_s, "adjoint_forward_ms": sa.kernel_ms, "adjoint_replay_ms": sa.replay_ms,
            "adjoint_pairs": r["pairs"], "adjoint_over_forward": adj_s / fwd_s}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=4)
    ap.add_argument("--warmup", type=int, default=2)
    ap.add_argument("--photons", type=int, default=100_000_000, help="photons per step (10^8: one wavelength of the "
                                                                     "headline run)")
    ap.add_argument("--wavelength", type=float, default=0.0, help="one wavelength only (0: alternate 660 / 940 nm)")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=20.0)
    ap.add_argument("--grid", type=int, default=0)
    ap.add_argument("--lds-nodes", type=int, default=-1)
    ap.add_argument("--safety-cells", type=int, default=-1)
    ap.add_argument("--opt", action="append", default=[], help="tuning option key=value (lp_set_option)")
    ap.add_argument("--tune", type=int, default=0, metavar="PHOTONS",
                    help="run the autotuner (trials of this many photons, with the step's tallies) before the "
                         "warm-up, untimed")
    ap.add_argument("--no-adjoint", action="store_true", help="skip the adjoint vs finite-difference measurement")
    ap.add_argument("--adjoint-photons", type=int, default=10_000_000)
    ap.add_argument("--c2", type=int, default=1_000_000, metavar="PHOTONS",
                    help="photons per instance of the batched 128-instance workload (10^7 in its definition; "
                         "0: skip)")
    ap.add_argument("--c2-per-geometry", action="store_true",
                    help="also run the batched workload as 4 launches, one per geometry, each autotuned")
    ap.add_argument("--c4", type=int, default=100_000_000, metavar="PHOTONS",
                    help="photons of the large 2e6-face workload (10^9 in its definition; 0: skip)")
    ap.add_argument("--fit", action="store_true",
                    help="also fit fetal SpO2 to synthetic two-wavelength data (torch, in a child process)")
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

    wls = [args.wavelength] if args.wavelength else [660.0, 940.0]
    ctxs, T = {}, 0
    for wl in wls:
        sc = scenes.maternal_fetal(wavelength_nm=wl)
        T = len(sc["tri"])
        ctx = lp.Context(0)
        setup(ctx, sc, seed=0x5EED0001)
        ctx.set_launch(args.grid, args.lds_nodes, args.safety_cells)
        for kv in args.opt:
            k, v = kv.split("=")
            ctx.set_option(k, float(v))
        ctxs[wl] = (ctx, sc)
    # lp_run blocks on a HIP event recorded after the kernel on the context's
    # stream, so the timed region is bracketed by completed device work.  torch
    # is not initialised on the GPU here: it ships its own HIP/HSA runtime, and
    # a second runtime on the device in one process is not supported.
    print("setup done: %d faces, wavelengths %s" % (T, wls), file=sys.stderr, flush=True)
    tuned = None
    if args.tune:
        r = ctxs[wls[0]][0].autotune(sample=args.tune, passes=2)
        tuned = {"sample": args.tune, "trials": r.trials, "ms_before": r.ms_before, "ms_after": r.ms_after,
                 "waves": r.waves, "blocks_per_cu": r.blocks_per_cu, "free_iters": r.free_iters,
                 "refill_min": r.refill_min, "list_min": r.list_min, "search_min": r.search_min}
        print("autotune: %s" % tuned, file=sys.stderr, flush=True)
        for wl in wls[1:]:
            c = ctxs[wl][0]
            for k, v in (("waves_per_simd", r.waves), ("free_iters", r.free_iters), ("refill_min", r.refill_min),
                         ("list_min", r.list_min), ("search_min", r.search_min)):
                c.set_option(k, v)

    first = {wl: 0 for wl in wls}
    for i in range(args.warmup):
        wl = wls[i % len(wls)]
        ctxs[wl][0].run(first[wl], args.photons)
        first[wl] += args.photons
        print("warmup step done (%g nm)" % wl, file=sys.stderr, flush=True)
    for wl in wls:
        ctxs[wl][0].reset()
    per = {wl: {"kms": [], "photons": 0} for wl in wls}
    t0 = time.perf_counter()
    for i in range(args.steps):
        wl = wls[i % len(wls)]
        ctx = ctxs[wl][0]
        ctx.run(first[wl], args.photons)
        first[wl] += args.photons
        per[wl]["kms"].append(ctx.stats().kernel_ms)
        per[wl]["photons"] += args.photons
        print("step %d (%g nm): %.1f ms kernel" % (i, wl, per[wl]["kms"][-1]), file=sys.stderr, flush=True)
    wall = time.perf_counter() - t0
    nph = args.steps * args.photons
    search = F_TRI + F_BOX * math.ceil(math.log2(max(T, 2)))
    flops = kernel_s = steps = 0.0
    counters, physics, by_wl = {}, {}, {}
    st = None
    for wl in wls:
        ctx, sc = ctxs[wl]
        st = ctx.stats()
        t = ctx.tallies()
        m = t["misc"]
        n = per[wl]["photons"]
        ks = sum(per[wl]["kms"]) / 1e3
        kernel_s += ks
        steps += m["steps"]
        flops += m["interact"] * (F_INT + search) + m["boundary"] * (F_BND + search)
        for k in ("interact", "boundary", "escaped", "capped", "nan", "node_visits", "tri_tests", "skipped",
                  "dist_queries", "dist_visits", "list_tests", "wave_iters", "search_rounds", "bvh_rounds",
                  "marches", "crawl_replayed"):
            counters[k] = counters.get(k, 0) + m[k]
        if n:
            R = t["sft"].reshape(-1, 3).sum(0) / lp.FIXED / n
            det = t["det"].reshape(-1, 2)[:, 0] / lp.FIXED / n
            physics["%g nm" % wl] = {"R_total": float(R[0]), "F_total": float(R[1]), "T_total": float(R[2]),
                                     "detector_R": [float(x) for x in det]}
            by_wl["%g nm" % wl] = {"photons": n, "kernel_ms_per_step": 1e3 * ks / len(per[wl]["kms"]),
                                   "photons_per_s_kernel": n / ks, "steps_per_photon": m["steps"] / n}
    achieved = flops / kernel_s / 1e12 if kernel_s > 0 else 0.0
    tpp = traffic_per_photon()
    out = {
        "metric": METRIC,
        "value": nph / wall if wall > 0 and nph else 0.0,
        "unit": "photons/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": wall * 1e3 / max(args.steps, 1),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": "maternal-fetal mesh, 10 regions, %d faces, %s nm alternating step by step, SpO2 0.97/0.50, "
                        "%d photons per step, 8 detectors + 256^3 int64 grid + region/radial/SFT/RTd tallies"
                        % (T, "/".join("%g" % w for w in wls), args.photons),
            "faces": T, "photons_per_step": args.photons, "wavelengths_nm": wls, "rng": "philox4x32-10",
        },
        "by_wavelength": by_wl,
        "photon_steps_per_s": steps / wall if nph else 0.0,
        "photons_per_min": nph / wall * 60 if nph else 0.0,
        "kernel_ms_per_step": 1e3 * kernel_s / max(args.steps, 1),
        "autotune": tuned,
        "steps_per_photon": steps / nph if nph else 0.0,
        "roofline": {"bound": "valu", "achieved": achieved, "peak": FP32_PEAK_TFLOPS, "unit": "TFLOP/s",
                     "frac": achieved / FP32_PEAK_TFLOPS,
                     "traffic": (tpp * args.photons) if tpp is not None else None,
                     "traffic_unit": "bytes per 660 nm launch (rocprofv3 FETCH_SIZE x2 + WRITE_SIZE x1, copy-calibrated)",
                     "work": "algorithmic fp32 per step: interaction %g + boundary %g, search %g (T=%d)" %
                             (F_INT, F_BND, search, T)},
        "counters": counters,
        "options": args.opt,
        "physics": physics,
        "grid_blocks": st.grid_blocks, "bvh_nodes": st.bvh_nodes, "bvh_depth": st.bvh_depth,
    }
    if args.calibrate_copy:
        ms = ctxs[wls[0]][0].debug_copy(args.calibrate_copy << 20)
        out["copy_calibration"] = {"bytes": args.calibrate_copy << 20, "ms": ms,
                                   "GB_per_s": 2 * (args.calibrate_copy << 20) / ms / 1e6}
    sc660 = ctxs[wls[0]][1]
    for ctx, _ in ctxs.values():
        ctx.close()
    if not args.no_adjoint:
        out["adjoint"] = adjoint_vs_fd(sc660, photons=args.adjoint_photons)
    if args.c2:
        out["batch_c2"] = batch_c2(args.c2, per_geometry=args.c2_per_geometry)
    if args.c4:
        out["large_c4"] = large_c4(args.c4)
    if args.fit:
        # torch and liblp.so bring two HIP runtimes: the fit runs in a process of its own
        import subprocess
        p = subprocess.run([sys.executable, "-m", "lightpropagation_amd.spo2_fit"], cwd=ROOT, stdout=subprocess.PIPE,
                           text=True, timeout=1800)
        out["spo2_fit"] = json.loads(p.stdout.strip().splitlines()[-1]) if p.returncode == 0 else {"error": p.returncode}
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(sc660, args.cpu_seconds)
    print(json.dumps(out))


if __name__ == "__main__":
    main()

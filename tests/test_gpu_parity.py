"""GPU parity: the HIP path through the C ABI against the CPU oracle.

Bit-exact, on the same seeded photons:
  * the kernel's math core (sqrt, divide vs IEEE float32; log, cos, acos vs
    the oracle's implementations of the same series),
  * per-photon step traces (event kind, face, tissue, w, p, v, step length,
    random draws consumed) in both RNG modes,
  * per-photon exit records and every tally (SFT, RTd, Arz, Phi_rz, per
    region, voxels, radial bins, detectors, counters),
  * determinism: repeated and split GPU runs give identical tallies.
Scenes: the reference fixtures model1 (MCML slab), model2 (mismatched n,
isotropic), model6/configFB (thick slab, n = 1.37), model4 (five concentric
icospheres, 25 600 faces) and the three-layer slab of the bench config 0.
"""
import math

import numpy as np
import pytest

import lightpropagation_amd as lp
import _oracle as O
import scenes

pytestmark = pytest.mark.gpu

TALLIES = dict(vox=dict(origin=[-1.0, -1.0, 0.0], size=[0.02, 0.02, 0.01], n=[100, 100, 120]),
               rad=dict(dr=0.05, n=200), det_rlo=[0.0, 0.2, 0.5, 1.0], det_rhi=[0.2, 0.5, 1.0, 3.0])
MAXSTEPS = 200000


@pytest.fixture(scope="module")
def ctx():
    c = lp.Context(0)
    yield c
    c.close()


def _scene(name):
    if name == "c0slab":
        return scenes.slab(scenes.C0_LAYERS)
    if name == "c0slab_g1":  # g = 1 - 2^-24 in the middle layer (forward scattering at the float limit)
        return scenes.slab([L if i != 1 else L[:4] + (0.99999994,) for i, L in enumerate(scenes.C0_LAYERS)])
    if name == "mf_coarse":  # maternal-fetal abdomen, 5 040 faces
        from lightpropagation_amd import scenes as S
        return S.maternal_fetal(freq_maternal=6, freq_fetal=6, freq_torso=4)
    if name == "model6FB":
        return scenes.load_fixture("model6", "configFB.xml")
    if name == "planar":  # model4/configPlanar.xml: 20 planar faces, 160 x 160 cm
        return scenes.load_fixture("model4", "configPlanar.xml")
    if name == "escape":  # model3 with a second start direction that misses the mesh
        return scenes.escape_scene(ESCAPE_UP)
    if name == "oblique":  # oblique photon through a slab at mismatched n
        return scenes.oblique_slab()
    return scenes.load_fixture(name)


ESCAPE_UP = 150  # photons [0, 150) head into the mesh, the rest away from it


def _configure(ctx, sc, seed, rng):
    ctx.set_scene(sc, seed=seed, rng=rng, max_steps=MAXSTEPS)
    ctx.set_voxel_grid(TALLIES["vox"]["origin"], TALLIES["vox"]["size"], TALLIES["vox"]["n"])
    ctx.set_radial(TALLIES["rad"]["dr"], TALLIES["rad"]["n"])
    ctx.set_detectors(TALLIES["det_rlo"], TALLIES["det_rhi"])
    ctx.set_tallies(record_exit=True)
    ctx.reset()


def test_math_core_bit_exact(ctx):
    rng = np.random.default_rng(0)
    x = np.concatenate([rng.random(100000), np.exp(rng.uniform(-90, 40, 50000))]).astype(np.float32)
    y = np.exp(rng.uniform(-20, 20, len(x))).astype(np.float32)
    assert np.array_equal(ctx.debug_math(0, x).view(np.uint32), np.sqrt(x).view(np.uint32))
    assert np.array_equal(ctx.debug_math(1, x, y).view(np.uint32), (x / y).view(np.uint32))
    xs = x[:20000]
    assert np.array_equal(ctx.debug_math(2, xs).view(np.uint32),
                          np.array([O.logf(v) for v in xs], np.float32).view(np.uint32))
    ps = (np.float32(6.28318548) * rng.random(20000).astype(np.float32)).astype(np.float32)
    assert np.array_equal(ctx.debug_math(3, ps).view(np.uint32),
                          np.array([O.cosf(v) for v in ps], np.float32).view(np.uint32))
    cs = rng.uniform(-1, 1, 20000).astype(np.float32)
    assert np.array_equal(ctx.debug_math(4, cs).view(np.uint32),
                          np.array([O.acosf(v) for v in cs], np.float32).view(np.uint32))


def test_scattering_sqrt_correctly_rounded(ctx):
    # op 7 (the scattering path's square root) is correctly rounded on its
    # domain: 0, and arguments from 2^-96 up (1 - c^2 of floats c in [-1, 1],
    # squared norms near 1), including every float in [0.5, 4)
    rng = np.random.default_rng(3)
    c = rng.uniform(-1, 1, 200000).astype(np.float32)
    x = np.concatenate([
        (np.float32(1) - c * c).astype(np.float32),
        np.exp(rng.uniform(-66, 88, 200000)).astype(np.float32),
        np.arange(0x3f000000, 0x40800000, 7, dtype=np.uint32).view(np.float32),
        np.array([0.0, 2.0 ** -96, 1.0, 2.0 ** -25, 3.4e38, np.inf], np.float32),
    ])
    x = x[(x == 0) | (x >= np.float32(2.0 ** -96))]
    got = ctx.debug_math(7, x)
    assert np.array_equal(got.view(np.uint32), np.sqrt(x).view(np.uint32))


def test_normalize_reciprocal_correctly_rounded(ctx):
    # op 8 (1 / y of normalize on the scattering path) equals the correctly
    # rounded 1 / y on y in [2^-60, 2^60], densely on [0.5, 4)
    rng = np.random.default_rng(4)
    y = np.concatenate([
        np.exp2(rng.uniform(-60, 60, 300000)).astype(np.float32),
        np.arange(0x3f000000, 0x40800000, 5, dtype=np.uint32).view(np.float32),
        np.array([2.0 ** -60, 2.0 ** 60, 1.0, 0.75, 3.0], np.float32),
    ])
    got = ctx.debug_math(8, y)
    assert np.array_equal(got.view(np.uint32), (np.float32(1) / y).view(np.uint32))


def test_step_log_equals_general_log(ctx):
    # op 9 (the step-length log, without special cases) equals op 2 on the
    # draws it meets: Philox words (k 2^-24, k >= 1) and ran3 values r / 1e9
    rng = np.random.default_rng(5)
    k = np.concatenate([np.arange(1, 5000), rng.integers(1, 1 << 24, 200000)]).astype(np.float32)
    r = np.concatenate([np.arange(1, 3000), rng.integers(1, 10 ** 9, 100000)]).astype(np.float32)
    x = np.concatenate([k * np.float32(5.96046448e-08), r / np.float32(1e9)]).astype(np.float32)
    assert np.array_equal(ctx.debug_math(9, x).view(np.uint32), ctx.debug_math(2, x).view(np.uint32))


CASES = [("model1", 512), ("model2", 512), ("model6FB", 256), ("c0slab", 256), ("model4", 48), ("model3", 256),
         ("planar", 128), ("escape", 256), ("oblique", 256), ("c0slab_g1", 256)]


@pytest.mark.parametrize("rng", [lp.RNG_PHILOX, lp.RNG_RAN3], ids=["philox", "ran3"])
@pytest.mark.parametrize("name,n", CASES, ids=[c[0] for c in CASES])
def test_traces_bit_exact(ctx, name, n, rng):
    sc = _scene(name)
    maxrec = 3000
    _configure(ctx, sc, 4242, rng)
    g, gn = ctx.trace(100, n, maxrec)
    orc = O.Run(sc, seed=4242, rng_mode=rng, max_steps=MAXSTEPS, tallies=TALLIES)
    _, o, on = orc.run(100, n, trace_max=maxrec, nthreads=8)
    assert np.array_equal(gn, on), np.nonzero(gn != on)[0][:5]
    bad = []
    for i in range(n):
        k = on[i]
        if not np.array_equal(g[i, :k].view(np.uint32), o[i, :k].view(np.uint32)):
            j = int(np.nonzero(np.any(g[i, :k].view(np.uint32) != o[i, :k].view(np.uint32), axis=1))[0][0])
            bad.append((i, j, g[i, j].tolist(), o[i, j].tolist()))
    assert not bad, bad[:3]


TALLY_CASES = [("model1", 20000), ("model2", 8000), ("c0slab", 3000), ("model4", 96), ("mf_coarse", 1500),
               ("model3", 4000), ("planar", 1500), ("escape", 3000), ("oblique", 4000), ("c0slab_g1", 3000)]


@pytest.mark.parametrize("rng", [lp.RNG_PHILOX, lp.RNG_RAN3], ids=["philox", "ran3"])
@pytest.mark.parametrize("name,n", TALLY_CASES, ids=[c[0] for c in TALLY_CASES])
def test_tallies_bit_exact(ctx, name, n, rng):
    sc = _scene(name)
    _configure(ctx, sc, 99, rng)
    ctx.run(0, n)
    ex = ctx.exit_records(n)
    gt = ctx.tallies()
    orc = O.Run(sc, seed=99, rng_mode=rng, max_steps=MAXSTEPS, tallies=TALLIES)
    oex, _, _ = orc.run(0, n, exit_rec=True, nthreads=8)
    ot = orc.tallies()
    assert np.array_equal(ex.view(np.uint32), oex.view(np.uint32))
    for k in ("sft", "rtd", "arz", "phirz", "region", "vox", "rad", "det"):
        a, b = gt[k], ot[k][:len(gt[k])]
        assert np.array_equal(a, b), k
    for k in ("photons", "steps", "interact", "boundary", "escaped", "capped", "arz_out", "vox_out", "nan"):
        assert gt["misc"][k] == ot["misc"][k], (k, gt["misc"][k], ot["misc"][k])
    if name == "escape":  # the photons heading away found no face ahead and left
        assert gt["misc"]["escaped"] == n - ESCAPE_UP
    # with every search shortcut off (each step searched), the same results
    if name in ("model4", "mf_coarse", "escape"):
        ctx.set_option("no_skip", 1)
        ctx.reset()
        ctx.run(0, n)
        assert np.array_equal(ctx.exit_records(n).view(np.uint32), oex.view(np.uint32))
        g2 = ctx.tallies()
        ctx.set_option("no_skip", 0)
        assert g2["misc"]["skipped"] == 0
        for k in ("sft", "rtd", "arz", "phirz", "region", "vox", "rad", "det"):
            assert np.array_equal(g2[k], gt[k]), k


def test_gpu_runs_are_reproducible_and_resumable(ctx):
    sc = _scene("c0slab")
    _configure(ctx, sc, 5, lp.RNG_PHILOX)
    ctx.set_tallies(record_exit=False)
    ctx.run(0, 200000)
    a = ctx.tallies()
    ctx.reset()
    ctx.run(0, 70000)
    ctx.run(70000, 130000)
    b = ctx.tallies()
    for k in ("sft", "rtd", "arz", "phirz", "region", "vox", "rad", "det"):
        assert np.array_equal(a[k], b[k]), k
    assert a["misc"]["steps"] == b["misc"]["steps"]


def test_mcml_slab_on_gpu(ctx):
    # 10^6 photons of the MCML verification slab: Rd, Tt within 3 sigma of
    # the published values (see tests/test_oracle_physics.py for sources)
    sc = _scene("model1")
    _configure(ctx, sc, 12345, lp.RNG_PHILOX)
    ctx.set_tallies(record_exit=True)
    n = 1000000
    ctx.run(0, n)
    ex = ctx.exit_records(n).astype(np.float64)
    R, T = ex[:, 3], ex[:, 5]
    sR = math.hypot(R.std() / math.sqrt(n), 0.00035)
    sT = math.hypot(T.std() / math.sqrt(n), 0.00020)
    assert abs(R.mean() - 0.09734) < 3 * sR
    assert abs(T.mean() - 0.66096) < 3 * sT
    t = ctx.tallies()
    A = t["region"].sum() / lp.FIXED / n
    assert abs(R.mean() + T.mean() + A - 1.0) < 1e-5
    # tallies agree with the per-photon records
    assert abs(t["sft"].reshape(-1, 3)[:, 0].sum() / lp.FIXED - R.sum()) < 1e-3


def test_bench_scene_photons_that_turned_nan():
    # Bench workload (202 160 faces, seed 0x5EED0001): photon 265 841 met a
    # face at |cos| = 1.00000012 and its direction became NaN in the
    # reference's Transmit (sqrt of 1 - cos^2 < 0); NaN rays then opened
    # every BVH box and the launch ran for minutes.  The range around it must
    # now finish quickly with no NaN photon and no capped photon.  (The
    # seven-region mesh of that workload.)
    from lightpropagation_amd import scenes as S
    sc = S.maternal_fetal(regions=7)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=0x5EED0001, max_steps=200000)
    ctx.set_tallies(record_exit=True)
    ctx.run(265000, 2000)
    m = ctx.tallies()["misc"]
    assert m["nan"] == 0 and m["capped"] == 0 and m["photons"] == 2000
    ex = ctx.exit_records(2000)
    assert np.all(np.isfinite(ex[:, :6]))
    assert ctx.stats().kernel_ms < 20000
    ctx.close()


def test_misc_counters_fit_their_tally(ctx):
    sc = _scene("c0slab")
    _configure(ctx, sc, 1, lp.RNG_PHILOX)
    assert len(lp.MISC) <= ctx.layout().n[lp.TALLY["misc"]]
    ctx.run(0, 1000)
    m = ctx.tallies()["misc"]
    assert m["photons"] == 1000 and m["wave_iters"] > 0 and m["search_rounds"] > 0


PHYSICS_MISC = ["photons", "steps", "interact", "boundary", "escaped", "capped", "arz_out", "vox_out", "nan"]


def test_tuning_options_do_not_change_results():
    # Every option of lp_set_option only changes how the kernel finds faces
    # (free regions, near-face lists, BVH in LDS, search scheduling, register
    # budget, steps per iteration, crawl replay), never which face a step meets: 10^6
    # photons of the bench workload (196 420 faces) give bit-identical
    # tallies under each setting.
    from lightpropagation_amd import scenes as S
    sc = S.maternal_fetal()
    settings = [{}, {"near_r0": 0.0}, {"near_r0": 0.03}, {"near_r0": 0.08, "near_cell": 0.1},
                {"free_iters": 4, "search_min": 0.5}, {"waves_per_simd": 1, "lds_nodes": 0},
                {"dist_cap": 0.0}, {"clear_step": -1.0, "refill_min": 4}, {"clear_step": 0.05},
                {"list_long": 0}, {"list_long": 0, "near_r0": 0.03, "dist_cap": 0.0}, {"no_skip": 1},
                {"no_skip": 1, "near_r0": 0.0}, {"second_step": 0}, {"second_step": 1, "list_min": 0.3},
                {"second_step": 8, "waves_per_simd": 5}, {"crawl": 0}]
    ref = None
    for opts in settings:
        ctx = lp.Context(0)
        for k, v in opts.items():
            ctx.set_option(k, v)
        ctx.set_scene(sc, seed=0x5EED0001, max_steps=200000)
        ctx.set_voxel_grid([-6.4, -6.4, 0.0], [0.05, 0.05, 0.05], [256, 256, 256])
        ctx.set_radial(0.05, 200)
        ctx.reset()
        print("options", opts, flush=True)
        ctx.run(0, 1_000_000)
        t = ctx.tallies()
        ctx.close()
        got = {k: np.asarray(v).copy() for k, v in t.items() if k != "misc"}
        got["misc"] = np.array([t["misc"][k] for k in PHYSICS_MISC])
        if ref is None:
            ref = got
            continue
        for k in ref:
            assert np.array_equal(ref[k], got[k]), (opts, k)


def test_face_met_at_exactly_the_step_length():
    # Photons of the bench workload whose search meets a face at t equal to
    # the step length: the face must come with its own tissue pair.  (Three
    # separate conditional updates of (t, id, med) compiled so that a face
    # taken through the t == best branch kept the previous med; the search
    # now keeps (t, id) as one key.)  Exit records against the oracle's,
    # committed as tests/golden/tie_photons.json, with the near-face lists
    # off (BVH search) and on.
    import json
    import os
    from lightpropagation_amd import scenes as S
    import bench
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "tie_photons.json")))
    sc = S.maternal_fetal(regions=7)  # the seven-region mesh the ties were found on
    for opts in ({"near_r0": 0.0}, {}):
        ctx = lp.Context(0)
        for k, v in opts.items():
            ctx.set_option(k, v)
        bench.setup(ctx, sc, seed=g["seed"])
        ctx.set_tallies(record_exit=True)
        for pid, bits in g["exit_bits"].items():
            print("options", opts, "photon", pid, flush=True)
            ctx.run(int(pid), 1)
            ex = ctx.exit_records(1)[0]
            assert ex.view(np.uint32).tolist() == bits, (opts, pid, ex.tolist())
        ctx.close()


@pytest.mark.parametrize("opts", [{}, {"near_r0": 0.0}, {"no_skip": 1}], ids=["default", "bvh_only", "no_skip"])
def test_headline_mesh_against_oracle_golden(opts):
    # 2048 photons of the headline workload (the ten-region maternal-fetal
    # mesh, 196 420 faces, 660 nm, bench.setup tallies) against the oracle's
    # results committed by tests/golden/make_c1_golden.py: every exit
    # record bit for bit, every nonzero of the SFT, RTd, region, radial and
    # detector tallies, the voxel tally's hash, and the physics counters --
    # with the default search shortcuts, with the BVH alone, and with every
    # step searched
    import hashlib
    import json
    import os
    from lightpropagation_amd import scenes as S
    import bench
    g = json.load(open(os.path.join(os.path.dirname(__file__), "golden", "c1_photons.json")))
    sc = S.maternal_fetal()
    assert len(sc["tri"]) == g["faces"]
    ctx = lp.Context(0)
    for k, v in opts.items():
        ctx.set_option(k, v)
    bench.setup(ctx, sc, seed=g["seed"])
    ctx.set_tallies(sum(1 << lp.TALLY[k] for k in ("sft", "rtd", "region", "vox", "rad", "det", "misc")),
                    record_exit=True)
    ctx.reset()
    print("options", opts, flush=True)
    ctx.run(g["first"], g["count"])
    ex = ctx.exit_records(g["count"])
    t = ctx.tallies()
    ctx.close()
    want = np.array(g["exit_bits"], np.uint32)
    bad = np.nonzero(np.any(ex.view(np.uint32) != want, axis=1))[0]
    assert len(bad) == 0, (len(bad), bad[:5])
    for k in ("sft", "rtd", "region", "rad", "det"):
        a = np.asarray(t[k]).reshape(-1)
        assert len(a) == g[k]["n"], k
        nz = np.nonzero(a)[0]
        assert nz.tolist() == g[k]["index"] and a[nz].tolist() == g[k]["value"], k
    v = np.ascontiguousarray(t["vox"], np.int64)
    assert int(np.count_nonzero(v)) == g["vox"]["nonzero"] and int(v.sum()) == g["vox"]["sum"]
    assert hashlib.sha256(v.tobytes()).hexdigest() == g["vox"]["sha256"]
    for k, val in g["misc"].items():
        if k != "arz_out":  # the Arz image is not among the tallies kept here
            assert t["misc"][k] == val, k
    if opts.get("no_skip"):
        assert t["misc"]["skipped"] == 0

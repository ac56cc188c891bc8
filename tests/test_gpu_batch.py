"""Batched launches (lp_run_batch) and the autotuner on the GPU.

* Every instance of a batch equals lp_set_materials(instance) + lp_run with
  the same seed and photon range, bit for bit (detector and region tallies):
  instances share the per-photon random streams.
* The captured HIP graph, replayed with new optics between eager launches
  and host synchronisations, gives the eager batch's tallies every time.
* The autotuner changes launch options only: the tallies after tuning equal
  those before, and its trial launches (which carry the context's tallies
  into stand-in buffers) leave the context's tallies untouched.
"""
import numpy as np
import pytest

import lightpropagation_amd as lp
from lightpropagation_amd import scenes as S

pytestmark = pytest.mark.gpu


def _ctx(seed=31):
    sc = S.maternal_fetal(freq_maternal=8, freq_fetal=8, freq_torso=4)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=seed)
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    ctx.set_detectors(c - 0.25, c + 0.25)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    return ctx, sc


def _table(pairs):
    return np.stack([S.mf_optics(wl, sm, sf) for wl, sm, sf in pairs])


def _single(ctx, sc, nm, first, count):
    ctx.set_materials(nm, sc["fetal_id"])
    ctx.reset()
    ctx.run(first, count)
    t = ctx.tallies()
    return t["det"].reshape(-1, 2), t["region"]


def test_batch_equals_single_runs():
    ctx, sc = _ctx()
    tab = _table([(660.0, 0.97, 0.5), (660.0, 0.97, 0.3), (940.0, 0.9, 0.75), (940.0, 0.6, 0.4)])
    first, count = 12345, 150_000
    print("batch of", len(tab), "x", count, flush=True)
    b = ctx.run_batch(tab, first, count)
    assert ctx.stats().photons == len(tab) * count
    for i, nm in enumerate(tab):
        det, reg = _single(ctx, sc, nm, first, count)
        assert np.array_equal(b["det"][i], det), i
        assert np.array_equal(b["region"][i], reg), i
        assert det[:, 0].sum() > 0
    # instances differ (the optics matter) and one instance is a batch of one
    assert not np.array_equal(b["det"][0], b["det"][1])
    one = ctx.run_batch(tab[2:3], first, count)
    assert np.array_equal(one["det"][0], b["det"][2]) and np.array_equal(one["region"][0], b["region"][2])
    ctx.close()


def test_graph_replays_equal_eager():
    ctx, sc = _ctx(seed=77)
    rng = np.random.default_rng(3)
    first, count = 0, 100_000
    for rep in range(4):
        pairs = [(wl, float(sm), float(sf)) for wl in (660.0, 940.0) for sm, sf in rng.uniform(0.3, 1.0, (3, 2))]
        tab = _table(pairs)
        ctx.set_option("graph", 1)
        g = ctx.run_batch(tab, first, count)
        ctx.reset()
        ctx.run(first + 7, 20_000)  # eager work between replays
        ctx.set_option("graph", 0)
        e = ctx.run_batch(tab, first, count)
        print("replay", rep, "det sums", g["det"][:, :, 0].sum(1), flush=True)
        assert np.array_equal(g["det"], e["det"]), rep
        assert np.array_equal(g["region"], e["region"]), rep
    ctx.close()


def test_autotune_keeps_results():
    ctx, sc = _ctx(seed=5)
    ctx.set_voxel_grid([-6.4, -6.4, 0.0], [0.1, 0.1, 0.1], [128, 128, 128])
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["vox"] | 1 << lp.TALLY["misc"])
    ctx.reset()
    ctx.run(0, 200_000)
    before = ctx.tallies()
    r = ctx.autotune(sample=400_000, passes=1)
    during = ctx.tallies()  # the trial launches tally into stand-ins
    for k in ("det", "region", "vox", "misc"):
        assert before[k] == during[k] if k == "misc" else np.array_equal(before[k], during[k]), k
    print("autotune: %d trials, %.1f -> %.1f ms; waves %d, blocks/CU %d, free %d, refill %d, list %.3g, search %.3g"
          % (r.trials, r.ms_before, r.ms_after, r.waves, r.blocks_per_cu, r.free_iters, r.refill_min, r.list_min,
             r.search_min), flush=True)
    assert r.trials >= 2 and r.ms_after <= r.ms_before
    ctx.reset()
    ctx.run(0, 200_000)
    after = ctx.tallies()
    for k in ("det", "region", "vox"):
        assert np.array_equal(before[k], after[k]), k
    assert before["vox"].sum() > 0
    ctx.close()


def test_grid_follows_occupancy():
    """The automatic grid holds as many work-groups per CU as one CU runs at
    once: 4 for the 4-waves variant (128 VGPRs), 5 for the 5-waves one; an
    explicit LP_OPT_BLOCKS_PER_CU overrides it."""
    ctx, sc = _ctx()
    grids = {}
    for key, opts in (("w4", {}), ("w5", {"waves_per_simd": 5}), ("bpc2", {"waves_per_simd": 4, "blocks_per_cu": 2})):
        for k, v in opts.items():
            ctx.set_option(k, v)
        ctx.run(0, 400_000)
        grids[key] = ctx.stats().grid_blocks
    print("grids", grids, flush=True)
    ctx.close()
    assert grids["w4"] % 4 == 0 and grids["w5"] * 4 == grids["w4"] * 5 and grids["bpc2"] * 2 == grids["w4"]


def test_batch_rejects_oversized_tables():
    ctx, sc = _ctx()
    tab = np.repeat(_table([(660.0, 0.97, 0.5)]), 200, axis=0)  # 200 x 12 rows > 2048
    with pytest.raises(lp.LPError):
        ctx.run_batch(tab, 0, 10)
    ctx.close()


def test_batch_against_oracle():
    # each instance of a batch against the oracle with that instance's
    # optics, same seed and photon range: detector and region tallies bit
    # for bit
    import _oracle as O
    ctx, sc = _ctx(seed=41)
    tab = _table([(660.0, 0.97, 0.5), (940.0, 0.9, 0.3), (660.0, 0.8, 0.7)])
    first, count = 4321, 300
    b = ctx.run_batch(tab, first, count)
    ctx.close()
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    for i, nm in enumerate(tab):
        run = O.Run(dict(sc, nmuasg=nm), seed=41, tallies={"det_rlo": c - 0.25, "det_rhi": c + 0.25})
        run.run(first, count, nthreads=8)
        t = run.tallies()
        assert np.array_equal(b["det"][i], t["det"].reshape(-1, 2)), i
        assert np.array_equal(b["region"][i], t["region"]), i

"""Several geometries in one batched launch (lp_set_geometries +
lp_run_batch_geom) on the GPU.

* Instances spread over three uneven maternal-fetal geometries (fetal head
  at 1.5, 2.5 and 3.5 cm; 8 960, 14 000 and 5 040 faces, overlapping in
  space) give
  the detector and region tallies of lp_run_batch on a mesh of each
  geometry alone, bit for bit: a photon meets only its own geometry's
  faces, through that geometry's BVH and near-face lists.  Checked with the
  near-face lists on and off (BVH only), through the captured graph and
  eagerly.
* lp_run in a mesh of several geometries transports in geometry 0.
* Bad geometry counts and instance geometries are refused.
"""
import numpy as np
import pytest

import lightpropagation_amd as lp
from lightpropagation_amd import scenes as S

pytestmark = pytest.mark.gpu

GEOMS = [(1.5, 8, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, 10, {}), (3.5, 6, {})]
TALLIES = 1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"]


def _scenes():
    return [S.maternal_fetal(fetal_depth=d, freq_maternal=f, freq_fetal=f, freq_torso=4, **kw) for d, f, kw in GEOMS]


def _ctx(sc, seed, counts=None, near=True):
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=seed)
    if counts is not None:
        ctx.set_geometries(counts)
    if not near:
        ctx.set_option("near_r0", 0.0)
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    ctx.set_detectors(c - 0.25, c + 0.25)
    ctx.set_tallies(TALLIES)
    return ctx


def _table():
    pairs = [(660.0, 0.97, 0.5), (940.0, 0.97, 0.3), (660.0, 0.8, 0.75), (940.0, 0.6, 0.45), (660.0, 0.9, 0.35),
             (940.0, 0.97, 0.6), (660.0, 0.7, 0.55)]
    return np.stack([S.mf_optics(wl, sm, sf) for wl, sm, sf in pairs])


@pytest.mark.parametrize("near", [True, False])
def test_geometries_equal_separate_meshes(near):
    scs = _scenes()
    merged, counts = S.geometries(scs)
    assert len(set(counts)) == 3
    tab = _table()
    geom = np.array([0, 1, 2, 2, 0, 1, 1], np.int32)
    first, count = 4242, 120_000
    ctx = _ctx(merged, 0x5EED0009, counts, near)
    print("merged: faces", counts, "near lists", near, flush=True)
    g = ctx.run_batch(tab, first, count, geom=geom)
    assert ctx.stats().photons == len(tab) * count
    ctx.set_option("graph", 0)
    e = ctx.run_batch(tab, first, count, geom=geom)
    assert np.array_equal(g["det"], e["det"]) and np.array_equal(g["region"], e["region"])
    ctx.close()
    for k, sc in enumerate(scs):
        sel = np.nonzero(geom == k)[0]
        one = _ctx(sc, 0x5EED0009, near=near)
        s = one.run_batch(tab[sel], first, count)
        one.close()
        print("geometry", k, "instances", sel.tolist(), "R", s["det"][:, :, 0].sum(1), flush=True)
        assert np.array_equal(g["det"][sel], s["det"]), k
        assert np.array_equal(g["region"][sel], s["region"]), k
        assert s["det"][:, :, 0].sum() > 0
    # the geometries matter: the same optics in two geometries differ
    both = _ctx(merged, 0x5EED0009, counts, near)
    x = both.run_batch(np.stack([tab[0], tab[0]]), first, count, geom=[0, 2])
    both.close()
    assert not np.array_equal(x["det"][0], x["det"][1])


def test_run_transports_in_geometry_zero():
    scs = _scenes()
    merged, counts = S.geometries(scs)
    a = _ctx(merged, 77, counts)
    a.run(100, 60_000)
    ta = a.tallies()
    a.close()
    b = _ctx(scs[0], 77)
    b.run(100, 60_000)
    tb = b.tallies()
    b.close()
    for k in ("det", "region"):
        assert np.array_equal(ta[k], tb[k]), k


def test_geometry_errors():
    scs = _scenes()
    merged, counts = S.geometries(scs)
    ctx = _ctx(merged, 1)
    with pytest.raises(lp.LPError):
        ctx.set_geometries([counts[0], counts[1]])  # do not cover the mesh
    with pytest.raises(lp.LPError):
        ctx.set_geometries([counts[0] + counts[1] + counts[2], 0])
    ctx.set_geometries(counts)
    with pytest.raises(lp.LPError):
        ctx.run_batch(_table()[:2], 0, 10, geom=[0, 3])
    ctx.set_launch(0, -1, 16)  # a safety grid needs one geometry
    with pytest.raises(lp.LPError):
        ctx.run_batch(_table()[:2], 0, 10, geom=[0, 1])
    ctx.close()


def test_batched_geometries_against_oracle():
    # the batched launch over three merged geometries against the oracle run
    # on each instance's own geometry with that instance's optics: detector
    # and region tallies bit for bit, for 200 photons per instance
    import _oracle as O
    scs = _scenes()
    merged, counts = S.geometries(scs)
    tab = _table()
    geom = np.array([0, 1, 2, 2, 0, 1, 1], np.int32)
    first, count = 777, 200
    ctx = _ctx(merged, 0x5EED0011, counts)
    g = ctx.run_batch(tab, first, count, geom=geom)
    ctx.close()
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    for i in range(len(tab)):
        sc = dict(scs[geom[i]], nmuasg=tab[i])
        run = O.Run(sc, seed=0x5EED0011, tallies={"det_rlo": c - 0.25, "det_rhi": c + 0.25})
        run.run(first, count, nthreads=8)
        t = run.tallies()
        print("instance", i, "geometry", geom[i], "R", t["det"].reshape(-1, 2)[:, 0].sum(), flush=True)
        assert np.array_equal(g["det"][i], t["det"].reshape(-1, 2)), i
        assert np.array_equal(g["region"][i], t["region"]), i


C2_GEOMS = [(1.5, 19, 14, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, 38, 27, {}), (3.5, 45, 32, {}), (4.5, 53, 38, {})]


def test_autotune_on_merged_geometries_keeps_results():
    # the autotuner's trial launches on a context holding several
    # geometries (transport in geometry 0) leave its tallies untouched, and
    # the batched launch over the geometries gives the same tallies with the
    # options it picks as with the defaults
    scs = _scenes()
    merged, counts = S.geometries(scs)
    tab = _table()
    geom = (np.arange(len(tab)) % len(counts)).astype(np.int32)
    ctx = _ctx(merged, 0x5EED0031, counts)
    before = ctx.run_batch(tab, 0, 30_000, geom=geom)
    r = ctx.autotune(sample=300_000, passes=1)
    print("autotune: %d trials, %.1f -> %.1f ms; waves %d, free %d, refill %d, list %.3g"
          % (r.trials, r.ms_before, r.ms_after, r.waves, r.free_iters, r.refill_min, r.list_min), flush=True)
    assert r.trials >= 2 and r.ms_after <= r.ms_before
    after = ctx.run_batch(tab, 0, 30_000, geom=geom)
    ctx.close()
    for k in ("det", "region"):
        assert np.array_equal(before[k], after[k]), k
    assert before["det"][:, :, 0].sum() > 0


def test_batched_workload_at_full_size():
    # the batched workload's real shape -- 4 merged geometries (56 480 to
    # 377 120 faces), 2 wavelengths x 16 fetal SpO2 per geometry = 128
    # instances in one launch -- with 20 000 photons per instance: every
    # instance equals lp_run of its optics on its geometry alone
    spo2 = np.linspace(0.30, 0.75, 16)
    scs = [S.maternal_fetal(fetal_depth=d, freq_maternal=fm, freq_fetal=ff, **kw) for d, fm, ff, kw in C2_GEOMS]
    merged, counts = S.geometries(scs)
    assert counts == [56_480, 196_420, 272_400, 377_120]
    one = np.stack([S.mf_optics(wl, 0.97, sf) for wl in (660.0, 940.0) for sf in spo2])
    tab = np.concatenate([one] * 4)
    geom = np.repeat(np.arange(4, dtype=np.int32), len(one))
    first, count = 0, 20_000
    ctx = _ctx(merged, 0x5EED0002, counts)
    print("merged", counts, "instances", len(tab), flush=True)
    g = ctx.run_batch(tab, first, count, geom=geom)
    assert ctx.stats().photons == 128 * count
    ctx.close()
    for k, sc in enumerate(scs):
        one_ctx = _ctx(sc, 0x5EED0002)
        for i in np.nonzero(geom == k)[0]:
            one_ctx.set_materials(tab[i], sc["fetal_id"])
            one_ctx.reset()
            one_ctx.run(first, count)
            t = one_ctx.tallies()
            assert np.array_equal(g["det"][i], t["det"].reshape(-1, 2)), (k, i)
            assert np.array_equal(g["region"][i], t["region"]), (k, i)
        print("geometry", k, "done", flush=True)
        one_ctx.close()
    assert g["det"][:, :, 0].sum() > 0

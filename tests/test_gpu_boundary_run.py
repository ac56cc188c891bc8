"""The boundary-run cap on the GPU: the trapped photon of
test_oracle_boundary_run.py (caught on the edge of faces 28171 and 28172 of
the thin-layer geometry at step 744) ends as capped at step 10 743 with a
boundary-run limit of 10 000, its whole step trace equal to the oracle's
bit for bit, with the crawl replay on (its crossings repeated from the
lane's records) and off; and a batched launch that holds it costs no more
than one photon's 10 000 extra steps.  On the bench workload (196 420
faces, 660 nm), photons crawling along an edge in cycles of two and of four
crossings give the same step traces, exit records and tallies with the
replay on and off, with the Philox streams and with ran3."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _oracle as O  # noqa: E402
import lightpropagation_amd as lp  # noqa: E402
from test_oracle_boundary_run import ONSET, PHOTON, RUN, SEED, trapped_scene  # noqa: E402

pytestmark = pytest.mark.gpu


@pytest.mark.parametrize("crawl", [1, 0], ids=["replay", "no_replay"])
def test_trapped_photon_trace_equals_oracle(crawl):
    sc = trapped_scene()
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=SEED)
    ctx.set_option("boundary_run", RUN)
    ctx.set_option("crawl", crawl)
    ctx.set_tallies(1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    print("photon", PHOTON, "crawl", crawl, flush=True)
    tr, nrec = ctx.trace(PHOTON, 1, ONSET + RUN + 10)
    m = ctx.tallies()["misc"]
    ctx.close()
    assert m["capped"] == 1 and m["steps"] == ONSET + RUN == nrec[0]
    # the replay takes nearly all of the run's crossings
    assert (m["crawl_replayed"] > RUN // 2) if crawl else (m["crawl_replayed"] == 0)
    orc = O.Run(sc, seed=SEED, max_steps=0, tallies={}, max_boundary_run=RUN)
    _, otr, on = orc.run(PHOTON, 1, trace_max=ONSET + RUN + 10)
    assert on[0] == nrec[0]
    assert np.array_equal(tr[0, :nrec[0]].view(np.uint32), otr[0, :on[0]].view(np.uint32))


# photons of the 660 nm bench stream (seed 0x5EED0001) caught on an edge
# until the boundary-run cap of 10^5: the crossings of 954 434 101 (faces
# 34405 and 34406, from step 9 144) repeat with period two, those of
# 949 166 754 (faces 50136 and 50137, from step 3 892) with period four
CRAWLERS = [954_434_101, 949_166_754]


def _bench_ctx(crawl):
    from lightpropagation_amd import scenes as S
    import bench
    ctx = lp.Context(0)
    ctx.set_option("crawl", crawl)
    bench.setup(ctx, S.maternal_fetal(wavelength_nm=660.0), seed=0x5EED0001)
    return ctx


def test_headline_crawls_replayed_exactly():
    tr, n, misc = {}, {}, {}
    for crawl in (1, 0):
        ctx = _bench_ctx(crawl)
        for pid in CRAWLERS:
            print("trace photon", pid, "crawl", crawl, flush=True)
            ctx.reset()
            t, k = ctx.trace(pid, 1, 120_000)
            tr[crawl, pid], n[crawl, pid] = t[0, :k[0]].copy(), int(k[0])
            misc[crawl, pid] = ctx.tallies()["misc"]
        ctx.close()
    for pid in CRAWLERS:
        assert n[1, pid] == n[0, pid] > 100_000
        assert np.array_equal(tr[1, pid].view(np.uint32), tr[0, pid].view(np.uint32)), pid
        assert misc[1, pid]["capped"] == misc[0, pid]["capped"] == 1
        assert misc[1, pid]["crawl_replayed"] > 90_000 and misc[0, pid]["crawl_replayed"] == 0
        # the run: 10^5 transmissions at one position
        run = tr[1, pid][-100_000:]
        assert np.all(run[:, 0] == 1.0)
        assert len({tuple(r) for r in run[:, 4:7].view(np.uint32).tolist()}) == 1


def test_headline_range_with_crawlers_same_tallies():
    # 2 * 10^5 photons around each crawler: every tally, the exit records
    # and the physics counters equal with the replay on and off
    out = {}
    for crawl in (1, 0):
        ctx = _bench_ctx(crawl)
        ctx.set_tallies(sum(1 << k for k in (0, 1, 4, 5, 6, 7, 8)), record_exit=True)
        res = []
        for pid in CRAWLERS:
            print("range around", pid, "crawl", crawl, flush=True)
            ctx.reset()
            ctx.run(pid - 100_000, 200_000)
            t = ctx.tallies()
            res.append(({k: np.asarray(v).copy() for k, v in t.items() if k != "misc"}, dict(t["misc"]),
                        ctx.exit_records(200_000).copy()))
        out[crawl] = res
        ctx.close()
    for (t1, m1, e1), (t0, m0, e0) in zip(out[1], out[0]):
        for k in t1:
            assert np.array_equal(t1[k], t0[k]), k
        assert np.array_equal(e1.view(np.uint32), e0.view(np.uint32))
        for k in ("photons", "steps", "interact", "boundary", "escaped", "capped", "nan"):
            assert m1[k] == m0[k], k
        assert m1["capped"] >= 1 and m1["crawl_replayed"] > 90_000 and m0["crawl_replayed"] == 0


# under the ran3 streams of the same workload, photons [93 650 000,
# 93 700 000) hold a crawler that reaches the cap: the replay peeks each
# Fresnel draw from the lane's ran3 state before consuming it
RAN3_RANGE = (93_650_000, 50_000)


def test_ran3_range_with_crawler_same_tallies():
    out = {}
    for crawl in (1, 0):
        ctx = _bench_ctx(crawl)
        from lightpropagation_amd import scenes as S
        sc = S.maternal_fetal(wavelength_nm=660.0)
        ctx.set_params(float(sc["chance"]), float(sc["w_th"]), 1.0, 0x5EED0001, lp.RNG_RAN3, 0)
        ctx.set_tallies(sum(1 << k for k in (0, 1, 4, 5, 6, 7, 8)), record_exit=True)
        print("ran3 range", RAN3_RANGE, "crawl", crawl, flush=True)
        ctx.reset()
        ctx.run(*RAN3_RANGE)
        t = ctx.tallies()
        out[crawl] = ({k: np.asarray(v).copy() for k, v in t.items() if k != "misc"}, dict(t["misc"]),
                      ctx.exit_records(RAN3_RANGE[1]).copy())
        ctx.close()
    (t1, m1, e1), (t0, m0, e0) = out[1], out[0]
    for k in t1:
        assert np.array_equal(t1[k], t0[k]), k
    assert np.array_equal(e1.view(np.uint32), e0.view(np.uint32))
    for k in ("photons", "steps", "interact", "boundary", "escaped", "capped", "nan"):
        assert m1[k] == m0[k], k
    assert m1["capped"] == 1 and m1["crawl_replayed"] > 90_000 and m0["crawl_replayed"] == 0


def test_batch_holding_the_trapped_photon():
    sc = trapped_scene()
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=SEED)
    ctx.set_option("boundary_run", RUN)
    ctx.set_tallies(1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    tab = np.stack([sc["nmuasg"]] * 4)
    ctx.reset()
    b = ctx.run_batch(tab, PHOTON - 50_000, 100_000)
    ms = ctx.stats().kernel_ms
    m = ctx.tallies()["misc"]
    print("4 x 100000 photons around the trapped one: %.1f ms, capped %d" % (ms, m["capped"]), flush=True)
    ctx.close()
    assert m["capped"] == 4  # the trapped photon in each instance
    assert np.array_equal(b["region"][0], b["region"][3])
    assert ms < 2000.0  # before the cap, each trapped photon ran 10^6 steps (about 14 s)

"""The boundary-run cap on the GPU: the trapped photon of
test_oracle_boundary_run.py (caught on the edge of faces 28171 and 28172 of
the thin-layer geometry at step 744) ends as capped at step 10 743 with a
boundary-run limit of 10 000, its whole step trace equal to the oracle's
bit for bit, and a batched launch that holds it costs no more than one
photon's 10 000 extra steps."""
import os
import sys

import numpy as np
import pytest

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _oracle as O  # noqa: E402
import lightpropagation_amd as lp  # noqa: E402
from test_oracle_boundary_run import ONSET, PHOTON, RUN, SEED, trapped_scene  # noqa: E402

pytestmark = pytest.mark.gpu


def test_trapped_photon_trace_equals_oracle():
    sc = trapped_scene()
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=SEED)
    ctx.set_option("boundary_run", RUN)
    ctx.set_tallies(1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    print("photon", PHOTON, flush=True)
    tr, nrec = ctx.trace(PHOTON, 1, ONSET + RUN + 10)
    m = ctx.tallies()["misc"]
    ctx.close()
    assert m["capped"] == 1 and m["steps"] == ONSET + RUN == nrec[0]
    orc = O.Run(sc, seed=SEED, max_steps=0, tallies={}, max_boundary_run=RUN)
    _, otr, on = orc.run(PHOTON, 1, trace_max=ONSET + RUN + 10)
    assert on[0] == nrec[0]
    assert np.array_equal(tr[0, :nrec[0]].view(np.uint32), otr[0, :on[0]].view(np.uint32))


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

"""The large maternal-fetal workload on the GPU: frequency-118 shells and a
frequency-40 fetal torso ellipsoid (1 992 320 faces, a BVH far deeper than
its LDS-staged top), 940 nm, a 512^3 int64 absorption grid (1 GiB) and 8
annular detectors, at test-sized photon counts.

* Exit records of 16 photons equal the oracle's (brute-force search over
  all 2e6 faces) bit for bit.
* 2e5 photons give identical exit records with the near-face lists on and
  off (BVH only).
* 10^6 photons: no NaN or capped photon; the grid's absorption plus what
  fell outside it equals the per-region total; R + T + absorbed = 1 - Rsp
  to the roulette's statistical error.
* The adjoint of 10^6 photons replays every detected path exactly.
"""
import numpy as np
import pytest

import lightpropagation_amd as lp
from lightpropagation_amd import scenes as S
import _oracle as O

pytestmark = pytest.mark.gpu
SEED = 0x5EED0004
_SCENE = {}


def _scene():
    if "sc" not in _SCENE:
        print("building the 2e6-face mesh", flush=True)
        _SCENE["sc"] = S.maternal_fetal(wavelength_nm=940.0, freq_maternal=118, freq_fetal=92, freq_torso=40)
    return _SCENE["sc"]


def _ctx(sc, **opts):
    ctx = lp.Context(0)
    for k, v in opts.items():
        ctx.set_option(k, v)
    ctx.set_scene(sc, seed=SEED)
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    ctx.set_detectors(c - 0.25, c + 0.25)
    return ctx


@pytest.fixture(scope="module")
def ctx():
    """One context on the large mesh (the BVH and near-face build takes tens
    of seconds), shared by the tests of this module."""
    c = _ctx(_scene())
    yield c
    c.close()


def test_large_mesh_exit_records_equal_oracle(ctx):
    sc = _scene()
    assert len(sc["tri"]) == 1_992_320
    # the oracle tests all 2e6 faces per step (72 MB of loads): take a block
    # of 16 consecutive photons that end within 3000 steps
    _, nrec = ctx.trace(0, 512, 3000)
    first = next(16 * b for b in range(32) if nrec[16 * b:16 * b + 16].max() < 3000)
    print("photons", first, "..", first + 15, "steps", nrec[first:first + 16].tolist(), flush=True)
    ctx.set_tallies(record_exit=True)
    ctx.reset()
    ctx.run(first, 16)
    got = ctx.exit_records(16).copy()
    print("oracle: 16 photons, brute force", flush=True)
    run = O.Run(sc, seed=SEED)
    want, _, _ = run.run(first, 16, exit_rec=True, nthreads=16)
    assert np.array_equal(got.view(np.uint32), want.view(np.uint32))


def test_large_mesh_near_lists_on_and_off(ctx):
    n = 200_000
    ctx.set_tallies(record_exit=True)
    ctx.reset()
    print("near-face lists on", flush=True)
    ctx.run(0, n)
    on = ctx.exit_records(n).copy()
    off = _ctx(_scene(), near_r0=0.0)
    off.set_tallies(record_exit=True)
    print("near-face lists off", flush=True)
    off.run(0, n)
    rec = off.exit_records(n).copy()
    off.close()
    assert np.array_equal(on.view(np.uint32), rec.view(np.uint32))


def test_large_mesh_grid_and_energy(ctx):
    ctx.set_voxel_grid([-12.8, -12.8, 0.0], [0.05, 0.05, 0.05], [512, 512, 512])
    ctx.set_tallies(sum(1 << lp.TALLY[k] for k in ("sft", "region", "vox", "det", "misc")))
    ctx.reset()
    N = 1_000_000
    print("1e6 photons with the 512^3 grid", flush=True)
    ctx.run(0, N)
    t = ctx.tallies()
    rsp = ctx.layout().rsp
    m = t["misc"]
    assert m["nan"] == 0 and m["capped"] == 0 and m["escaped"] == 0 and m["photons"] == N
    region = int(t["region"].sum())
    assert int(t["vox"].sum()) + int(m["vox_out"]) == region
    assert int(t["vox"].sum()) > 0.5 * region
    sft = t["sft"].reshape(-1, 3).sum(0) / lp.FIXED
    # a photon starts with 1 - Rsp and leaves with w, recorded as w (1 - Rsp)
    # (the reference's double factor); roulette keeps the expectation
    total = (region / lp.FIXED + (sft[0] + sft[2]) / (1 - rsp)) / N
    assert abs(total - (1 - rsp)) < 2e-3, (total, 1 - rsp)


def test_large_mesh_adjoint_replays_exactly(ctx):
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
    ctx.reset()
    print("adjoint, 1e6 photons", flush=True)
    r = ctx.adjoint(0, 1_000_000)
    st = ctx.stats()
    assert st.replay_mismatch == 0 and r["pairs"] > 0
    # the forward half adds to the tallies as lp_run does
    assert np.array_equal(r["signals"], ctx.tallies()["det"].reshape(-1, 2)[:, 0])
    g = r["grad"] / lp.FIXED / 1e6
    assert np.all(np.isfinite(g)) and np.all(g[:, 1:, 0] <= 0)  # more absorption, less light
    print("pairs %d, forward %.0f ms, replay %.0f ms" % (r["pairs"], st.kernel_ms, st.replay_ms), flush=True)

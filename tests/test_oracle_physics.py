"""Whole-path checks of the oracle.

* MCML verification slab (reference fixture models/model1: mua = 10/cm,
  mus = 90/cm, g = 0.75, n = 1, d = 0.02 cm).  Published values: van de
  Hulst (1980) Rd = 0.09739, Tt = 0.66096; MCML (Wang, Jacques & Zheng,
  Comput. Methods Programs Biomed. 47:131-146, 1995, Table 2)
  Rd = 0.09734 +- 0.00035, Tt = 0.66096 +- 0.00020.  With matched n every
  faithful-mode quirk of the reference is inert, so the oracle must agree
  within 3 combined standard errors.
* Beer-Lambert: the unscattered fraction of the same slab is exp(-mu_t d).
* Energy balance: E[R + T]/(1 - Rsp) + E[A] = 1 - Rsp.
* Determinism: int64 fixed-point tallies do not depend on thread count.
"""
import math

import numpy as np
import pytest

import _oracle as O
import scenes

FX = 2.0 ** 32


def _run(sc, n, seed=11, rng=O.RNG_PHILOX, threads=8, **kw):
    r = O.Run(sc, seed=seed, rng_mode=rng, **kw)
    ex, _, _ = r.run(0, n, exit_rec=True, nthreads=threads)
    return r, ex


@pytest.mark.parametrize("rng", [O.RNG_PHILOX, O.RNG_RAN3F])
def test_mcml_slab_rd_tt(rng):
    sc = scenes.load_fixture("model1")
    n = 200000
    r, ex = _run(sc, n, rng=rng)
    R = ex[:, 3].astype(np.float64)
    T = ex[:, 5].astype(np.float64)
    sR = math.hypot(R.std() / math.sqrt(n), 0.00035)
    sT = math.hypot(T.std() / math.sqrt(n), 0.00020)
    assert abs(R.mean() - 0.09734) < 3 * sR, R.mean()
    assert abs(T.mean() - 0.66096) < 3 * sT, T.mean()
    t = r.tallies()
    A = t["region"].sum() / FX / n
    assert abs(R.mean() + T.mean() + A - 1.0) < 1e-5


def test_beer_lambert_unscattered_fraction():
    sc = scenes.load_fixture("model1")
    n = 100000
    r = O.Run(sc, seed=3)
    _, tr, tn = r.run(0, n, trace_max=3, nthreads=8)
    # entry (transmit), then the exit face with no interaction in between
    ball = (tn >= 2) & (tr[:, 0, 0] == O.EV_TRANSMIT) & (tr[:, 1, 0] == O.EV_TRANSMIT) & \
           (np.abs(tr[:, 1, 6] - 0.02) < 1e-6)
    frac = ball.mean()
    p = math.exp(-100.0 * 0.02)
    assert abs(frac - p) < 3 * math.sqrt(p * (1 - p) / n)


def test_energy_balance_mismatched_layers():
    sc = scenes.slab(scenes.C0_LAYERS)
    n = 4000
    r, ex = _run(sc, n)
    rsp = r.L.rsp
    t = r.tallies()
    RT = (ex[:, 3].astype(np.float64) + ex[:, 5]).sum() / n
    A = t["region"].sum() / FX / n
    per = (ex[:, 3].astype(np.float64) + ex[:, 5]) / (1 - rsp)
    lhs = per.mean() + A
    se = per.std() / math.sqrt(n) + 0.02 / math.sqrt(n)
    assert abs(lhs - (1 - rsp)) < 4 * se + 1e-3, (lhs, 1 - rsp)
    assert t["misc"]["escaped"] == 0 and t["misc"]["capped"] == 0


def test_thread_count_does_not_change_tallies():
    sc = scenes.load_fixture("model1")
    tl = dict(vox=dict(origin=[-0.5, -0.5, 0.0], size=[0.01, 0.01, 0.001], n=[100, 100, 20]),
              rad=dict(dr=0.05, n=200), det_rlo=[0.0, 0.1], det_rhi=[0.1, 0.5])
    a = O.Run(sc, seed=5, tallies=tl)
    a.run(0, 20000, nthreads=1)
    b = O.Run(sc, seed=5, tallies=tl)
    b.run(0, 7000, nthreads=8)
    b.run(7000, 13000, nthreads=3)  # split run == whole run (resume is exact)
    ta, tb = a.tallies(), b.tallies()
    for k in ("sft", "rtd", "arz", "phirz", "region", "vox", "rad", "det"):
        assert np.array_equal(ta[k], tb[k]), k
    assert ta["misc"] == tb["misc"]

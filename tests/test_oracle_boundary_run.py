"""The boundary-run cap, on the oracle.

Photon 9 904 387 of the thin-layer maternal-fetal geometry (fetal head at
1.5 cm, fat 0.3 cm, uterine wall 0.5 cm, frequency 19: 50 540 faces; 940 nm,
maternal SpO2 0.97, fetal 0.60; Philox seed 0x5EED0002) crosses at step 744
onto the shared edge of faces 28171 and 28172 and then crosses from one to
the other at t ~ 0 with no interaction between: the last-face exclusion
cannot separate two faces, so under the reference's loop it never ends.
With no step cap and a boundary-run limit of 10 000 (the library's default
is 100 000; the smaller limit keeps the brute-force oracle fast) the oracle
ends it as capped at step 743 + 10 000, its trace alternating between the
two faces.
The GPU side of the same photon is in test_gpu_boundary_run.py.
"""
import os
import sys

import numpy as np

sys.path.insert(0, os.path.dirname(os.path.abspath(__file__)))
import _oracle as O  # noqa: E402
from lightpropagation_amd import scenes as S  # noqa: E402

PHOTON = 9_904_387
SEED = 0x5EED0002
ONSET = 743  # steps before the first crossing onto the edge
RUN = 10_000


def trapped_scene():
    # the seven-region thin geometry of the batched workload it was found in
    sc = S.maternal_fetal(regions=7, fetal_depth=1.5, freq_maternal=19, freq_fetal=19, t_fat=0.3, t_uterus=0.5)
    sc["nmuasg"] = S.mf_optics(940.0, 0.97, float(np.linspace(0.30, 0.75, 16)[10]), regions=7)
    return sc


def test_trapped_photon_ends_after_the_boundary_run():
    sc = trapped_scene()
    assert len(sc["tri"]) == 50_540
    run = O.Run(sc, seed=SEED, max_steps=0, tallies={}, max_boundary_run=RUN)
    _, tr, tn = run.run(PHOTON, 1, trace_max=ONSET + RUN + 10)
    m = run.tallies()["misc"]
    assert m["capped"] == 1 and m["photons"] == 1
    assert m["steps"] == ONSET + RUN == tn[0]
    t = tr[0, :tn[0]]
    assert t[ONSET - 1, 0] == 0  # the last interaction
    kinds, faces = t[ONSET:, 0], t[ONSET:, 1].astype(int)
    assert np.all(kinds != 0) and len(kinds) == RUN
    assert set(faces.tolist()) == {28171, 28172} and np.all(faces[1:] != faces[:-1])
    # the photon does not move: every crossing is at t ~ 0
    assert np.ptp(t[ONSET:, 4:7], axis=0).max() < 1e-5


def test_cap_spares_ordinary_photons():
    """The neighbouring photons of the same scene scatter and end normally."""
    sc = trapped_scene()
    run = O.Run(sc, seed=SEED, max_steps=0, tallies={}, max_boundary_run=RUN)
    run.run(PHOTON + 1, 6, nthreads=6)
    m = run.tallies()["misc"]
    assert m["capped"] == 0 and m["photons"] == 6 and m["interact"] > 0

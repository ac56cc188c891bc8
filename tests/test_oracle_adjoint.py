"""The oracle's adjoint (path-length estimator of the detector-signal
gradients) on the CPU, pinned by what the geometry and the estimator fix:

* the voxel traversal gives each voxel the length of the segment inside it
  (exact box/segment overlaps in double on random segments), and the lengths
  add up to the segment;
* per-voxel mua gradients summed over a grid that holds every path equal the
  per-tissue mua gradients summed over tissues (both are -sum L W);
* the per-tissue gradients equal the estimator evaluated on the oracle's own
  step traces (path length per tissue from the positions, interactions per
  tissue from the event kinds), so a dropped move, a move booked to the
  tissue after a crossing instead of before, or a miscounted interaction
  fails;
* dS/dmus for a tissue with n interactions and length L is (n/mus - L) W:
  checked on the same traces.
"""
import numpy as np
import pytest

import _oracle as O
import scenes

FX = 2.0 ** 32


def _overlap(p0, p1, lo, hi):
    """Length of segment p0 -> p1 inside the box [lo, hi) (double)."""
    d = p1 - p0
    t0, t1 = 0.0, 1.0
    for a in range(3):
        if d[a] == 0:
            if not (lo[a] <= p0[a] < hi[a]):
                return 0.0
            continue
        ta, tb = (lo[a] - p0[a]) / d[a], (hi[a] - p0[a]) / d[a]
        ta, tb = min(ta, tb), max(ta, tb)
        t0, t1 = max(t0, ta), min(t1, tb)
    return max(0.0, t1 - t0) * float(np.linalg.norm(d))


def test_dda_matches_exact_overlaps():
    rng = np.random.default_rng(5)
    origin, size, n = [-1.0, -0.5, 0.0], [0.25, 0.2, 0.1], [8, 6, 10]
    for trial in range(300):
        p0 = rng.uniform([-1.5, -1.0, -0.3], [1.5, 1.0, 1.3])
        if trial % 3 == 0:  # axis-parallel and in-plane segments too
            p1 = p0.copy()
            p1[trial % 2] += rng.uniform(-1, 1)
        else:
            p1 = p0 + rng.normal(0, 0.5, 3)
        p0f, p1f = p0.astype(np.float32), p1.astype(np.float32)
        s = np.float32(np.linalg.norm(p1f.astype(np.float64) - p0f))
        if s == 0:
            continue
        seg = np.concatenate([p0f, p1f, [s]]).astype(np.float32)
        jac = O.dda(seg, 1.0, origin, size, n).reshape(n[2], n[1], n[0]) / -FX
        want = np.zeros_like(jac)
        for z in range(n[2]):
            for y in range(n[1]):
                for x in range(n[0]):
                    lo = np.array(origin) + np.array([x, y, z]) * np.array(size)
                    want[z, y, x] = _overlap(p0f.astype(np.float64), p1f.astype(np.float64), lo,
                                             lo + np.array(size))
        assert np.allclose(jac, want, rtol=1e-4, atol=2e-6 * max(float(s), 1.0)), trial


def _c0_run(seed=3, layers=None):
    sc = scenes.slab(layers or scenes.C0_LAYERS)
    tall = {"det_rlo": [0.0, 0.3, 1.0], "det_rhi": [0.3, 1.0, 3.0]}
    return sc, O.Run(sc, seed=seed, tallies=tall)


def test_voxel_and_tissue_gradients_agree():
    _, run = _c0_run()
    # a grid holding every path: the 10 x 10 cm slab 3.15 cm deep, and the
    # 0.002 cm of ambient above it where the pencil beam starts
    grid = ([-5.0, -5.0, -0.01], [0.25, 0.25, 0.0205], [40, 40, 155])
    sig, grad, jac, pairs = run.adjoint(0, 3000, jac=grid, nthreads=8)
    assert pairs > 0
    for d in range(3):
        tissue = grad[d, :, 0].sum() / FX
        voxels = jac[d].sum() / FX
        assert abs(voxels - tissue) <= 1e-5 * abs(tissue) + 1e-6, (d, voxels, tissue)


@pytest.mark.parametrize("layers", ["c0", "slab10"])
def test_gradients_equal_estimator_on_traces(layers):
    # the three-layer slab, and ten layers: 10 regions x (mua, mus) = 20 parameters
    sc, run = _c0_run(seed=11, layers=scenes.SLAB10_LAYERS if layers == "slab10" else None)
    N = 400
    sig, grad, _, pairs = run.adjoint(0, N)
    ex, tr, tn = run.run(0, N, exit_rec=True, trace_max=20000)
    nt = len(sc["nmuasg"])
    lo, hi = [0.0, 0.3, 1.0], [0.3, 1.0, 3.0]
    want = np.zeros((3, nt, 2))
    want_sig = np.zeros(3)
    p0 = np.asarray(sc["pstart"], np.float64)
    for i in range(N):
        W, radius = ex[i, 3], ex[i, 7]
        if W == 0:
            continue
        L = np.zeros(nt)
        n = np.zeros(nt)
        prev, tissue = p0, 0
        for k in range(tn[i]):
            rec = tr[i, k]
            pos = rec[4:7].astype(np.float64)
            L[tissue] += np.linalg.norm(pos - prev)  # the move happens in the tissue before the event
            if rec[0] == 0:
                n[tissue] += 1
            prev, tissue = pos, int(rec[2])
        for d in range(3):
            if lo[d] <= radius < hi[d]:
                want_sig[d] += W
                for t in range(nt):
                    mus = sc["nmuasg"][t][2]
                    want[d, t, 0] += -L[t] * W
                    want[d, t, 1] += ((n[t] / mus - L[t]) * W) if mus != 0 else -L[t] * W
    assert np.allclose(sig / FX, want_sig, rtol=1e-6, atol=1e-6)
    # the ambient start (tissue 0) is 0.002 cm long; the escape move of a
    # photon is never in a detected path
    assert np.allclose(grad / FX, want, rtol=2e-4, atol=1e-4)
    assert np.any(grad[:, 1:, 1] != 0)
    if layers == "slab10":
        assert np.all((grad[:, 1:, :] != 0).any(axis=0))  # every region reached by a detected path


def test_oracle_adjoint_refuses_a_pixel_per_unit_factor():
    import ctypes as C
    sc = dict(scenes.slab(scenes.C0_LAYERS), factor=2.0)
    run = O.Run(sc, seed=1, tallies={"det_rlo": [0.0], "det_rhi": [1.0]})
    A = O.AdjointOut()
    sig = np.zeros(1, np.int64)
    A.signals = O.lptr(sig)
    assert O.lib().orc_adjoint(C.byref(run.P), 0, 10, None, C.byref(A), 1) == 3


def _coords32(pos, pb):
    """The oracle's emitter-frame coordinates in float32: (p - POI) . e with
    each basis vector normalised as p * (1 / sqrt(p . p))."""
    f = np.float32
    d = (np.asarray(pos, f) - np.asarray(pb[:3], f)).astype(f)
    out = []
    for k in range(3):
        e = np.asarray(pb[3 + 3 * k:6 + 3 * k], f)
        inv = f(1.0) / np.sqrt(f((e[0] * e[0] + e[1] * e[1]) + e[2] * e[2]))
        e = (inv * e).astype(f)
        out.append(f(f(e[0] * d[0] + e[1] * d[1]) + f(e[2] * d[2])))
    return np.array(out, f)


def test_voxel_mus_gradient_on_traces():
    # dS/dmus_v = (sum over the interactions in v of 1 / mus) W - L_v W: the
    # difference from the mua Jacobian is W / mus booked at each detected
    # path's interaction points, recomputed here from the step traces
    sc, run = _c0_run(seed=13)
    grid = ([-5.0, -5.0, -0.01], [0.25, 0.25, 0.0205], [40, 40, 155])
    N = 500
    sig, grad, jac, pairs, jac_s = run.adjoint(0, N, jac=grid, jac_mus=True)
    ex, tr, tn = run.run(0, N, exit_rec=True, trace_max=20000)
    lo, hi = [0.0, 0.3, 1.0], [0.3, 1.0, 3.0]
    o, sz, n = (np.asarray(x, np.float32) for x in grid)
    inv = (np.float32(1.0) / sz).astype(np.float32)
    want = np.zeros_like(jac_s)
    pb = np.asarray(sc["poi_basis"], np.float32)
    for i in range(N):
        W, radius = ex[i, 3], ex[i, 7]
        if W == 0:
            continue
        tissue = 0
        for k in range(tn[i]):
            rec = tr[i, k]
            if rec[0] == 0:  # an interaction, in the tissue the photon is in
                g = ((_coords32(rec[4:7], pb) - o) * inv).astype(np.float32)
                if np.all(g >= 0) and np.all(g < n):
                    v = (int(g[2]) * int(n[1]) + int(g[1])) * int(n[0]) + int(g[0])
                    q = int(np.float32(W / np.float32(sc["nmuasg"][tissue][2])) * np.float32(FX))
                    for d in range(3):
                        if lo[d] <= radius < hi[d]:
                            want[d, v] += q
            tissue = int(rec[2])
    assert np.array_equal(jac_s - jac, want)
    # summed over a grid that holds every path: the per-tissue mus gradients
    for d in range(3):
        tissue = grad[d, :, 1].sum() / FX
        voxels = jac_s[d].sum() / FX
        assert abs(voxels - tissue) <= 1e-5 * abs(tissue) + 1e-5, (d, voxels, tissue)


def test_path_records_are_the_traced_paths():
    # per-photon path records (the perturbation Monte Carlo inputs): the
    # length per tissue from the traced positions, the interactions per
    # tissue from the event kinds, and the gradients are their sums
    sc, run = _c0_run(seed=17)
    N = 400
    nt = len(sc["nmuasg"])
    sig, grad, _, pairs, P = run.adjoint(0, N, paths=True)
    ex, tr, tn = run.run(0, N, exit_rec=True, trace_max=20000)
    lo, hi = [0.0, 0.3, 1.0], [0.3, 1.0, 3.0]
    assert np.array_equal(P["w"], ex[:, 3]) and np.array_equal(P["radius"], ex[:, 7])
    p0 = np.asarray(sc["pstart"], np.float64)
    want_g = np.zeros((3, nt, 2), np.int64)
    npairs = 0
    for i in range(N):
        L = np.zeros(nt)
        n = np.zeros(nt, np.int32)
        prev, tissue = p0, 0
        for k in range(tn[i]):
            rec = tr[i, k]
            pos = rec[4:7].astype(np.float64)
            L[tissue] += np.linalg.norm(pos - prev)
            if rec[0] == 0:
                n[tissue] += 1
            prev, tissue = pos, int(rec[2])
        W = P["w"][i]
        if W == 0:
            continue
        assert np.array_equal(P["interactions"][i], n), i
        assert np.allclose(P["length"][i], L, rtol=2e-4, atol=1e-5), i
        for d in range(3):
            if lo[d] <= P["radius"][i] < hi[d]:
                npairs += 1
                for t in range(nt):
                    Lt, nt_ = P["length"][i, t], P["interactions"][i, t]
                    if Lt == 0 and nt_ == 0:
                        continue
                    mus = np.float32(sc["nmuasg"][t][2])
                    gm = np.float32(-(Lt * W))
                    gs = np.float32(np.float32(np.float32(nt_) / mus) - Lt) * W if mus != 0 else gm
                    want_g[d, t, 0] += int(np.float32(gm) * np.float32(FX))
                    want_g[d, t, 1] += int(np.float32(gs) * np.float32(FX))
    assert npairs == pairs > 0
    assert np.array_equal(grad, want_g)


def test_path_records_reweight_to_other_albedos():
    # optics with the same mu_t = mua + mus in every layer give every photon
    # the same path (same step lengths, same draws) and change only the
    # albedo of each interaction, so a photon's weight at the new optics is
    # its weight at the old times prod_t (mus'_t / mus_t)^n_t: the records'
    # interaction counts per tissue, and nothing else, decide it
    sc, run = _c0_run(seed=19)
    N = 600
    _, _, _, _, P = run.adjoint(0, N, paths=True)
    nm = np.array(sc["nmuasg"], np.float32)
    nm2 = nm.copy()
    f = np.float32
    for t in range(1, len(nm)):
        mus2 = f(nm[t, 2] - f(0.25))
        mua2 = f(f(nm[t, 1] + nm[t, 2]) - mus2)
        assert f(mua2 + mus2) == f(nm[t, 1] + nm[t, 2])
        nm2[t, 1], nm2[t, 2] = mua2, mus2
    sc2 = dict(sc, nmuasg=nm2)
    run2 = O.Run(sc2, seed=19, tallies={"det_rlo": [0.0, 0.3, 1.0], "det_rhi": [0.3, 1.0, 3.0]})
    ex2, _, _ = run2.run(0, N, exit_rec=True)
    ratio = np.ones(len(nm))
    ratio[1:] = nm2[1:, 2].astype(np.float64) / nm[1:, 2]
    same = (P["w"] > 0) & (ex2[:, 7] == P["radius"])  # same exit: no roulette decided otherwise
    assert same.sum() > 0.8 * (P["w"] > 0).sum() > 100
    pred = P["w"].astype(np.float64) * np.prod(ratio[None, :] ** P["interactions"], axis=1)
    assert np.allclose(ex2[same, 3], pred[same], rtol=2e-4, atol=0)
    # one interaction booked to the wrong tissue moves the prediction by
    # about 0.1 %: the check sees it
    bad = P["interactions"].copy()
    k = np.argmax(same & (bad[:, 1] > 0) & (bad[:, 2] > 0))
    bad[k, 1] -= 1
    bad[k, 2] += 1
    wrong = P["w"][k] * np.prod(ratio ** bad[k])
    assert abs(wrong - ex2[k, 3]) > 5e-4 * ex2[k, 3]

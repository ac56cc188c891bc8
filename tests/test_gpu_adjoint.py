"""The adjoint replay on the GPU (lp_adjoint): detector-signal gradients by
the path-length estimator, replayed from the detected photons' (seed, index)
random streams.

* Bit-exact against the oracle's restatement (orc_adjoint): signals,
  per-tissue (mua, mus) gradients and the per-voxel mua Jacobian, and every
  replay ends with the forward photon's exit weight.
* Against central finite differences with common random numbers (step 2 %,
  four photon batches for the standard errors) on the 202 160-face bench
  workload.
* Against the scaling identity of a semi-infinite homogeneous medium: total
  reflectance does not change when every mu is multiplied by k (lengths
  scale by 1/k), so sum_t (mua_t dR/dmua_t + mus_t dR/dmus_t) = 0, i.e.
  E[(n - mut L) W] = 0 over the detected photons.
"""
import math

import numpy as np
import pytest

import lightpropagation_amd as lp
import _oracle as O
import scenes

pytestmark = pytest.mark.gpu
FX = 2.0 ** 32


def _scene(name):
    if name == "c0slab":
        return scenes.slab(scenes.C0_LAYERS), [0.0, 0.3, 1.0], [0.3, 1.0, 3.0], \
            ([-5.0, -5.0, -0.01], [0.25, 0.25, 0.0205], [40, 40, 155])
    if name == "slab10":  # ten regions: 20 (mua, mus) parameters
        return scenes.slab(scenes.SLAB10_LAYERS), [0.0, 0.3, 1.0], [0.3, 1.0, 3.0], \
            ([-5.0, -5.0, -0.01], [0.25, 0.25, 0.02], [40, 40, 52])
    from lightpropagation_amd import scenes as S
    return S.maternal_fetal(freq_maternal=6, freq_fetal=6, freq_torso=4), [0.0, 1.0, 3.0], [1.0, 3.0, 10.0], \
        ([-6.4, -6.4, 0.0], [0.2, 0.2, 0.2], [64, 64, 64])


@pytest.mark.parametrize("name,n", [("c0slab", 20000), ("slab10", 20000), ("mf_coarse", 3000)])
def test_adjoint_bit_exact_vs_oracle(name, n):
    sc, lo, hi, grid = _scene(name)
    print("adjoint", name, n, flush=True)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=21)
    ctx.set_detectors(lo, hi)
    ctx.set_jacobian_grid(*grid)
    got = ctx.adjoint(0, n, jac=True, jac_mus=True, paths=True)
    st = ctx.stats()
    ctx.close()
    run = O.Run(sc, seed=21, tallies={"det_rlo": lo, "det_rhi": hi})
    sig, grad, jac, pairs, jac_s, P = run.adjoint(0, n, jac=grid, nthreads=8, jac_mus=True, paths=True)
    assert st.replay_mismatch == 0 and st.replayed == pairs == got["pairs"] > 0
    assert np.array_equal(got["signals"], sig)
    assert np.array_equal(got["grad"], grad)
    assert np.array_equal(got["jac"].reshape(len(lo), -1), jac)
    assert np.array_equal(got["jac_mus"].reshape(len(lo), -1), jac_s)
    assert np.any(jac_s != jac)
    # the per-pair path records, sorted by (photon, detector), against the
    # oracle's per-photon records
    G = got["paths"]
    want = [(i, d) for i in range(n) for d in range(len(lo))
            if P["w"][i] != 0 and lo[d] <= P["radius"][i] < hi[d]]
    assert [(int(a), int(b)) for a, b in zip(G["photon"], G["det"])] == want
    ids = G["photon"]
    assert np.array_equal(G["w"], P["w"][ids])
    assert np.array_equal(G["length"], P["length"][ids])
    assert np.array_equal(G["interactions"], P["interactions"][ids].astype(np.uint32))


def test_adjoint_replay_with_fewer_lanes_than_pairs():
    # More detected pairs than replay lanes: lanes that finish a pair take
    # the next from the list.  The outputs (integer sums, path records
    # sorted by (photon, detector)) equal those of a grid that holds every
    # pair at once (checked against the oracle above).
    sc, lo, hi, grid = _scene("c0slab")
    out = []
    for blocks in (0, None):
        ctx = lp.Context(0)
        ctx.set_scene(sc, seed=21)
        ctx.set_detectors(lo, hi)
        ctx.set_jacobian_grid(*grid)
        if blocks is None:  # about a quarter as many lanes as pairs
            blocks = max(1, out[0]["pairs"] // (4 * 256))
            ctx.set_option("grid_blocks", blocks)
        print("adjoint c0slab 20000, grid_blocks", blocks, flush=True)
        got = ctx.adjoint(0, 20000, jac=True, jac_mus=True, paths=True)
        got["lanes"] = blocks * 256
        assert ctx.stats().replay_mismatch == 0
        ctx.close()
        out.append(got)
    a, b = out
    assert 2 * b["lanes"] < b["pairs"]
    assert a["pairs"] == b["pairs"]
    for k in ("signals", "grad", "jac", "jac_mus"):
        assert np.array_equal(a[k], b[k]), k
    for k in ("photon", "det", "w", "length", "interactions"):
        assert np.array_equal(a["paths"][k], b["paths"][k]), k


def _bench_ctx():
    from lightpropagation_amd import scenes as S
    import bench
    sc = S.maternal_fetal()
    ctx = lp.Context(0)
    bench.setup(ctx, sc, seed=0x5EED0001)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
    return ctx, sc


def test_adjoint_against_central_differences():
    ctx, sc = _bench_ctx()
    nm0 = np.asarray(sc["nmuasg"], np.float32)
    params = [(1, 1), (2, 1), (1, 2), (2, 2)]  # (tissue, 1 = mua / 2 = mus): maternal skin and fat
    dets = [0, 2]
    B, N, rel = 4, 2_000_000, 0.02
    adj = np.zeros((B, len(params), len(dets)))
    for b in range(B):
        g = ctx.adjoint(b * N, N)["grad"] / FX / N
        for i, (t, c) in enumerate(params):
            adj[b, i] = g[dets, t, c - 1]
    st = ctx.stats()
    print("forward %.1f ms, replay %.1f ms for %d pairs" % (st.kernel_ms, st.replay_ms, st.replayed), flush=True)
    assert st.replay_mismatch == 0
    fd = np.zeros_like(adj)
    for i, (t, c) in enumerate(params):
        h = float(nm0[t, c]) * rel
        S = {}
        for sgn in (-1, 1):
            nm = nm0.copy()
            nm[t, c] = np.float32(nm0[t, c] + sgn * h)
            ctx.set_materials(nm, sc["fetal_id"])
            S[sgn] = []
            for b in range(B):
                ctx.reset()
                ctx.run(b * N, N)
                S[sgn].append(ctx.tallies()["det"].reshape(-1, 2)[dets, 0] / FX / N)
        dh = float(np.float32(nm0[t, c] + h)) - float(np.float32(nm0[t, c] - h))
        for b in range(B):
            fd[b, i] = (S[1][b] - S[-1][b]) / dh
    ctx.close()
    a_mean, f_mean = adj.mean(0), fd.mean(0)
    se = np.sqrt(adj.var(0, ddof=1) / B + fd.var(0, ddof=1) / B)
    print("adjoint", a_mean, "\nfinite differences", f_mean, "\nse", se, flush=True)
    assert np.all(np.abs(a_mean - f_mean) <= 4 * se + 0.03 * np.abs(a_mean)), (a_mean, f_mean, se)


def test_scaling_identity_semi_infinite():
    # one tissue (mua 0.5, mus 10, g 0.9, n 1.4), 40 x 40 x 20 cm: a
    # semi-infinite medium for photons that travel centimetres.  The mus part
    # of the estimator, (n/mus - L) W, has a per-photon standard deviation of
    # about four times its mean here (and grows as sqrt(interactions) at higher
    # mus), so the standard error comes from 16 batches of 10^6 photons, and
    # the test also requires it to be small enough that a term dropped from
    # the estimator (one interaction, one segment: of the order of R itself)
    # would fail.
    layers = [(20.0, 1.40, 0.5, 10.0, 0.9)]
    sc = scenes.slab(layers, half=20.0)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=5)
    ctx.set_detectors([0.0], [1000.0])  # every reflected photon
    B, N = 16, 1_000_000
    mua, mus = float(np.float32(layers[0][2])), float(np.float32(layers[0][3]))
    vals, scales, R = [], [], []
    for b in range(B):
        r = ctx.adjoint(b * N, N)
        g = r["grad"][0] / FX / N
        vals.append(mua * g[1, 0] + mus * g[1, 1])
        scales.append(abs(mua * g[1, 0]) + abs(mus * g[1, 1]))
        R.append(r["signals"][0] / FX / N)
    assert ctx.stats().replay_mismatch == 0
    ctx.close()
    vals = np.array(vals)
    scale, se = float(np.mean(scales)), vals.std(ddof=1) / math.sqrt(B)
    print("sum mu dR/dmu: mean %.3e se %.3e scale %.4f R %.5f" % (vals.mean(), se, scale, np.mean(R)), flush=True)
    assert se < 0.01 * scale and se < 0.02 * np.mean(R)
    assert abs(vals.mean()) <= 4 * se + 1e-3 * scale


def test_adjoint_outputs_are_cleared_when_nothing_is_detected():
    # a call that fills the gradient and Jacobian buffers, then one whose
    # detectors no photon reaches, then one with no photons: the later calls
    # must return zeros (their buffers are cleared on the context's stream
    # and copied back only after it has drained)
    sc, lo, hi, grid = _scene("c0slab")
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=21)
    ctx.set_detectors(lo, hi)
    ctx.set_jacobian_grid(*grid)
    first = ctx.adjoint(0, 4000, jac=True)
    assert first["pairs"] > 0 and np.any(first["grad"] != 0) and np.any(first["jac"] != 0)
    ctx.set_detectors([500.0, 600.0, 700.0], [501.0, 601.0, 701.0])
    for count in (4000, 0):
        r = ctx.adjoint(0, count, jac=True)
        st = ctx.stats()
        assert r["pairs"] == 0 and st.replay_mismatch == 0
        assert not np.any(r["signals"]) and not np.any(r["grad"]) and not np.any(r["jac"])
    ctx.close()


def _reweighted(P, nm0, nm1, ndet, count):
    """Signals per photon and their variances at optics nm1 from the path
    records P of photons that ran at nm0 (perturbation Monte Carlo)."""
    nm0, nm1 = nm0.astype(np.float64), nm1.astype(np.float64)
    ok = nm0[:, 2] > 0
    lr = P["interactions"] @ np.where(ok, np.log(np.where(ok, nm1[:, 2] / np.where(ok, nm0[:, 2], 1), 1)), 0)
    lr -= P["length"].astype(np.float64) @ ((nm1[:, 1] - nm0[:, 1]) + (nm1[:, 2] - nm0[:, 2]))
    w = P["w"].astype(np.float64) * np.exp(lr)
    s1 = np.bincount(P["det"], w, ndet) / count
    s2 = np.bincount(P["det"], w * w, ndet) / count
    return s1, (s2 - s1 * s1) / count, np.bincount(P["det"], minlength=ndet), w


def test_reweighted_paths_match_runs_at_other_optics():
    # the path records of photons run at one set of optics, reweighted to
    # another (fetal SpO2 0.50 -> 0.30 and maternal 0.97 -> 0.90 at 660 nm,
    # and a 1 % change of the maternal fat's mus), agree with the signals
    # of independent photons run at the other optics within 4 standard
    # errors at every detector with 100 or more pairs: the estimator is
    # unbiased (roulette and Fresnel decisions do not depend on the optics).
    # A mus change spreads the log weights by about sqrt(n) |ln(mus'/mus)|
    # (n: interactions in the tissue, hundreds in the fat): at 10 % the
    # weights are heavy-tailed and 4e6 photons no longer sample the tail,
    # so the sample mean falls short and its variance estimate with it.
    from lightpropagation_amd import scenes as S
    ctx, sc = _bench_ctx()
    N = 4_000_000
    nm0 = S.mf_optics(660.0, 0.97, 0.50)
    others = [S.mf_optics(660.0, 0.97, 0.30), S.mf_optics(660.0, 0.90, 0.50)]
    fat = nm0.copy()
    fat[2, 2] *= 1.01
    others.append(fat)
    ctx.set_materials(nm0, sc["fetal_id"])
    A = ctx.adjoint(0, N, paths=True)
    nd = len(A["signals"])
    s0, _, _, w0 = _reweighted(A["paths"], nm0, nm0, nd, N)
    assert np.allclose(s0, A["signals"] / FX / N, rtol=1e-5, atol=0)  # the identity at the records' optics
    for k, nm1 in enumerate(others):
        ctx.set_materials(nm1, sc["fetal_id"])
        B = ctx.adjoint((k + 1) * N, N, paths=True)
        sa, va, na, wa = _reweighted(A["paths"], nm0, nm1, nd, N)
        sb, vb, nb, _ = _reweighted(B["paths"], nm1, nm1, nd, N)
        use = (na >= 100) & (nb >= 100)
        z = (sa - sb) / np.sqrt(va + vb)
        print("optics", k, "detectors", int(use.sum()), "z", np.round(z[use], 2), flush=True)
        assert use.sum() >= 4 and np.all(np.abs(z[use]) < 4.0), z
        # the change itself is seen: the reweighted signals differ from the
        # records' own by more than 4 standard errors of that paired difference
        dd = wa - w0
        m1 = np.bincount(A["paths"]["det"], dd, nd) / N
        vd = (np.bincount(A["paths"]["det"], dd * dd, nd) / N - m1 * m1) / N
        assert np.any(np.abs(sa[use] - s0[use]) > 4 * np.sqrt(vd[use]))
    ctx.close()


def test_adjoint_refuses_a_pixel_per_unit_factor():
    # the path-length estimator assumes step lengths -ln(xi) / mu_t
    sc, lo, hi, grid = _scene("c0slab")
    sc = dict(sc, factor=np.float32(2.0))
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=21)
    ctx.set_detectors(lo, hi)
    with pytest.raises(lp.LPError, match="factor"):
        ctx.adjoint(0, 100)
    ctx.close()


def _voxel_inclusion_scene():
    """Two layers (0.5 cm, mus 100 and 80, g 0.9, one index) with an
    inclusion of its own tissue (3: mua 0.3, mus 10, g 0, so that its mus
    matters enough for differences to resolve) that coincides with one voxel
    of the Jacobian grid: the box [0.4, 0.6] x [0, 0.2] x [0.2, 0.4] in the
    emitter frame (the grid's frame), mapped to the world by the emitter
    basis."""
    from lightpropagation_amd import scenes as S
    layers = [(0.5, 1.40, 0.3, 100.0, 0.9), (2.0, 1.40, 0.2, 80.0, 0.9)]
    sc = S.slab(layers, half=5.0)
    pb = np.asarray(sc["poi_basis"], np.float64)
    poi, E = pb[:3], pb[3:].reshape(3, 3)
    lo, hi = np.array([0.4, 0.0, 0.2]), np.array([0.6, 0.2, 0.4])
    corner = lambda i, j, k: poi + E.T @ np.array([(lo, hi)[i][0], (lo, hi)[j][1], (lo, hi)[k][2]])
    quads = [((0, 0, 0), (0, 1, 0), (1, 1, 0), (1, 0, 0)), ((0, 0, 1), (1, 0, 1), (1, 1, 1), (0, 1, 1)),
             ((0, 0, 0), (1, 0, 0), (1, 0, 1), (0, 0, 1)), ((0, 1, 0), (0, 1, 1), (1, 1, 1), (1, 1, 0)),
             ((0, 0, 0), (0, 0, 1), (0, 1, 1), (0, 1, 0)), ((1, 0, 0), (1, 1, 0), (1, 1, 1), (1, 0, 1))]
    tri = []
    centre = poi + E.T @ ((lo + hi) / 2)
    for q in quads:
        a, b, c, d = (corner(*v) for v in q)
        for t in ((a, b, c), (a, c, d)):
            n = np.cross(t[1] - t[0], t[2] - t[0])
            if np.dot(n, np.mean(t, axis=0) - centre) < 0:  # wind outward
                t = (t[0], t[2], t[1])
            tri.append(np.concatenate(t))
    tri = np.array(tri, np.float32)
    n = np.cross(tri[:, 3:6].astype(np.float64) - tri[:, 0:3], tri[:, 6:9].astype(np.float64) - tri[:, 0:3])
    n /= np.linalg.norm(n, axis=1, keepdims=True)
    sc["tri"] = np.ascontiguousarray(np.concatenate([sc["tri"], tri]))
    sc["nrm"] = np.ascontiguousarray(np.concatenate([sc["nrm"], n.astype(np.float32)]))
    sc["med"] = np.ascontiguousarray(np.concatenate([sc["med"], np.tile(np.array([[3, 1]], np.int32), (12, 1))]))
    sc["nmuasg"] = np.concatenate([sc["nmuasg"], [[1.40, 0.3, 10.0, 0.0]]]).astype(np.float32)
    grid = ([-1.0, -1.0, 0.0], [0.2, 0.2, 0.2], [10, 10, 5])
    voxel = (1 * 10 + 5) * 10 + 7  # iz 1, iy 5, ix 7
    return sc, grid, voxel


def test_voxel_jacobians_against_differences_on_one_voxel():
    # the mua and mus Jacobians at the inclusion's voxel against central
    # differences of the detector signals in the inclusion's mua and mus,
    # with common random numbers (the voxel's optics are the inclusion's)
    sc, grid, voxel = _voxel_inclusion_scene()
    lo, hi = [0.2, 0.6, 1.2], [0.6, 1.2, 2.5]
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=0x5EED0021)
    ctx.set_detectors(lo, hi)
    ctx.set_jacobian_grid(*grid)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
    B, N, rel = 8, 4_000_000, 0.05
    adj = np.zeros((B, 2, 3))
    tis = np.zeros((B, 2, 3))
    for b in range(B):
        r = ctx.adjoint(b * N, N, jac=True, jac_mus=True)
        adj[b, 0] = r["jac"].reshape(3, -1)[:, voxel] / FX / N
        adj[b, 1] = r["jac_mus"].reshape(3, -1)[:, voxel] / FX / N
        tis[b] = r["grad"][:, 3, :].T / FX / N  # the inclusion tissue's own gradients
    assert ctx.stats().replay_mismatch == 0
    nm0 = np.asarray(sc["nmuasg"], np.float32)
    fd = np.zeros_like(adj)
    for i, c in enumerate((1, 2)):
        h = float(nm0[3, c]) * rel
        S = {}
        for sgn in (-1, 1):
            nm = nm0.copy()
            nm[3, c] = np.float32(nm0[3, c] + sgn * h)
            ctx.set_materials(nm, -1)
            S[sgn] = []
            for b in range(B):
                ctx.reset()
                ctx.run(b * N, N)
                S[sgn].append(ctx.tallies()["det"].reshape(-1, 2)[:, 0] / FX / N)
        dh = float(np.float32(nm0[3, c] + h)) - float(np.float32(nm0[3, c] - h))
        for b in range(B):
            fd[b, i] = (S[1][b] - S[-1][b]) / dh
    ctx.close()
    a, f = adj.mean(0), fd.mean(0)
    se = np.sqrt(adj.var(0, ddof=1) / B + fd.var(0, ddof=1) / B)
    print("voxel adjoint", a, "\nfinite differences", f, "\nse", se, flush=True)
    # the voxel holds the inclusion: its Jacobian is the inclusion's gradient
    assert np.allclose(a, tis.mean(0), rtol=2e-3, atol=1e-9)
    assert np.all(np.abs(a) > 3 * se)  # resolved
    assert np.all(np.abs(a - f) <= 4 * se + 0.05 * np.abs(a)), (a, f, se)

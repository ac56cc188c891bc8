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
    return S.maternal_fetal(freq_maternal=6, freq_fetal=6), [0.0, 1.0, 3.0], [1.0, 3.0, 10.0], \
        ([-6.4, -6.4, 0.0], [0.2, 0.2, 0.2], [64, 64, 64])


@pytest.mark.parametrize("name,n", [("c0slab", 20000), ("slab10", 20000), ("mf_coarse", 3000)])
def test_adjoint_bit_exact_vs_oracle(name, n):
    sc, lo, hi, grid = _scene(name)
    print("adjoint", name, n, flush=True)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=21)
    ctx.set_detectors(lo, hi)
    ctx.set_jacobian_grid(*grid)
    got = ctx.adjoint(0, n, jac=True)
    st = ctx.stats()
    ctx.close()
    run = O.Run(sc, seed=21, tallies={"det_rlo": lo, "det_rhi": hi})
    sig, grad, jac, pairs = run.adjoint(0, n, jac=grid, nthreads=8)
    assert st.replay_mismatch == 0 and st.replayed == pairs == got["pairs"] > 0
    assert np.array_equal(got["signals"], sig)
    assert np.array_equal(got["grad"], grad)
    assert np.array_equal(got["jac"].reshape(len(lo), -1), jac)


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

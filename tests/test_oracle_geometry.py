"""Pins the oracle's ray-triangle search (FindTriangleInPath,
MCLightPropagation.cl:1037-1093) and its optics (Rspecular :273-299,
GetFresnel :857-907, Transmit/Reflect :506-549, VScatter :597-655)."""
import math

import numpy as np
import pytest

import _oracle as O
import scenes

f32 = np.float32


def brute_double(tri, p, v, ref=-1):
    """Moeller-Trumbore style barycentric solve in double precision."""
    best, bid = 10000.0, -1
    T = np.asarray(tri, np.float64).reshape(-1, 9)
    p = np.asarray(p, np.float64)
    v = np.asarray(v, np.float64)
    for j, t9 in enumerate(T):
        if j == ref:
            continue
        K, L, M = t9[0:3], t9[3:6], t9[6:9]
        A = np.stack([K - M, L - M, -v], axis=1)
        if abs(np.linalg.det(A)) < 1e-14:
            continue
        k, l, t = np.linalg.solve(A, p - M)
        if 0 < t < best and k >= 0 and l >= 0 and k + l < 1:
            best, bid = t, j
    return bid, best


def test_random_rays_match_double_brute_force():
    sc = scenes.load_fixture("model4")  # 25600 triangles: 5 concentric spheres
    rng = np.random.default_rng(5)
    tri = sc["tri"]
    sub = tri[rng.choice(len(tri), 400, replace=False)]
    agree = 0
    for i in range(300):
        p = rng.uniform(-12, 12, 3).astype(f32)
        v = rng.normal(size=3)
        v = (v / np.linalg.norm(v)).astype(f32)
        tid, t = O.find_triangle(sub, p, v)
        bid, bt = brute_double(sub, p, v)
        assert tid == bid, (i, tid, bid)
        if tid >= 0:
            assert abs(t - bt) <= 1e-5 * max(1.0, bt)
            agree += 1
    assert agree > 5


def test_known_answer_model1_box():
    sc = scenes.load_fixture("model1")
    # from the source point straight down +z: first face is the z=0 top
    tid, t = O.find_triangle(sc["tri"], sc["pstart"], [0, 0, 1])
    assert tid in (0, 6) and abs(t - 0.002) < 1e-8
    assert sc["med"][tid].tolist() == [1, 0]
    # excluding it (refTri) from a point on the face finds the z=0.02 face
    tid2, t2 = O.find_triangle(sc["tri"], [0.001, -0.003, 0.0], [0, 0, 1], ref_tri=tid)
    assert abs(t2 - 0.02) < 1e-7 and abs(sc["nrm"][tid2][2] - 1) < 1e-7
    # nothing ahead when leaving the box
    assert O.find_triangle(sc["tri"], [0, 0, 0.03], [0, 0, 1])[0] == -1


def test_tie_break_lowest_index():
    # two identical triangles: the strict < keeps the first one
    t9 = [-1, -1, 1, 1, -1, 1, 0, 1, 1]
    tri = np.array([t9, t9, t9], f32)
    assert O.find_triangle(tri, [0, 0, 0], [0, 0, 1])[0] == 0
    assert O.find_triangle(tri, [0, 0, 0], [0, 0, 1], ref_tri=0)[0] == 1


def test_fixture_normals_match_file_vn():
    for m in ("model1", "model4"):
        sc = scenes.load_fixture(m)
        vn = sc["file_vn"]
        ok = ~np.isnan(vn[:, 0])
        dots = np.sum(sc["nrm"][ok] * vn[ok], axis=1)
        assert np.all(dots > 0.99), m


def test_fresnel_closed_forms():
    for n1, n2 in [(1.0, 1.4), (1.4, 1.0), (1.37, 1.44)]:
        r0 = ((n1 - n2) / (n1 + n2)) ** 2
        assert abs(O.fresnel(n1, n2, 1.0) - r0) < 1e-7
        assert abs(O.fresnel(n1, n2, -1.0) - r0) < 1e-7
        # unpolarised Fresnel at an oblique angle (cos_in > 0)
        c = 0.6
        s = math.sqrt(1 - c * c)
        st = n1 * s / n2
        if st < 1:
            ct = math.sqrt(1 - st * st)
            rs = ((n1 * c - n2 * ct) / (n1 * c + n2 * ct)) ** 2
            rp = ((n1 * ct - n2 * c) / (n1 * ct + n2 * c)) ** 2
            assert abs(O.fresnel(n1, n2, c) - 0.5 * (rs + rp)) < 2e-6
        else:
            assert O.fresnel(n1, n2, c) == 1.0
    assert O.fresnel(1.3, 1.3, 0.3) == 0.0
    assert O.fresnel(1.0, 1.4, 0.0) == 1.0
    # the reference's signed-cosine quirk: oblique cos_in < 0 with n1 != n2
    # yields r > 1 (always reflect)
    assert O.fresnel(1.0, 1.4, -0.6) > 1.0


def test_rspecular():
    nm = np.array([[1.0, 0, 0, 0], [1.4, 0.4, 200, 0.9]], f32)
    assert abs(O.lib().orc_rspecular(O.fptr(nm), 2) - (0.4 / 2.4) ** 2) < 1e-8
    # glass layer (mua = mus = 0) between ambient and tissue: MCML two-interface form
    nm = np.array([[1.0, 0, 0, 0], [1.5, 0, 0, 0], [1.33, 1, 10, 0.9]], f32)
    r1 = (0.5 / 2.5) ** 2
    r2 = (0.17 / 2.83) ** 2
    ref = r1 + (1 - r1) ** 2 * r2 / (1 - r1 * r2)
    assert abs(O.lib().orc_rspecular(O.fptr(nm), 3) - ref) < 1e-7


def test_vscatter_unit_and_hg_mean():
    rng = np.random.default_rng(7)
    for g in (0.0, 0.75, 0.9, -0.5):
        cs = []
        for _ in range(20000):
            v = rng.normal(size=3)
            v /= np.linalg.norm(v)
            xt, xp = rng.random(2)
            w = O.vscatter(v.astype(f32), g, f32(xt), f32(xp))
            assert abs(np.linalg.norm(w) - 1) < 1e-5
            cs.append(float(np.dot(w, v)))
        cs = np.array(cs)
        # E[cos theta] = g for Henyey-Greenstein; isotropic has E[cos^2] = 1/3
        se = cs.std() / math.sqrt(len(cs))
        assert abs(cs.mean() - g) < 4 * se + 2e-3, (g, cs.mean())
        if g == 0.0:
            assert abs(np.mean(cs ** 2) - 1 / 3) < 0.01


def test_hg_ks():
    # KS test of cos(theta) against the Henyey-Greenstein CDF, g = 0.9
    g = 0.9
    rng = np.random.default_rng(8)
    v0 = np.array([0.3, -0.4, np.sqrt(1 - 0.25)], f32)
    cs = np.sort([float(np.dot(O.vscatter(v0, g, f32(a), f32(b)), v0)) for a, b in rng.random((5000, 2))])
    mu = np.clip(cs, -1, 1)
    cdf = (1 - g * g) / (2 * g) * (1 / np.sqrt(1 + g * g - 2 * g * mu) - 1 / (1 + g))
    d = np.max(np.abs(cdf - (np.arange(1, len(cs) + 1) / len(cs))))
    assert d < 1.63 / math.sqrt(len(cs))  # 1% level


def test_fresnel_rounded_normal_incidence_is_finite():
    # |cos| slightly above 1 (two unit vectors whose squared norms round above
    # 1): the normal-incidence value, not NaN (the reference takes sqrt of a
    # negative number here and the photon continues as NaN)
    for n1, n2 in [(1.4, 1.0), (1.0, 1.4), (1.4, 1.4)]:
        for c in (1.0000001, -1.0000001):
            r = O.fresnel(n1, n2, c)
            assert abs(r - ((n1 - n2) / (n1 + n2)) ** 2) < 1e-7

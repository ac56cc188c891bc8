"""The maternal-fetal mesh generator (lpm_mf_build) on the CPU: face counts,
every surface closed (each edge used once in each direction), normals
pointing out of the tissue they bound, the tissue tags, and the optional
fetal torso ellipsoid inside the amniotic cavity and clear of the head."""
import numpy as np
import pytest

from lightpropagation_amd import scenes as S


def _surfaces(tri, med):
    """Groups faces by (inner, outer) tag in the generator's emission order."""
    out = {}
    for k in range(len(tri)):
        out.setdefault((int(med[k, 0]), int(med[k, 1])), []).append(k)
    return out


def _closed(tri_sub):
    """Each directed edge appears exactly once and its reverse exactly once."""
    v = np.round(tri_sub.reshape(-1, 3, 3).astype(np.float64), 6)
    edges = {}
    for f in v:
        for a, b in ((0, 1), (1, 2), (2, 0)):
            e = (tuple(f[a]), tuple(f[b]))
            edges[e] = edges.get(e, 0) + 1
    return all(n == 1 and edges.get((e[1], e[0]), 0) == 1 for e, n in edges.items())


@pytest.mark.parametrize("regions", [10, 7])
@pytest.mark.parametrize("torso", [0, 6])
def test_mesh_counts_tags_closure_and_normals(torso, regions):
    tri, nrm, med, p = S.mf_mesh(freq_maternal=5, freq_fetal=4, freq_torso=torso, regions=regions)
    head = 5 if regions == 10 else 3
    assert len(tri) == 4 * 20 * 25 + head * 20 * 16 + 20 * torso * torso
    surf = _surfaces(tri, med)
    want = {(1, 0), (2, 1), (3, 2), (4, 3), (5, 4), (6, 5), (7, 6)}
    if regions == 10:
        want |= {(8, 7), (9, 8)} | ({(10, 4)} if torso else set())
    assert set(surf) == want
    torso_tag = (10, 4) if regions == 10 else (5, 4)
    cm = np.array([0.0, 0.0, p.r_maternal])
    ch = np.array([0.0, 0.0, p.fetal_depth + p.r_head])
    tc = np.array(p.torso_center[:])
    for tag, ks in surf.items():
        f = tri[ks].reshape(-1, 3, 3).astype(np.float64)
        c = f.mean(1)
        if tag == (5, 4) and torso and regions == 7:
            # the head's outer sphere and the torso share the tag: split them
            in_head = np.linalg.norm(c - ch, axis=1) < p.r_head + 0.5
            parts = [f[in_head], f[~in_head]]
            centre = np.where(in_head[:, None], ch, tc)
        else:
            parts = [f]
            centre = tc if tag == torso_tag and torso and regions == 10 else (cm if tag[0] <= 4 else ch)
        for part in parts:
            assert _closed(part.reshape(len(part), 9)), tag
        # normals point away from the centre of the shape they bound (outward)
        n = nrm[ks].astype(np.float64)
        assert np.all(np.einsum("ij,ij->i", n, c - centre) > 0), tag
        assert np.allclose(np.linalg.norm(n, axis=1), 1.0, atol=1e-6)
    if regions == 10:  # the head's shells at the stated radii
        r = [p.r_head, p.r_head - p.t_scalp, p.r_head - p.t_scalp - p.t_skull,
             p.r_head - p.t_scalp - p.t_skull - p.t_csf, p.r_head - p.t_scalp - p.t_skull - p.t_csf - p.t_grey]
        for tag, rr in zip([(5, 4), (6, 5), (7, 6), (8, 7), (9, 8)], r):
            ks = [k for k in surf[tag] if np.linalg.norm(tri[k, :3] - ch) < p.r_head + 0.5]
            d = np.linalg.norm(tri[ks].reshape(-1, 3).astype(np.float64) - ch, axis=1)
            assert np.allclose(d, rr, atol=1e-5), tag


def test_torso_lies_in_the_cavity_and_clear_of_the_head():
    tri, nrm, med, p = S.mf_mesh(freq_maternal=4, freq_fetal=4, freq_torso=12)
    f = tri[-20 * 144:].reshape(-1, 3).astype(np.float64)
    assert np.all(med[-20 * 144:] == [10, 4])  # the fetal body, its own region
    r_cav = p.r_maternal - p.t_skin - p.t_fat - p.t_uterus
    assert np.all(np.linalg.norm(f - [0, 0, p.r_maternal], axis=1) < r_cav)
    assert np.all(np.linalg.norm(f - [0, 0, p.fetal_depth + p.r_head], axis=1) > p.r_head)
    # the vertices lie on the ellipsoid
    q = (f - np.array(p.torso_center[:])) / np.array(p.torso_axes[:])
    assert np.allclose(np.linalg.norm(q, axis=1), 1.0, atol=1e-5)
    # the torso follows the head down: 0.5 cm below it at every C2 depth
    for depth, kw in ((1.5, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, {}), (3.5, {}), (4.5, {})):
        _, _, _, q = S.mf_mesh(freq_maternal=4, freq_fetal=4, freq_torso=4, fetal_depth=depth, **kw)
        assert q.torso_center[2] == depth + 2 * q.r_head + 0.5 + q.torso_axes[2]


def test_misplaced_torso_is_rejected():
    with pytest.raises(ValueError):
        S.mf_mesh(freq_maternal=4, freq_fetal=4, freq_torso=4, torso_center=(0.0, 0.0, 8.0))  # in the head
    with pytest.raises(ValueError):
        S.mf_mesh(freq_maternal=4, freq_fetal=4, freq_torso=4, torso_center=(0.0, 0.0, 26.0))  # past the cavity


def test_large_mesh_size():
    # the 2e6-face mesh: frequency-118 maternal spheres, frequency-92 fetal
    # head spheres and a frequency-40 torso; the default 10-region mesh
    import ctypes as C
    p = S.MFParams()
    S.lib().lpm_mf_default(C.byref(p))
    assert p.regions == 10
    assert S.lib().lpm_mf_num_triangles(C.byref(p)) == 4 * 20 * 38 ** 2 + 5 * 20 * 27 ** 2 + 20 * 20 ** 2 == 196_420
    p.freq_maternal, p.freq_fetal, p.freq_torso = 118, 92, 40
    n = 4 * 20 * 118 * 118 + 5 * 20 * 92 * 92 + 20 * 40 * 40
    assert S.lib().lpm_mf_num_triangles(C.byref(p)) == n == 1_992_320
    p.regions, p.freq_fetal = 7, 118  # the seven-region 2e6-face mesh
    assert S.lib().lpm_mf_num_triangles(C.byref(p)) == 7 * 20 * 118 * 118 + 20 * 40 * 40 == 1_981_360


def test_geometries_merge():
    """scenes.geometries concatenates the meshes in order and refuses scenes
    that differ in anything but the mesh."""
    a = S.maternal_fetal(fetal_depth=2.5, freq_maternal=6, freq_fetal=6)
    b = S.maternal_fetal(fetal_depth=3.5, freq_maternal=8, freq_fetal=8)
    m, counts = S.geometries([a, b])
    assert counts == [len(a["tri"]), len(b["tri"])] and counts[0] != counts[1]
    assert np.array_equal(m["tri"][:counts[0]], a["tri"]) and np.array_equal(m["tri"][counts[0]:], b["tri"])
    assert np.array_equal(m["med"][counts[0]:], b["med"]) and np.array_equal(m["nrm"][:counts[0]], a["nrm"])
    assert np.array_equal(m["pstart"], a["pstart"])
    with pytest.raises(ValueError):
        S.geometries([a, S.slab(S.C0_LAYERS)])
    c = dict(b)
    c["chance"] = 0.2
    with pytest.raises(ValueError):
        S.geometries([a, c])

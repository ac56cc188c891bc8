"""Scene construction through the model layer of liblp.so
(include/lightprop_model.h): the synthetic benchmark scenes.

A scene is a dict: tri (T, 9), nrm (T, 3), med (T, 2), nmuasg (R, 4),
fetal_id, pstart, vstart (A, 3), vocc (A,), poi_basis (12,), model_dims,
det_grid, factor, chance, w_th.
"""
import ctypes as C

import numpy as np

from ._capi import fp, ip, lib

MF_TISSUES = ["ambient", "maternal skin", "fat", "uterine wall", "amniotic fluid", "fetal scalp",
              "fetal skull", "CSF", "grey matter", "white matter", "fetal body"]
MF_TISSUES_7 = MF_TISSUES[:7] + ["fetal brain"]  # the regions=7 layout
MF_FETAL_ID = 5


class MFParams(C.Structure):
    _fields_ = [("r_maternal", C.c_double), ("t_skin", C.c_double), ("t_fat", C.c_double),
                ("t_uterus", C.c_double), ("fetal_depth", C.c_double), ("r_head", C.c_double),
                ("t_scalp", C.c_double), ("t_skull", C.c_double), ("freq_maternal", C.c_int32),
                ("freq_fetal", C.c_int32), ("freq_torso", C.c_int32), ("torso_axes", C.c_double * 3),
                ("torso_center", C.c_double * 3), ("regions", C.c_int32), ("t_csf", C.c_double),
                ("t_grey", C.c_double)]


def _chk(rc, what):
    if rc != 0:
        raise ValueError("%s failed (%d)" % (what, rc))


def emitter(p0, heading, distance):
    p0 = np.ascontiguousarray(p0, np.float32)
    h = np.ascontiguousarray(heading, np.float32)
    ps = np.zeros(3, np.float32)
    pb = np.zeros(12, np.float32)
    _chk(lib().lpm_emitter(fp(p0), fp(h), float(distance), fp(ps), fp(pb)), "lpm_emitter")
    return ps, pb


def mf_optics(wavelength_nm=660.0, spo2_maternal=0.97, spo2_fetal=0.50, regions=10):
    """(n, mua, mus, g) of the maternal-fetal tissues: 11 rows (regions 10)
    or 8 (regions 7)."""
    nt = 11 if regions == 10 else 8
    out = np.zeros((nt, 4), np.float32)
    _chk(lib().lpm_mf_optics(float(wavelength_nm), float(spo2_maternal), float(spo2_fetal), fp(out), nt),
         "lpm_mf_optics")
    return out


def mf_mesh(**kw):
    p = MFParams()
    lib().lpm_mf_default(C.byref(p))
    for k, v in kw.items():
        if k in ("torso_axes", "torso_center"):
            v = (C.c_double * 3)(*v)
        setattr(p, k, v)
    if "torso_center" not in kw:  # 0.5 cm below the fetal head, whatever its depth
        p.torso_center[2] = p.fetal_depth + 2 * p.r_head + 0.5 + p.torso_axes[2]
    n = lib().lpm_mf_num_triangles(C.byref(p))
    tri = np.zeros((n, 9), np.float32)
    nrm = np.zeros((n, 3), np.float32)
    med = np.zeros((n, 2), np.int32)
    _chk(lib().lpm_mf_build(C.byref(p), fp(tri), fp(nrm), ip(med), n), "lpm_mf_build")
    return tri, nrm, med, p


def maternal_fetal(wavelength_nm=660.0, spo2_maternal=0.97, spo2_fetal=0.50, chance=0.1, w_th=1e-4, **mesh_kw):
    """Maternal-fetal abdomen (about 2e5 faces by default: 10 regions),
    pencil beam at the origin heading +z into the body.  regions=7 gives
    the earlier layout (skin, fat, uterus, fluid, scalp, skull, brain; seven
    spheres of frequency 38 unless given otherwise, 202 160 faces)."""
    if mesh_kw.get("regions", 10) == 7:
        mesh_kw.setdefault("freq_fetal", 38)
        mesh_kw.setdefault("freq_torso", 0)
    tri, nrm, med, p = mf_mesh(**mesh_kw)
    ps, pb = emitter([0, 0, 0], [0, 0, 1], 0.002)
    D = 2 * p.r_maternal
    return dict(tri=tri, nrm=nrm, med=med, nmuasg=mf_optics(wavelength_nm, spo2_maternal, spo2_fetal, p.regions),
                fetal_id=MF_FETAL_ID, pstart=ps, vstart=np.array([[0, 0, 1]], np.float32),
                vocc=np.array([1], np.int32), poi_basis=pb, model_dims=np.array([D, D, D], np.float32),
                det_grid=np.array([0.1, 0.1, 0.1, 0.05, 1.0, 1.0], np.float32), factor=1.0, chance=chance,
                w_th=w_th)


def geometries(scenes):
    """One scene holding several geometries (for lp_set_geometries): their
    meshes concatenated in order, everything else taken from the first,
    which the others must share (source, grid, tissue count, parameters).
    Returns (scene, triangle count of each geometry)."""
    base = scenes[0]
    for sc in scenes[1:]:
        for k in ("pstart", "vstart", "vocc", "poi_basis", "model_dims", "det_grid"):
            if not np.array_equal(np.asarray(sc[k]), np.asarray(base[k])):
                raise ValueError("geometries differ in %s" % k)
        for k in ("fetal_id", "factor", "chance", "w_th"):
            if sc.get(k) != base.get(k):
                raise ValueError("geometries differ in %s" % k)
        if np.asarray(sc["nmuasg"]).shape != np.asarray(base["nmuasg"]).shape:
            raise ValueError("geometries differ in their tissue count")
    out = dict(base)
    for k in ("tri", "nrm", "med"):
        out[k] = np.ascontiguousarray(np.concatenate([np.asarray(sc[k]) for sc in scenes]))
    return out, [len(sc["tri"]) for sc in scenes]


def slab(layers, half=5.0, chance=0.1, w_th=1e-4, det_grid=(0.05, 0.05, 0.01, 0.05, 1, 1), distance=0.002):
    """Layered slab in the reference convention; layers = [(thickness, n, mua, mus, g), ...]."""
    nl = len(layers)
    th = (C.c_double * nl)(*[L[0] for L in layers])
    n = lib().lpm_slab_num_triangles(nl)
    tri = np.zeros((n, 9), np.float32)
    nrm = np.zeros((n, 3), np.float32)
    med = np.zeros((n, 2), np.int32)
    _chk(lib().lpm_slab_build(nl, th, half, half, fp(tri), fp(nrm), ip(med), n), "lpm_slab_build")
    nm = [[1.0, 0.0, 0.0, 0.0]] + [[L[1], L[2], L[3], L[4]] for L in layers]
    ps, pb = emitter([0, 0, 0], [0, 0, 1], distance)
    return dict(tri=tri, nrm=nrm, med=med, nmuasg=np.array(nm, np.float32), fetal_id=-1, pstart=ps,
                vstart=np.array([[0, 0, 1]], np.float32), vocc=np.array([1], np.int32), poi_basis=pb,
                model_dims=np.array([2 * half, 2 * half, sum(L[0] for L in layers)], np.float32),
                det_grid=np.array(det_grid, np.float32), factor=1.0, chance=chance, w_th=w_th)


# Bench configuration 0: three layers (thickness, n, mua, mus, g)
C0_LAYERS = [(0.15, 1.40, 0.4, 200.0, 0.9), (1.0, 1.44, 0.1, 120.0, 0.9), (2.0, 1.37, 0.3, 80.0, 0.9)]


def c0_closed_boxes():
    """Config 0 as three stacked closed boxes (36 faces) merged to the
    reference convention (one surface per interface, 32 faces)."""
    boxes = []
    z = 0.0
    for i, L in enumerate(C0_LAYERS):
        lo = (C.c_double * 3)(-5.0, -5.0, z)
        hi = (C.c_double * 3)(5.0, 5.0, z + L[0])
        t = np.zeros((12, 9), np.float32)
        nr = np.zeros((12, 3), np.float32)
        md = np.zeros((12, 2), np.int32)
        _chk(lib().lpm_box(lo, hi, i + 1, 0, fp(t), fp(nr), ip(md)), "lpm_box")
        boxes.append((t, nr, md))
        z += L[0]
    tri = np.ascontiguousarray(np.concatenate([b[0] for b in boxes]))
    nrm = np.ascontiguousarray(np.concatenate([b[1] for b in boxes]))
    med = np.ascontiguousarray(np.concatenate([b[2] for b in boxes]))
    n = lib().lpm_merge_coincident(fp(tri), fp(nrm), ip(med), len(tri))
    return tri[:n].copy(), nrm[:n].copy(), med[:n].copy()

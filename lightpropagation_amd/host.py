"""ctypes binding of the host layer (include/lightprop_host.h): scene
ingestion from the reference's configuration/OBJ/emitter files, the emitter
model, tally images, the Logs/Analysis post-processing, and the reference's
whole main flow (lph_run_config).

All of it runs in liblp.so; like the transport binding, this module raises
when the library has not been built.
"""
import ctypes as C

import numpy as np

from ._capi import LPError, fp, ip, lib

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_i64p = C.POINTER(C.c_int64)
_dp = C.POINTER(C.c_double)

PER_RADIUS, PER_ANGLE, PER_Z = 0, 1, 2
TALLY_SFT, TALLY_RTD, TALLY_ARZ, TALLY_PHIRZ = 0, 1, 2, 3

_SIGS = {
    "lph_last_error": (C.c_char_p, []),
    "lph_scene_load": (C.c_int, [C.c_char_p, C.POINTER(C.c_void_p)]),
    "lph_scene_free": (None, [C.c_void_p]),
    "lph_scene_sizes": (C.c_int, [C.c_void_p, _i64p, _ip, _ip]),
    "lph_scene_mesh": (C.c_int, [C.c_void_p, _fp, _fp, _ip]),
    "lph_scene_materials": (C.c_int, [C.c_void_p, _fp, _ip]),
    "lph_scene_source": (C.c_int, [C.c_void_p, _fp, _fp, _ip, _fp]),
    "lph_scene_params": (C.c_int, [C.c_void_p, _fp, _fp, _fp, _fp, _fp, _i64p]),
    "lph_scene_apply": (C.c_int, [C.c_void_p, C.c_void_p, C.c_uint64, C.c_int32, C.c_int64]),
    "lph_emitter_model": (C.c_int, [_fp, _fp, C.c_float, _fp, _fp, C.c_int32, C.c_int64, _fp, _fp, _ip, _fp,
                                    _fp, _fp]),
    "lph_image": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, _ip, _ip, _fp]),
    "lph_scale_per_radius_1d": (C.c_float, [_fp, C.c_int64, C.c_int32]),
    "lph_scale_per_angle_1d": (C.c_float, [_fp, C.c_int64, C.c_int32]),
    "lph_scale_2d": (C.c_float, [_fp, C.c_int64, C.c_int32, C.c_int32]),
    "lph_profile": (C.c_int, [C.c_int32, _fp, _ip, _fp, C.c_int64, _ip, _fp, _fp]),
    "lph_surface_sums": (C.c_int, [_fp, _ip, _fp]),
    "lph_run_config": (C.c_int, [C.c_char_p, C.c_int64, C.c_uint64, C.c_int32, C.c_char_p, _dp]),
}
_bound = None


def hlib():
    global _bound
    if _bound is None:
        L = lib()
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _bound = L
    return _bound


def _chk(rc, what):
    if rc != 0:
        raise LPError("%s failed (%d): %s" % (what, rc, hlib().lph_last_error().decode()))


def load_scene(config_path):
    """The scene dict (same keys as lightpropagation_amd.scenes) of a
    reference configuration file, read by the C++ importer."""
    L = hlib()
    s = C.c_void_p()
    _chk(L.lph_scene_load(str(config_path).encode(), C.byref(s)), "lph_scene_load")
    try:
        nt, nr, na = C.c_int64(), C.c_int32(), C.c_int32()
        _chk(L.lph_scene_sizes(s, C.byref(nt), C.byref(nr), C.byref(na)), "lph_scene_sizes")
        tri = np.zeros((nt.value, 9), np.float32)
        nrm = np.zeros((nt.value, 3), np.float32)
        med = np.zeros((nt.value, 2), np.int32)
        _chk(L.lph_scene_mesh(s, fp(tri), fp(nrm), ip(med)), "lph_scene_mesh")
        nm = np.zeros((nr.value, 4), np.float32)
        fid = C.c_int32()
        _chk(L.lph_scene_materials(s, fp(nm), C.byref(fid)), "lph_scene_materials")
        ps = np.zeros(3, np.float32)
        vs = np.zeros((na.value, 3), np.float32)
        vo = np.zeros(na.value, np.int32)
        pb = np.zeros(12, np.float32)
        _chk(L.lph_scene_source(s, fp(ps), fp(vs), ip(vo), fp(pb)), "lph_scene_source")
        md = np.zeros(3, np.float32)
        dg = np.zeros(6, np.float32)
        fac, ch, wt, nph = C.c_float(), C.c_float(), C.c_float(), C.c_int64()
        _chk(L.lph_scene_params(s, fp(md), fp(dg), C.byref(fac), C.byref(ch), C.byref(wt), C.byref(nph)),
             "lph_scene_params")
    finally:
        L.lph_scene_free(s)
    return dict(tri=tri, nrm=nrm, med=med, nmuasg=nm, fetal_id=fid.value, pstart=ps, vstart=vs, vocc=vo,
                poi_basis=pb, model_dims=md, det_grid=dg, factor=np.float32(fac.value),
                chance=np.float32(ch.value), w_th=np.float32(wt.value), nphotons=nph.value)


def emitter_model(p0, heading, distance, angles_deg=(0.0,), rri=(1.0,), nphotons=1):
    """(pstart, vstart (A, 3), vocc (A,), ex, ey, ez) of the emitter model."""
    a = np.ascontiguousarray(angles_deg, np.float32)
    r = np.ascontiguousarray(rri, np.float32)
    ps = np.zeros(3, np.float32)
    vs = np.zeros((len(a), 3), np.float32)
    vo = np.zeros(len(a), np.int32)
    ex, ey, ez = (np.zeros(3, np.float32) for _ in range(3))
    _chk(hlib().lph_emitter_model(fp(np.asarray(p0, np.float32)), fp(np.asarray(heading, np.float32)),
                                  float(distance), fp(a), fp(r), len(a), int(nphotons), fp(ps), fp(vs), ip(vo),
                                  fp(ex), fp(ey), fp(ez)), "lph_emitter_model")
    return ps, vs, vo, ex, ey, ez


def image(ctx, which, component=-1):
    """(image (ny, nx[, ncomp]) float32, extent [x0, x1, y0, y1]) of a tally."""
    ext = np.zeros(4, np.int32)
    nc = C.c_int32()
    _chk(hlib().lph_image(ctx._c, which, component, ip(ext), C.byref(nc), None), "lph_image")
    ny, nx = ext[3] - ext[2] + 1, ext[1] - ext[0] + 1
    out = np.zeros((ny, nx, nc.value), np.float32)
    _chk(hlib().lph_image(ctx._c, which, component, ip(ext), C.byref(nc), fp(out)), "lph_image")
    return (out[..., 0] if nc.value == 1 else out), ext


def scale_per_radius_1d(det_grid, n, ir):
    return hlib().lph_scale_per_radius_1d(fp(np.asarray(det_grid, np.float32)), int(n), int(ir))


def scale_per_angle_1d(det_grid, n, ia):
    return hlib().lph_scale_per_angle_1d(fp(np.asarray(det_grid, np.float32)), int(n), int(ia))


def scale_2d(det_grid, n, ir, ia):
    return hlib().lph_scale_2d(fp(np.asarray(det_grid, np.float32)), int(n), int(ir), int(ia))


def profile(kind, img, extent, det_grid, n):
    """The two columns the reference's _perRadius / _perAngle / _perZ file
    holds for a one-component image."""
    img = np.ascontiguousarray(img, np.float32)
    ext = np.ascontiguousarray(extent, np.int32)
    dg = np.ascontiguousarray(det_grid, np.float32)
    rows = C.c_int32()
    _chk(hlib().lph_profile(kind, fp(img), ip(ext), fp(dg), int(n), C.byref(rows), None, None), "lph_profile")
    c0 = np.zeros(rows.value, np.float32)
    c1 = np.zeros(rows.value, np.float32)
    _chk(hlib().lph_profile(kind, fp(img), ip(ext), fp(dg), int(n), C.byref(rows), fp(c0), fp(c1)), "lph_profile")
    return c0, c1


def surface_sums(sft_img, extent):
    """Analysis::CalculateSurfaceSums over a 3-component SFT image: (R, F, T)."""
    img = np.ascontiguousarray(sft_img, np.float32)
    out = np.zeros(3, np.float32)
    _chk(hlib().lph_surface_sums(fp(img), ip(np.ascontiguousarray(extent, np.int32)), fp(out)), "lph_surface_sums")
    return out


def run_config(config_path, nphotons=0, seed=1, device=0, out_dir=None):
    """The reference's main flow on a configuration file; returns (R, F, T)
    as fractions of the photons run."""
    sums = (C.c_double * 3)()
    _chk(hlib().lph_run_config(str(config_path).encode(), int(nphotons), int(seed), int(device),
                               out_dir.encode() if out_dir else None, sums), "lph_run_config")
    return tuple(sums)

"""ctypes binding of include/lightprop.h (liblp.so).

The library is the product; this module only marshals numpy arrays across
the C ABI.  It raises if liblp.so has not been built -- there is no fallback.
"""
import ctypes as C
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(HERE, "liblp.so")

LP_OK = 0
RNG_PHILOX, RNG_RAN3 = 0, 1
TALLY = {"sft": 0, "rtd": 1, "arz": 2, "phirz": 3, "region": 4, "vox": 5, "rad": 6, "det": 7, "misc": 8}
MISC = ["photons", "steps", "interact", "boundary", "escaped", "capped", "arz_out", "vox_out",
        "node_visits", "tri_tests", "skipped", "nan", "dist_queries", "dist_visits", "list_tests", "wave_iters", "search_rounds",
        "bvh_rounds", "replay_mismatch", "marches", "crawl_replayed"]
OPTIONS = {"grid_blocks": 0, "lds_nodes": 1, "safety_cells": 2, "dist_cap": 3, "dist_cap_steps": 4,
           "blocks_per_cu": 5, "free_iters": 6, "waves_per_simd": 7, "near_r0": 8, "near_cell": 9, "search_min": 10,
           "list_min": 11, "refill_min": 12, "graph": 13, "clear_step": 14, "boundary_run": 15, "list_long": 16,
           "no_skip": 17, "march": 18, "second_step": 19, "crawl": 20}
FIXED = 2.0 ** 32
FIXED_PHI = 2.0 ** 24

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("dimx", "dimy", "dimz", "dimR", "dimAlpha", "sft_nx", "sft_ny",
                                         "arz_nr", "arz_nz")] + \
               [("dist_min", C.c_float), ("rsp", C.c_float), ("n", C.c_int64 * 9)]


class TuneResult(C.Structure):
    _fields_ = [("sample_photons", C.c_int64), ("trials", C.c_int32), ("ms_before", C.c_double),
                ("ms_after", C.c_double), ("waves", C.c_int32), ("blocks_per_cu", C.c_int32),
                ("free_iters", C.c_int32), ("refill_min", C.c_int32), ("list_min", C.c_double),
                ("search_min", C.c_double)]


class Stats(C.Structure):
    _fields_ = [("kernel_ms", C.c_double), ("photons", C.c_int64), ("bvh_nodes", C.c_int64),
                ("bvh_depth", C.c_int32), ("grid_blocks", C.c_int32), ("device_bytes", C.c_int64),
                ("replay_ms", C.c_double), ("replayed", C.c_int64), ("replay_mismatch", C.c_int64)]


class LPError(RuntimeError):
    pass


_lib = None

_SIGS = {
    "lp_create": (C.c_int, [C.POINTER(C.c_void_p), C.c_int32]),
    "lp_destroy": (None, [C.c_void_p]),
    "lp_last_error": (C.c_char_p, [C.c_void_p]),
    "lp_set_mesh": (C.c_int, [C.c_void_p, _fp, _fp, _ip, C.c_int64]),
    "lp_set_materials": (C.c_int, [C.c_void_p, _fp, C.c_int32, C.c_int32]),
    "lp_set_source": (C.c_int, [C.c_void_p, _fp, _fp, _ip, C.c_int32, _fp]),
    "lp_set_grid": (C.c_int, [C.c_void_p, _fp, _fp]),
    "lp_set_params": (C.c_int, [C.c_void_p, C.c_float, C.c_float, C.c_float, C.c_uint64, C.c_int32, C.c_int64]),
    "lp_set_voxel_grid": (C.c_int, [C.c_void_p, _fp, _fp, _ip]),
    "lp_set_radial": (C.c_int, [C.c_void_p, C.c_float, C.c_int32]),
    "lp_set_detectors": (C.c_int, [C.c_void_p, _fp, _fp, C.c_int32]),
    "lp_set_tallies": (C.c_int, [C.c_void_p, C.c_uint32, C.c_int32]),
    "lp_get_layout": (C.c_int, [C.c_void_p, C.POINTER(Layout)]),
    "lp_reset_tallies": (C.c_int, [C.c_void_p]),
    "lp_run": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64]),
    "lp_get_tally": (C.c_int, [C.c_void_p, C.c_int32, _lp, C.c_int64]),
    "lp_get_exit_records": (C.c_int, [C.c_void_p, _fp, C.c_int64]),
    "lp_trace": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.c_int32, _fp, _ip]),
    "lp_get_stats": (C.c_int, [C.c_void_p, C.POINTER(Stats)]),
    "lp_set_jacobian_grid": (C.c_int, [C.c_void_p, _fp, _fp, _ip]),
    "lp_adjoint": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _lp, _lp, _lp, _lp]),
    "lp_adjoint_voxels": (C.c_int, [C.c_void_p, C.c_int64, C.c_int64, _lp, _lp, _lp, _lp, _lp]),
    "lp_set_adjoint_paths": (C.c_int, [C.c_void_p, C.c_int32]),
    "lp_get_adjoint_paths": (C.c_int, [C.c_void_p, C.c_int64, _lp, _ip, _fp, _fp, C.c_void_p]),
    "lp_run_batch": (C.c_int, [C.c_void_p, _fp, C.c_int32, C.c_int64, C.c_int64, _lp, _lp]),
    "lp_set_geometries": (C.c_int, [C.c_void_p, _lp, C.c_int32]),
    "lp_run_batch_geom": (C.c_int, [C.c_void_p, _fp, C.c_int32, _ip, C.c_int64, C.c_int64, _lp, _lp]),
    "lp_run_batch_noalloc": (C.c_int, [C.c_void_p, _fp, C.c_int32, _ip, C.c_int64, C.c_int64, C.c_void_p,
                                       C.c_void_p]),
    "lp_autotune": (C.c_int, [C.c_void_p, C.c_int64, C.c_int32, C.POINTER(TuneResult)]),
    "lp_debug_math": (C.c_int, [C.c_void_p, C.c_int32, _fp, _fp, C.c_int64]),
    "lp_set_launch": (C.c_int, [C.c_void_p, C.c_int32, C.c_int32, C.c_int32]),
    "lp_set_option": (C.c_int, [C.c_void_p, C.c_int32, C.c_double]),
    "lp_debug_copy": (C.c_int, [C.c_void_p, C.c_int64, C.POINTER(C.c_double)]),
    "lpm_mf_default": (None, [C.c_void_p]),
    "lpm_mf_num_triangles": (C.c_int64, [C.c_void_p]),
    "lpm_mf_build": (C.c_int, [C.c_void_p, _fp, _fp, _ip, C.c_int64]),
    "lpm_mf_optics": (C.c_int, [C.c_double, C.c_double, C.c_double, _fp, C.c_int32]),
    "lpm_slab_num_triangles": (C.c_int64, [C.c_int32]),
    "lpm_slab_build": (C.c_int, [C.c_int32, C.POINTER(C.c_double), C.c_double, C.c_double, _fp, _fp, _ip,
                                 C.c_int64]),
    "lpm_merge_coincident": (C.c_int64, [_fp, _fp, _ip, C.c_int64]),
    "lpm_box": (C.c_int, [C.POINTER(C.c_double), C.POINTER(C.c_double), C.c_int32, C.c_int32, _fp, _fp, _ip]),
    "lpm_emitter": (C.c_int, [_fp, _fp, C.c_float, _fp, _fp]),
    "lpm_cell_normals": (C.c_int, [_fp, C.c_int64, _fp]),
}


def lib():
    """Load liblp.so (raises if it has not been built)."""
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise LPError("liblp.so not built: run python -c 'import __graft_entry__ as g; g.build()'")
        L = C.CDLL(LIB_PATH)
        for name, (res, args) in _SIGS.items():
            f = getattr(L, name)
            f.restype = res
            f.argtypes = args
        _lib = L
    return _lib


def _f(a):
    return np.ascontiguousarray(a, np.float32)


def _i(a):
    return np.ascontiguousarray(a, np.int32)


def fp(a):
    return a.ctypes.data_as(_fp)


def ip(a):
    return a.ctypes.data_as(_ip)


class Context:
    """One lp_context on one device."""

    def __init__(self, device=0):
        self._c = C.c_void_p()
        rc = lib().lp_create(C.byref(self._c), device)
        if rc != LP_OK:
            raise LPError("lp_create(device=%d) failed with %d (no HIP device?)" % (device, rc))
        self._keep = {}

    def close(self):
        if self._c:
            lib().lp_destroy(self._c)
            self._c = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def _chk(self, rc, what):
        if rc != LP_OK:
            raise LPError("%s failed (%d): %s" % (what, rc, lib().lp_last_error(self._c).decode()))

    # --- configuration ----------------------------------------------------
    def set_mesh(self, tri, nrm, med):
        tri, nrm, med = _f(tri).reshape(-1), _f(nrm).reshape(-1), _i(med).reshape(-1)
        self._chk(lib().lp_set_mesh(self._c, fp(tri), fp(nrm), ip(med), len(tri) // 9), "lp_set_mesh")

    def set_geometries(self, tri_counts):
        """lp_set_geometries: the mesh is consecutive geometries of these
        triangle counts, each searched on its own."""
        n = np.ascontiguousarray(tri_counts, np.int64)
        self._chk(lib().lp_set_geometries(self._c, n.ctypes.data_as(_lp), len(n)), "lp_set_geometries")

    def set_materials(self, nmuasg, fetal_id):
        nm = _f(nmuasg).reshape(-1)
        self._chk(lib().lp_set_materials(self._c, fp(nm), len(nm) // 4, int(fetal_id)), "lp_set_materials")
        self._keep["ntissue"] = len(nm) // 4

    def set_source(self, pstart, vstart, vocc, poi_basis):
        ps, vs, vo, pb = _f(pstart), _f(vstart).reshape(-1), _i(vocc).reshape(-1), _f(poi_basis)
        self._chk(lib().lp_set_source(self._c, fp(ps), fp(vs), ip(vo), len(vo), fp(pb)), "lp_set_source")

    def set_grid(self, model_dims, det_grid):
        self._chk(lib().lp_set_grid(self._c, fp(_f(model_dims)), fp(_f(det_grid))), "lp_set_grid")

    def set_params(self, chance, w_th, factor=1.0, seed=1, rng=RNG_PHILOX, max_steps=0):
        self._chk(lib().lp_set_params(self._c, chance, w_th, factor, int(seed) & (2 ** 64 - 1), rng, max_steps),
                  "lp_set_params")

    def set_voxel_grid(self, origin, size, n):
        self._chk(lib().lp_set_voxel_grid(self._c, fp(_f(origin)), fp(_f(size)), ip(_i(n))), "lp_set_voxel_grid")

    def set_radial(self, dr, n):
        self._chk(lib().lp_set_radial(self._c, dr, n), "lp_set_radial")

    def set_detectors(self, rlo, rhi):
        lo, hi = _f(rlo), _f(rhi)
        self._chk(lib().lp_set_detectors(self._c, fp(lo), fp(hi), len(lo)), "lp_set_detectors")
        self._keep["ndet"] = len(lo)

    def set_tallies(self, mask=0xFFFFFFFF, record_exit=False):
        self._chk(lib().lp_set_tallies(self._c, mask, int(bool(record_exit))), "lp_set_tallies")

    def set_launch(self, grid_blocks=0, lds_nodes=-1, safety_cells=-1):
        self._chk(lib().lp_set_launch(self._c, grid_blocks, lds_nodes, safety_cells), "lp_set_launch")

    def set_option(self, key, value):
        k = OPTIONS[key] if isinstance(key, str) else int(key)
        self._chk(lib().lp_set_option(self._c, k, float(value)), "lp_set_option(%s)" % key)

    def set_scene(self, sc, seed=1, rng=RNG_PHILOX, max_steps=0):
        """Configure from a scene dict (tri, nrm, med, nmuasg, fetal_id, pstart,
        vstart, vocc, poi_basis, model_dims, det_grid, factor, chance, w_th)."""
        self.set_mesh(sc["tri"], sc["nrm"], sc["med"])
        self.set_materials(sc["nmuasg"], sc.get("fetal_id", -1))
        self.set_source(sc["pstart"], sc["vstart"], sc["vocc"], sc["poi_basis"])
        self.set_grid(sc["model_dims"], sc["det_grid"])
        self.set_params(float(sc["chance"]), float(sc["w_th"]), float(sc.get("factor", 1.0)), seed, rng, max_steps)

    # --- execution ----------------------------------------------------------
    def layout(self):
        L = Layout()
        self._chk(lib().lp_get_layout(self._c, C.byref(L)), "lp_get_layout")
        return L

    def reset(self):
        self._chk(lib().lp_reset_tallies(self._c), "lp_reset_tallies")

    def run(self, first, count):
        self._chk(lib().lp_run(self._c, int(first), int(count)), "lp_run")

    def tally(self, name):
        k = TALLY[name]
        n = self.layout().n[k]
        out = np.zeros(max(n, 0), np.int64)
        self._chk(lib().lp_get_tally(self._c, k, out.ctypes.data_as(_lp), n), "lp_get_tally")
        return out

    def tallies(self):
        out = {k: self.tally(k) for k in TALLY if k != "misc"}
        m = self.tally("misc")
        out["misc"] = dict(zip(MISC, [int(x) for x in m[:len(MISC)]]))
        return out

    def exit_records(self, n):
        out = np.zeros((n, 8), np.float32)
        self._chk(lib().lp_get_exit_records(self._c, fp(out), n), "lp_get_exit_records")
        return out

    def trace(self, first, count, max_rec):
        out = np.zeros((count, max_rec, 12), np.float32)
        nrec = np.zeros(count, np.int32)
        self._chk(lib().lp_trace(self._c, first, count, max_rec, fp(out), ip(nrec)), "lp_trace")
        return out, nrec

    def set_jacobian_grid(self, origin, size, n):
        self._jac_n = [int(x) for x in n]
        self._chk(lib().lp_set_jacobian_grid(self._c, fp(_f(origin)), fp(_f(size)), ip(_i(n))),
                  "lp_set_jacobian_grid")

    def adjoint(self, first, count, jac=False, ntissue=None, ndet=None, jac_mus=False, paths=False):
        """lp_adjoint (lp_adjoint_voxels with jac_mus): dict(signals (ndet,),
        grad (ndet, ntissue, 2), jac (ndet, nz, ny, nx) or None, jac_mus
        (same shape) or None, pairs) as int64 fixed point (2^32 per unit);
        the photons also add to the tallies.  paths: also the per-pair path
        records (lp_set_adjoint_paths / lp_get_adjoint_paths) as dict(photon,
        det, w, length (npairs, ntissue), interactions (npairs, ntissue))
        under "paths"."""
        nt = ntissue if ntissue is not None else self._keep["ntissue"]
        nd = ndet if ndet is not None else self._keep["ndet"]
        sig = np.zeros(nd, np.int64)
        grad = np.zeros((nd, nt, 2), np.int64)
        J = None
        if jac:
            nx, ny, nz = getattr(self, "_jac_n", (0, 0, 0))
            J = np.zeros((nd, nz, ny, nx), np.int64)
        Js = None
        if jac_mus:
            nx, ny, nz = getattr(self, "_jac_n", (0, 0, 0))
            Js = np.zeros((nd, nz, ny, nx), np.int64)
        pairs = C.c_int64()
        self._chk(lib().lp_set_adjoint_paths(self._c, int(bool(paths))), "lp_set_adjoint_paths")
        self._chk(lib().lp_adjoint_voxels(self._c, int(first), int(count), sig.ctypes.data_as(_lp),
                                          grad.ctypes.data_as(_lp), J.ctypes.data_as(_lp) if J is not None else None,
                                          Js.ctypes.data_as(_lp) if Js is not None else None, C.byref(pairs)),
                  "lp_adjoint_voxels")
        out = dict(signals=sig, grad=grad, jac=J, jac_mus=Js, pairs=pairs.value)
        if paths:
            n = pairs.value
            P = dict(photon=np.zeros(n, np.int64), det=np.zeros(n, np.int32), w=np.zeros(n, np.float32),
                     length=np.zeros((n, nt), np.float32), interactions=np.zeros((n, nt), np.uint32))
            self._chk(lib().lp_get_adjoint_paths(self._c, n, P["photon"].ctypes.data_as(_lp), ip(P["det"]),
                                                 fp(P["w"]), fp(P["length"]), P["interactions"].ctypes.data),
                      "lp_get_adjoint_paths")
            self._chk(lib().lp_set_adjoint_paths(self._c, 0), "lp_set_adjoint_paths")
            out["paths"] = P
        return out

    def run_batch(self, nmuasg, first, count, geom=None):
        """lp_run_batch (geom None) or lp_run_batch_geom (geom: geometry of
        each instance): nmuasg (ninst, ntissue, 4) -> dict(det (ninst, ndet,
        2), region (ninst, ntissue)) as int64 fixed point (2^32 per unit)."""
        m = np.ascontiguousarray(nmuasg, np.float32)
        if m.ndim != 3 or m.shape[2] != 4:
            raise ValueError("nmuasg must be (ninst, ntissue, 4)")
        ni, nt = m.shape[0], m.shape[1]
        nd = self._keep.get("ndet", 0)
        det = np.zeros((ni, max(nd, 0), 2), np.int64)
        reg = np.zeros((ni, nt), np.int64)
        dp = det.ctypes.data_as(_lp) if nd else None
        if geom is None:
            self._chk(lib().lp_run_batch(self._c, fp(m), ni, int(first), int(count), dp, reg.ctypes.data_as(_lp)),
                      "lp_run_batch")
        else:
            g = np.ascontiguousarray(geom, np.int32)
            if g.shape != (ni,):
                raise ValueError("geom must hold one geometry per instance")
            self._chk(lib().lp_run_batch_geom(self._c, fp(m), ni, ip(g), int(first), int(count), dp,
                                              reg.ctypes.data_as(_lp)), "lp_run_batch_geom")
        return dict(det=det, region=reg)

    def autotune(self, sample=2_000_000, passes=2):
        """lp_autotune: returns the TuneResult (options left on the context)."""
        r = TuneResult()
        self._chk(lib().lp_autotune(self._c, int(sample), int(passes), C.byref(r)), "lp_autotune")
        return r

    def stats(self):
        s = Stats()
        self._chk(lib().lp_get_stats(self._c, C.byref(s)), "lp_get_stats")
        return s

    def debug_copy(self, nbytes):
        ms = C.c_double()
        self._chk(lib().lp_debug_copy(self._c, int(nbytes), C.byref(ms)), "lp_debug_copy")
        return ms.value

    def debug_math(self, op, x, y=None):
        x = _f(x)
        y = np.zeros_like(x) if y is None else _f(y)
        xy = np.ascontiguousarray(np.stack([x, y], axis=1), np.float32)
        out = np.zeros(len(x), np.float32)
        self._chk(lib().lp_debug_math(self._c, op, fp(xy), fp(out), len(x)), "lp_debug_math")
        return out

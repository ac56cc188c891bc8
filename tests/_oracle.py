"""ctypes binding to the CPU oracle (oracle/liboracle.so).

Test infrastructure: only tests/, __graft_entry__.smoke() and bench.py's
cpu_baseline leg load this module.  The oracle checks the GPU results; it is
never the product path.
"""
import ctypes as C
import os
import subprocess

import numpy as np

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
LIB_PATH = os.path.join(ROOT, "oracle", "liboracle.so")

RNG_PHILOX = 0
RNG_RAN3F = 1
EV_INTERACT, EV_TRANSMIT, EV_REFLECT, EV_ESCAPE = 0, 1, 2, 3
MISC = ["photons", "steps", "interact", "boundary", "escaped", "capped", "arz_out", "vox_out", "nan"]

_fp = C.POINTER(C.c_float)
_ip = C.POINTER(C.c_int32)
_lp = C.POINTER(C.c_int64)


class Params(C.Structure):
    _fields_ = [
        ("ntri", C.c_int32), ("tri", _fp), ("nrm", _fp), ("med", _ip),
        ("ntissue", C.c_int32), ("nmuasg", _fp), ("fetal_id", C.c_int32),
        ("pstart", C.c_float * 3), ("vstart", _fp), ("vocc", _ip), ("nstart", C.c_int32),
        ("poi_basis", C.c_float * 12), ("model_dims", C.c_float * 3), ("det_grid", C.c_float * 6),
        ("factor", C.c_float), ("chance", C.c_float), ("w_th", C.c_float),
        ("rng_mode", C.c_int32), ("seed", C.c_uint64), ("max_steps", C.c_int64), ("max_boundary_run", C.c_int64),
        ("vox_origin", C.c_float * 3), ("vox_size", C.c_float * 3), ("vox_n", C.c_int32 * 3),
        ("rad_dr", C.c_float), ("rad_n", C.c_int32),
        ("ndet", C.c_int32), ("det_rlo", _fp), ("det_rhi", _fp),
    ]


class Layout(C.Structure):
    _fields_ = [(n, C.c_int32) for n in ("dimx", "dimy", "dimz", "dimR", "dimAlpha",
                                         "sft_nx", "sft_ny", "arz_nr", "arz_nz")] + \
               [("dist_min", C.c_float), ("rsp", C.c_float)] + \
               [(n, C.c_int64) for n in ("n_sft", "n_rtd", "n_arz", "n_region", "n_vox", "n_rad", "n_det")]


class Tallies(C.Structure):
    _fields_ = [(n, _lp) for n in ("sft", "rtd", "arz", "phirz", "region", "vox", "rad", "det")] + \
               [("misc", C.c_int64 * 9)]


_lib = None


class AdjointOut(C.Structure):
    _fields_ = [("jac_origin", C.c_float * 3), ("jac_size", C.c_float * 3), ("jac_n", C.c_int32 * 3),
                ("signals", _lp), ("grad", _lp), ("jac", _lp), ("ndetected", C.c_int64), ("jac_mus", _lp),
                ("path_len", C.POINTER(C.c_float)), ("path_int", C.POINTER(C.c_int32)),
                ("path_w", C.POINTER(C.c_float)), ("path_r", C.POINTER(C.c_float))]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            subprocess.run(["make", "-C", os.path.join(ROOT, "oracle")], check=True,
                           capture_output=True)
        _lib = _bind(C.CDLL(LIB_PATH))
    return _lib


class planted:
    """Context manager: the oracle built with one planted mistake (the
    ORC_PLANT_* macro `macro`, see mcml_oracle.c) into `tmpdir` stands in for
    the real one, so a test can show that its checks catch that mistake."""

    def __init__(self, macro, tmpdir):
        self.path = os.path.join(str(tmpdir), "liboracle_%s.so" % macro.lower())
        src = os.path.join(ROOT, "oracle", "mcml_oracle.c")
        subprocess.run(["gcc", "-O2", "-fPIC", "-ffp-contract=off", "-fno-fast-math", "-fopenmp", "-std=c11",
                        "-shared", "-D" + macro, "-o", self.path, src, "-lm"], check=True, capture_output=True)

    def __enter__(self):
        global _lib
        self.saved = lib()
        _lib = _bind(C.CDLL(self.path))
        return _lib

    def __exit__(self, *exc):
        global _lib
        _lib = self.saved
        return False


def _bind(L):
    """Declares the argument and result types of the oracle's functions."""
    if True:
        L.orc_logf.restype = C.c_float
        L.orc_logf.argtypes = [C.c_float]
        L.orc_cosf.restype = C.c_float
        L.orc_cosf.argtypes = [C.c_float]
        L.orc_acosf.restype = C.c_float
        L.orc_acosf.argtypes = [C.c_float]
        L.orc_sign.restype = C.c_float
        L.orc_sign.argtypes = [C.c_float]
        L.orc_philox4x32_10.argtypes = [C.POINTER(C.c_uint32), C.POINTER(C.c_uint32), C.POINTER(C.c_uint32)]
        L.orc_ran3_init.argtypes = [C.c_int64, _lp, _ip, _ip]
        L.orc_ran3_next.restype = C.c_int64
        L.orc_ran3_next.argtypes = [_lp, _ip, _ip]
        L.orc_ran3f_next.restype = C.c_float
        L.orc_ran3f_next.argtypes = [_fp, _ip, _ip]
        L.orc_find_triangle.restype = C.c_int32
        L.orc_find_triangle.argtypes = [_fp, C.c_int32, _fp, _fp, C.c_int32, _fp]
        L.orc_rspecular.restype = C.c_float
        L.orc_rspecular.argtypes = [_fp, C.c_int32]
        L.orc_fresnel.restype = C.c_float
        L.orc_fresnel.argtypes = [C.c_float, C.c_float, C.c_float]
        L.orc_vscatter.argtypes = [_fp, C.c_float, C.c_float, C.c_float]
        L.orc_transmit_or_reflect.restype = C.c_int
        L.orc_transmit_or_reflect.argtypes = [_fp, _fp, C.c_float, C.c_float, C.c_float]
        L.orc_layout_compute.argtypes = [C.POINTER(Params), C.POINTER(Layout)]
        L.orc_run.restype = C.c_int
        L.orc_run.argtypes = [C.POINTER(Params), C.c_int64, C.c_int64, C.POINTER(Tallies), _fp, _fp,
                              C.c_int32, _ip, C.c_int]
        L.orc_adjoint.restype = C.c_int
        L.orc_adjoint.argtypes = [C.POINTER(Params), C.c_int64, C.c_int64, C.POINTER(Tallies),
                                  C.POINTER(AdjointOut), C.c_int]
        L.orc_dda.argtypes = [C.POINTER(AdjointOut), _fp, C.c_float, _lp]
    return L


def dda(seg, w, origin, size, n):
    """orc_dda: -(length in voxel) * w per voxel (2^32 fixed point)."""
    A = AdjointOut()
    A.jac_origin[:] = list(origin)
    A.jac_size[:] = list(size)
    A.jac_n[:] = list(n)
    jac = np.zeros(int(np.prod(n)), np.int64)
    seg = np.ascontiguousarray(seg, np.float32)
    lib().orc_dda(C.byref(A), fptr(seg), float(w), lptr(jac))
    return jac


def fptr(a):
    return a.ctypes.data_as(_fp)


def iptr(a):
    return a.ctypes.data_as(_ip)


def lptr(a):
    return a.ctypes.data_as(_lp)


# --- scalar building blocks ---------------------------------------------
def logf(x):
    return lib().orc_logf(float(x))


def cosf(x):
    return lib().orc_cosf(float(x))


def acosf(x):
    return lib().orc_acosf(float(x))


def philox(ctr, key):
    c = (C.c_uint32 * 4)(*ctr)
    k = (C.c_uint32 * 2)(*key)
    o = (C.c_uint32 * 4)()
    lib().orc_philox4x32_10(c, k, o)
    return list(o)


def ran3_init(idum):
    ma = np.zeros(56, np.int64)
    i = C.c_int32()
    j = C.c_int32()
    lib().orc_ran3_init(idum, lptr(ma), C.byref(i), C.byref(j))
    return ma, i.value, j.value


def ran3_stream(idum, n):
    ma, i, j = ran3_init(idum)
    ii, jj = C.c_int32(i), C.c_int32(j)
    return [lib().orc_ran3_next(lptr(ma), C.byref(ii), C.byref(jj)) for _ in range(n)]


def ran3f_stream(idum, n):
    ma, i, j = ran3_init(idum)
    s = ma.astype(np.float32)
    ii, jj = C.c_int32(i), C.c_int32(j)
    return np.array([lib().orc_ran3f_next(fptr(s), C.byref(ii), C.byref(jj)) for _ in range(n)],
                    np.float32)


def find_triangle(tri, p, v, ref_tri=-1):
    tri = np.ascontiguousarray(tri, np.float32).reshape(-1)
    p = np.ascontiguousarray(p, np.float32)
    v = np.ascontiguousarray(v, np.float32)
    t = C.c_float()
    tid = lib().orc_find_triangle(fptr(tri), len(tri) // 9, fptr(p), fptr(v), ref_tri, C.byref(t))
    return tid, t.value


def fresnel(n1, n2, c):
    return lib().orc_fresnel(n1, n2, c)


def transmit_or_reflect(nv, v, n1, n2, x):
    """orc_transmit_or_reflect: (kind, new direction)."""
    a = np.ascontiguousarray(nv, np.float32).copy()
    b = np.ascontiguousarray(v, np.float32).copy()
    k = lib().orc_transmit_or_reflect(fptr(a), fptr(b), n1, n2, x)
    return k, b


def vscatter(v, g, xt, xp):
    a = np.ascontiguousarray(v, np.float32).copy()
    lib().orc_vscatter(fptr(a), g, xt, xp)
    return a


# --- whole runs ------------------------------------------------------------
class Run:
    """Keeps the numpy buffers of one oracle call alive."""

    def __init__(self, scene, rng_mode=RNG_PHILOX, seed=1, max_steps=0, tallies=None, max_boundary_run=100_000):
        s = scene
        self.s = s
        self.tri = np.ascontiguousarray(s["tri"], np.float32).reshape(-1)
        self.nrm = np.ascontiguousarray(s["nrm"], np.float32).reshape(-1)
        self.med = np.ascontiguousarray(s["med"], np.int32).reshape(-1)
        self.nm = np.ascontiguousarray(s["nmuasg"], np.float32).reshape(-1)
        self.vstart = np.ascontiguousarray(s["vstart"], np.float32).reshape(-1)
        self.vocc = np.ascontiguousarray(s["vocc"], np.int32).reshape(-1)
        tl = dict(tallies or {})
        self.det_lo = np.ascontiguousarray(tl.get("det_rlo", np.zeros(1)), np.float32)
        self.det_hi = np.ascontiguousarray(tl.get("det_rhi", np.zeros(1)), np.float32)
        P = Params()
        P.ntri = len(self.tri) // 9
        P.tri, P.nrm, P.med = fptr(self.tri), fptr(self.nrm), iptr(self.med)
        P.ntissue = len(self.nm) // 4
        P.nmuasg = fptr(self.nm)
        P.fetal_id = int(s.get("fetal_id", -1))
        P.pstart[:] = [float(x) for x in s["pstart"]]
        P.vstart, P.vocc, P.nstart = fptr(self.vstart), iptr(self.vocc), len(self.vocc)
        P.poi_basis[:] = [float(x) for x in s["poi_basis"]]
        P.model_dims[:] = [float(x) for x in s["model_dims"]]
        P.det_grid[:] = [float(x) for x in s["det_grid"]]
        P.factor, P.chance, P.w_th = s.get("factor", 1.0), s["chance"], s["w_th"]
        P.rng_mode, P.seed, P.max_steps = rng_mode, seed, max_steps
        P.max_boundary_run = max_boundary_run  # the library's default
        vox = tl.get("vox")
        if vox:
            P.vox_origin[:] = list(vox["origin"])
            P.vox_size[:] = list(vox["size"])
            P.vox_n[:] = list(vox["n"])
        rad = tl.get("rad")
        if rad:
            P.rad_dr, P.rad_n = rad["dr"], rad["n"]
        P.ndet = len(tl.get("det_rlo", []))
        P.det_rlo, P.det_rhi = fptr(self.det_lo), fptr(self.det_hi)
        self.P = P
        L = Layout()
        lib().orc_layout_compute(C.byref(P), C.byref(L))
        self.L = L
        self.t = {
            "sft": np.zeros(L.n_sft, np.int64), "rtd": np.zeros(max(L.n_rtd, 1), np.int64),
            "arz": np.zeros(L.n_arz, np.int64), "phirz": np.zeros(L.n_arz, np.int64),
            "region": np.zeros(L.n_region, np.int64), "vox": np.zeros(max(L.n_vox, 1), np.int64),
            "rad": np.zeros(max(L.n_rad, 1), np.int64), "det": np.zeros(max(L.n_det, 1), np.int64),
        }
        T = Tallies()
        for k, a in self.t.items():
            setattr(T, k, lptr(a))
        self.T = T

    def run(self, first, count, exit_rec=False, trace_max=0, nthreads=1):
        ex = np.zeros((count, 8), np.float32) if exit_rec else None
        tr = np.zeros((count, max(trace_max, 1), 12), np.float32) if trace_max else None
        tn = np.zeros(count, np.int32) if trace_max else None
        rc = lib().orc_run(C.byref(self.P), first, count, C.byref(self.T),
                           fptr(ex) if ex is not None else None,
                           fptr(tr) if tr is not None else None, trace_max,
                           iptr(tn) if tn is not None else None, nthreads)
        assert rc == 0
        return ex, tr, tn

    def adjoint(self, first, count, jac=None, nthreads=1, tallies=False, jac_mus=False, paths=False):
        """orc_adjoint: (signals (ndet,), grad (ndet, ntissue, 2), jac
        (ndet, nvox) or None, pairs), 2^32 fixed point; jac = (origin,
        size, n) of the voxel grid or None.  With jac_mus, the per-voxel mus
        Jacobian (ndet, nvox) is appended as a fifth element; with paths,
        the per-photon path records dict(length (count, ntissue),
        interactions (count, ntissue), w (count,), radius (count,)) last."""
        nd = self.P.ndet
        nt = self.P.ntissue
        A = AdjointOut()
        sig = np.zeros(max(nd, 1), np.int64)
        grad = np.zeros(max(nd, 1) * nt * 2, np.int64)
        J = None
        if jac is not None:
            o, sz, n = jac
            A.jac_origin[:] = list(o)
            A.jac_size[:] = list(sz)
            A.jac_n[:] = list(n)
            J = np.zeros(max(nd, 1) * int(np.prod(n)), np.int64)
        A.signals, A.grad = lptr(sig), lptr(grad)
        A.jac = lptr(J) if J is not None else None
        Js = np.zeros_like(J) if (jac_mus and J is not None) else None
        A.jac_mus = lptr(Js) if Js is not None else None
        if paths:
            P = dict(length=np.zeros((count, nt), np.float32), interactions=np.zeros((count, nt), np.int32),
                     w=np.zeros(count, np.float32), radius=np.zeros(count, np.float32))
            A.path_len = P["length"].ctypes.data_as(C.POINTER(C.c_float))
            A.path_int = P["interactions"].ctypes.data_as(C.POINTER(C.c_int32))
            A.path_w = P["w"].ctypes.data_as(C.POINTER(C.c_float))
            A.path_r = P["radius"].ctypes.data_as(C.POINTER(C.c_float))
        rc = lib().orc_adjoint(C.byref(self.P), first, count, C.byref(self.T) if tallies else None, C.byref(A),
                               nthreads)
        assert rc == 0
        out = (sig[:nd], grad[:nd * nt * 2].reshape(nd, nt, 2),
               J.reshape(max(nd, 1), -1)[:nd] if J is not None else None, A.ndetected)
        if jac_mus:
            out += (Js.reshape(max(nd, 1), -1)[:nd] if Js is not None else None,)
        if paths:
            out += (P,)
        return out

    def tallies(self):
        out = {k: v.copy() for k, v in self.t.items()}
        out["misc"] = dict(zip(MISC, list(self.T.misc)))
        return out

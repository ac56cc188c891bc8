"""lp_run_batch_noalloc: the batched launch that allocates no device memory
(its instance table and work counter live in a static device array), called
through the C ABI with device output buffers the test allocates itself.

* Two calls in a row, the second with fewer instances and photons, each equal
  to lp_run_batch_geom's det and region tallies bit for bit -- over three
  uneven geometries in one launch, so the geometry rows of the static table
  are read too;
* sizes beyond the static table are refused, and so is a context whose
  scene is not on the device yet.
"""
import ctypes as C

import numpy as np
import pytest

import lightpropagation_amd as lp
from lightpropagation_amd import scenes as S
from lightpropagation_amd._capi import lib

pytestmark = pytest.mark.gpu

GEOMS = [(1.5, 8, {"t_fat": 0.3, "t_uterus": 0.5}), (2.5, 10, {}), (3.5, 6, {})]


def _hip():
    """The HIP runtime liblp.so links: the very file already mapped into this
    process with it (a second copy of the runtime would be a second runtime)."""
    lib()
    with open("/proc/self/maps") as f:
        paths = {ln.split()[-1] for ln in f if "libamdhip64.so" in ln}
    assert len(paths) == 1, paths
    h = C.CDLL(paths.pop(), mode=C.RTLD_GLOBAL)
    h.hipMalloc.argtypes = [C.POINTER(C.c_void_p), C.c_size_t]
    h.hipFree.argtypes = [C.c_void_p]
    h.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
    h.hipMemset.argtypes = [C.c_void_p, C.c_int, C.c_size_t]
    h.hipDeviceSynchronize.argtypes = []
    return h


class DevArray:
    """An int64 device buffer (hipMalloc by the test, not the library)."""

    def __init__(self, hip, n):
        self.hip, self.n, self.p = hip, n, C.c_void_p()
        assert hip.hipMalloc(C.byref(self.p), max(n, 1) * 8) == 0
        assert hip.hipMemset(self.p, 0x5A, max(n, 1) * 8) == 0  # garbage the call must clear
        assert hip.hipDeviceSynchronize() == 0

    def get(self):
        out = np.zeros(self.n, np.int64)
        assert self.hip.hipDeviceSynchronize() == 0
        assert self.hip.hipMemcpy(out.ctypes.data_as(C.c_void_p), self.p, self.n * 8, 2) == 0  # D2H
        return out

    def free(self):
        self.hip.hipFree(self.p)


def _table(pairs):
    return np.ascontiguousarray(np.stack([S.mf_optics(wl, sm, sf) for wl, sm, sf in pairs]), np.float32)


def _noalloc(ctx, tab, geom, first, count, det, reg):
    g = np.ascontiguousarray(geom, np.int32)
    return lib().lp_run_batch_noalloc(ctx._c, tab.ctypes.data_as(C.POINTER(C.c_float)), len(tab),
                                      g.ctypes.data_as(C.POINTER(C.c_int32)), first, count, det.p if det else None,
                                      reg.p)


def test_noalloc_equals_batch_on_two_calls():
    hip = _hip()
    scs = [S.maternal_fetal(fetal_depth=d, freq_maternal=f, freq_fetal=f, freq_torso=4, **kw) for d, f, kw in GEOMS]
    merged, counts = S.geometries(scs)
    ctx = lp.Context(0)
    ctx.set_scene(merged, seed=0x5EED000A)
    ctx.set_geometries(counts)
    c = 2.0 + 8.0 * np.arange(8) / 7.0
    ctx.set_detectors(c - 0.25, c + 0.25)
    ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["region"] | 1 << lp.TALLY["misc"])
    nt, nd = len(merged["nmuasg"]), 8
    pairs = [(660.0, 0.97, 0.5), (940.0, 0.97, 0.3), (660.0, 0.8, 0.75), (940.0, 0.6, 0.45), (660.0, 0.9, 0.35),
             (940.0, 0.97, 0.6)]
    calls = [(_table(pairs), [0, 1, 2, 2, 0, 1], 1000, 90_000),  # then a smaller call
             (_table(pairs[1:4]), [2, 0, 1], 5000, 40_000)]
    det = DevArray(hip, len(pairs) * nd * 2)
    reg = DevArray(hip, len(pairs) * nt)
    for k, (tab, geom, first, count) in enumerate(calls):
        want = ctx.run_batch(tab, first, count, geom=geom)  # also puts the scene on the device
        print("call", k, "instances", len(tab), "photons", count, flush=True)
        rc = _noalloc(ctx, tab, geom, first, count, det, reg)
        assert rc == 0, lib().lp_last_error(ctx._c)
        assert ctx.stats().photons == len(tab) * count
        got_det = det.get()[:len(tab) * nd * 2].reshape(len(tab), nd, 2)
        got_reg = reg.get()[:len(tab) * nt].reshape(len(tab), nt)
        assert np.array_equal(got_det, want["det"]), k
        assert np.array_equal(got_reg, want["region"]), k
        assert want["det"][:, :, 0].sum() > 0
    det.free()
    reg.free()
    ctx.close()


def test_noalloc_refusals():
    hip = _hip()
    sc = S.maternal_fetal(freq_maternal=6, freq_fetal=6, freq_torso=4)
    ctx = lp.Context(0)
    ctx.set_scene(sc, seed=3)
    ctx.set_detectors([2.0], [3.0])
    det, reg = DevArray(hip, 200 * 2), DevArray(hip, 200 * 8)
    tab = _table([(660.0, 0.97, 0.5)])
    rc = _noalloc(ctx, tab, [0], 0, 100, det, reg)  # the scene is not on the device yet
    assert rc == 6 and b"not on the device" in lib().lp_last_error(ctx._c)
    ctx.run(0, 1000)
    big = np.ascontiguousarray(np.repeat(tab, 200, axis=0))  # 200 x 12 rows > 2048
    rc = _noalloc(ctx, big, np.zeros(200, np.int32), 0, 10, det, reg)
    assert rc == 3 and b"static table" in lib().lp_last_error(ctx._c)
    assert _noalloc(ctx, tab, [0], 0, 100, det, reg) == 0
    det.free()
    reg.free()
    ctx.close()

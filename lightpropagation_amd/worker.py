"""A process that holds one liblp.so context and serves requests over two
pipes: the device side of the torch binding (lightpropagation_amd.torch_adjoint).

torch ships its own HIP runtime (torch/lib/libamdhip64.so) under the same
soname as the ROCm runtime liblp.so links, so a process that loads both ends
up with two runtimes on one device.  The torch side therefore never loads
liblp.so: it starts this module as a child process (`python -m
lightpropagation_amd.worker <read fd> <write fd>`) and sends it requests.
This module refuses to run once torch is imported.

Wire format: each message is an 8-byte little-endian length followed by a
pickle of a tuple.  Requests are (command, kwargs); replies are ("ok", value)
or ("err", message).  Both ends are this package's own code.
"""
import os
import pickle
import struct
import sys
import time

import numpy as np


def send(f, obj):
    b = pickle.dumps(obj, protocol=pickle.HIGHEST_PROTOCOL)
    f.write(struct.pack("<Q", len(b)))
    f.write(b)
    f.flush()


def recv(f):
    h = f.read(8)
    if len(h) < 8:
        raise EOFError("pipe closed")
    (n,) = struct.unpack("<Q", h)
    b = f.read(n)
    if len(b) < n:
        raise EOFError("pipe closed")
    return pickle.loads(b)


class Server:
    """Requests: scene, optics, materials, adjoint, run, stats, ping, close."""

    def __init__(self):
        self.ctx = None
        self.sc = None

    def scene(self, spec, seed, det_lo, det_hi, jac_grid=None, device=0, options=None):
        """spec: a scene dict (see lightpropagation_amd.scenes) or (builder
        name, kwargs) for a builder of lightpropagation_amd.scenes."""
        import lightpropagation_amd as lp
        from lightpropagation_amd import scenes
        if isinstance(spec, (tuple, list)):
            name, kw = spec
            if name not in ("maternal_fetal", "slab"):
                raise ValueError("unknown scene builder %r" % (name,))
            sc = getattr(scenes, name)(**kw)
        else:
            sc = spec
        if self.ctx is not None:
            self.ctx.close()
        self.ctx = lp.Context(device)
        self.ctx.set_scene(sc, seed=seed)
        self.ctx.set_detectors(det_lo, det_hi)
        if jac_grid is not None:
            self.ctx.set_jacobian_grid(*jac_grid)
        self.ctx.set_tallies(1 << lp.TALLY["det"] | 1 << lp.TALLY["misc"])
        for k, v in (options or {}).items():
            self.ctx.set_option(k, float(v))
        self.sc = sc
        return dict(nmuasg=np.asarray(sc["nmuasg"], np.float32), fetal_id=int(sc["fetal_id"]),
                    ntri=len(sc["tri"]), ndet=len(det_lo))

    def optics(self, wavelength_nm, spo2_maternal, spo2_fetal):
        from lightpropagation_amd import scenes
        return scenes.mf_optics(wavelength_nm, spo2_maternal, spo2_fetal,
                                regions=10 if len(self.sc["nmuasg"]) == 11 else 7)

    def materials(self, nmuasg):
        self.ctx.set_materials(np.asarray(nmuasg, np.float32), self.sc["fetal_id"])
        return None

    def adjoint(self, first, count, jac=False, nmuasg=None, paths=False):
        """Per-photon detector signals and gradients (float64, divided by the
        photon count) of photons [first, first + count); paths: also the
        per-pair path records (lp_get_adjoint_paths) as they come."""
        import lightpropagation_amd as lp
        if nmuasg is not None:
            self.materials(nmuasg)
        self.ctx.reset()
        t0 = time.perf_counter()
        r = self.ctx.adjoint(first, count, jac=jac, paths=paths)
        wall = time.perf_counter() - t0
        st = self.ctx.stats()
        if st.replay_mismatch:
            raise RuntimeError("adjoint replay ended %d paths with another weight" % st.replay_mismatch)
        s = lp.FIXED * float(count)
        return dict(signals=r["signals"] / s, grad=r["grad"] / s,
                    jac=None if r["jac"] is None else r["jac"] / s, pairs=r["pairs"],
                    forward_ms=st.kernel_ms, replay_ms=st.replay_ms, wall_s=wall, paths=r.get("paths"))

    def run(self, first, count, nmuasg=None):
        """Per-photon detector signals of a forward run (float64)."""
        import lightpropagation_amd as lp
        if nmuasg is not None:
            self.materials(nmuasg)
        self.ctx.reset()
        t0 = time.perf_counter()
        self.ctx.run(first, count)
        wall = time.perf_counter() - t0
        det = self.ctx.tallies()["det"].reshape(-1, 2)[:, 0]
        return dict(signals=det / (lp.FIXED * float(count)), forward_ms=self.ctx.stats().kernel_ms, wall_s=wall)

    def ping(self):
        return dict(pid=os.getpid(), torch_loaded="torch" in sys.modules)

    def close(self):
        if self.ctx is not None:
            self.ctx.close()
            self.ctx = None
        return None


def serve(rfd, wfd):
    if "torch" in sys.modules:
        raise RuntimeError("lightpropagation_amd.worker must run in a process without torch")
    srv = Server()
    with os.fdopen(rfd, "rb") as fin, os.fdopen(wfd, "wb") as fout:
        while True:
            try:
                cmd, kw = recv(fin)
            except EOFError:
                break
            if cmd not in ("scene", "optics", "materials", "adjoint", "run", "ping", "close"):
                send(fout, ("err", "unknown command %r" % (cmd,)))
                continue
            try:
                send(fout, ("ok", getattr(srv, cmd)(**kw)))
            except Exception as e:  # reported to the caller, which raises it
                send(fout, ("err", "%s: %s" % (type(e).__name__, e)))
            if cmd == "close":
                break
    srv.close()


if __name__ == "__main__":
    serve(int(sys.argv[1]), int(sys.argv[2]))

"""torch binding of the adjoint (lp_adjoint): annular detector signals as a
differentiable function of every tissue's (mua, mus), and the fit of
maternal and fetal arterial oxygen saturation to two-wavelength detector
data with Adam.

The transport runs in a child process (lightpropagation_amd.worker) that
holds liblp.so; this process holds torch.  torch ships its own HIP runtime
under the soname of the ROCm runtime liblp.so links, so the two are never
loaded into one process.  Tensors cross as numpy arrays over a pipe; the
work on the device is one forward launch and one replay launch per call.

Estimator (the path-length or perturbation Monte Carlo estimator, evaluated
on every detected path by replaying the photon's counter-based random
stream): for a photon with exit weight W, path length L_t and n_t
interactions in tissue t,
    dW/dmua_t = -L_t W,    dW/dmus_t = (n_t / mus_t - L_t) W.
The signals and gradients are means over the launched photons, so the
backward pass is exact for the sample: grad_mua = g @ dS/dmua.
"""
import os
import pickle
import select
import struct
import subprocess
import sys

import numpy as np
import torch

from . import worker as _w

_PKG_PARENT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


class WorkerError(RuntimeError):
    pass


class Simulator:
    """A transport context in a child process.

    scene: a scene dict or (builder, kwargs) with builder "maternal_fetal" or
    "slab" of lightpropagation_amd.scenes (built in the child, which holds
    the library).  det_lo/det_hi: annular detectors (exit radius in the
    emitter frame).  jac_grid: (origin, size, n) of the voxel mua Jacobian.
    timeout: seconds to wait for any one reply before WorkerError.
    """

    def __init__(self, scene, seed, det_lo, det_hi, jac_grid=None, device=0, options=None, timeout=600):
        r1, w1 = os.pipe()
        r2, w2 = os.pipe()
        env = dict(os.environ)
        env["PYTHONPATH"] = _PKG_PARENT + (os.pathsep + env["PYTHONPATH"] if env.get("PYTHONPATH") else "")
        self.proc = subprocess.Popen([sys.executable, "-u", "-m", "lightpropagation_amd.worker", str(r1), str(w2)],
                                     pass_fds=(r1, w2), env=env)
        os.close(r1)
        os.close(w2)
        self._out = os.fdopen(w1, "wb")
        self._in = r2  # raw descriptor: replies are read with a deadline
        self.timeout = timeout
        info = self.call("scene", spec=scene, seed=int(seed), det_lo=np.asarray(det_lo, np.float64),
                         det_hi=np.asarray(det_hi, np.float64), jac_grid=jac_grid, device=device,
                         options=options)
        self.nmuasg = info["nmuasg"]
        self.fetal_id = info["fetal_id"]
        self.ndet = info["ndet"]
        self.ntissue = len(self.nmuasg)

    def call(self, cmd, **kw):
        if self.proc is None:
            raise WorkerError("simulator is closed")
        try:
            _w.send(self._out, (cmd, kw))
            (n,) = struct.unpack("<Q", self._read(8))
            status, val = pickle.loads(self._read(n))
        except (EOFError, BrokenPipeError) as e:
            rc = self.proc.poll()
            raise WorkerError("worker process ended (exit status %s) during %r" % (rc, cmd)) from e
        if status != "ok":
            raise WorkerError(val)
        return val

    def _read(self, n):
        """n bytes of the reply, or WorkerError once `timeout` passes
        without any."""
        buf = bytearray()
        while len(buf) < n:
            ready, _, _ = select.select([self._in], [], [], self.timeout)
            if not ready:  # a late reply would answer the next request: end the worker
                self.proc.kill()
                self.proc.wait()
                self.proc = None
                raise WorkerError("no reply from the worker process in %g s" % self.timeout)
            chunk = os.read(self._in, n - len(buf))
            if not chunk:
                raise EOFError("pipe closed")
            buf += chunk
        return bytes(buf)

    def optics(self, wavelength_nm, spo2_maternal, spo2_fetal):
        """(n, mua, mus, g) of the maternal-fetal tissues (lpm_mf_optics)."""
        return self.call("optics", wavelength_nm=float(wavelength_nm), spo2_maternal=float(spo2_maternal),
                         spo2_fetal=float(spo2_fetal))

    def adjoint(self, nmuasg, first, count, jac=False, paths=False):
        """Per-photon signals (ndet,), grad (ndet, ntissue, 2) and, with
        jac, the voxel Jacobian (ndet, nz, ny, nx), as float64 numpy; with
        paths, the per-pair path records (see PathRecords) under "paths"."""
        return self.call("adjoint", first=int(first), count=int(count), jac=bool(jac),
                         nmuasg=np.asarray(nmuasg, np.float32), paths=bool(paths))

    def forward(self, nmuasg, first, count):
        """Per-photon detector signals of a forward run (no replay)."""
        return self.call("run", first=int(first), count=int(count), nmuasg=np.asarray(nmuasg, np.float32))

    def close(self):
        if self.proc is None:
            return
        try:
            self.call("close")
        except WorkerError:
            pass
        try:
            self._out.close()
        except OSError:
            pass
        try:
            os.close(self._in)
        except OSError:
            pass
        try:
            self.proc.wait(timeout=60)
        except subprocess.TimeoutExpired:
            self.proc.kill()
            self.proc.wait()
        self.proc = None

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


class DetectorSignals(torch.autograd.Function):
    """S = signals(mua, mus): mean detector weight per launched photon of
    photons [first, first + count), shape (ndet,).  mua, mus: (ntissue,)
    tensors replacing the scene's columns (taken as float32, the transport's
    type).  The same (first, count) in every call gives common random
    numbers: S is then a deterministic function of (mua, mus)."""

    @staticmethod
    def forward(ctx, mua, mus, sim, first, count):
        nm = np.array(sim.nmuasg, np.float32)
        nm[:, 1] = mua.detach().to("cpu", torch.float32).numpy()
        nm[:, 2] = mus.detach().to("cpu", torch.float32).numpy()
        r = sim.adjoint(nm, first, count)
        G = torch.from_numpy(np.ascontiguousarray(r["grad"])).to(device=mua.device, dtype=mua.dtype)
        ctx.save_for_backward(G)
        return torch.from_numpy(r["signals"]).to(device=mua.device, dtype=mua.dtype)

    @staticmethod
    def backward(ctx, g):
        (G,) = ctx.saved_tensors
        ga = g @ G[..., 0] if ctx.needs_input_grad[0] else None
        gs = g @ G[..., 1] if ctx.needs_input_grad[1] else None
        return ga, gs, None, None, None


def detector_signals(mua, mus, sim, first, count):
    return DetectorSignals.apply(mua, mus, sim, first, count)


class SpO2Optics:
    """mua of every tissue as a function of (maternal, fetal) SpO2 at one
    wavelength: mua_t = a_t + b_t * SpO2 of the tissue's side, exact for the
    maternal-fetal optics (haemoglobin absorption is linear in saturation);
    a, b and the fetal mask come from three evaluations of lpm_mf_optics."""

    def __init__(self, sim, wavelength_nm):
        n00 = sim.optics(wavelength_nm, 0.0, 0.0)
        n10 = sim.optics(wavelength_nm, 1.0, 0.0)
        n01 = sim.optics(wavelength_nm, 0.0, 1.0)
        self.wavelength = wavelength_nm
        self.nmuasg = n00
        a = n00[:, 1].astype(np.float64)
        self.fetal = torch.from_numpy(n01[:, 1] != n00[:, 1])
        b = np.where(self.fetal.numpy(), n01[:, 1], n10[:, 1]).astype(np.float64) - a
        self.a = torch.from_numpy(a)
        self.b = torch.from_numpy(b)
        self.mus = torch.from_numpy(n00[:, 2].astype(np.float64))

    def mua(self, spo2):
        """spo2: tensor (2,) = (maternal, fetal)."""
        s = torch.where(self.fetal, spo2[1], spo2[0])
        return self.a.to(spo2) + self.b.to(spo2) * s


def _residuals(x, sim, optics, measured, weights, first, count):
    """sqrt(w_d) (log S_d - log M_d) over wavelengths and the detectors with
    a measured signal and a nonzero weight."""
    out = []
    for i, (o, m) in enumerate(zip(optics, measured)):
        wts = None if weights is None else weights[i]
        m = torch.as_tensor(np.asarray(m, np.float64))
        w = torch.ones_like(m) if wts is None else torch.as_tensor(np.asarray(wts, np.float64))
        keep = (m > 0) & (w > 0)
        S = detector_signals(o.mua(x), o.mus.clone(), sim, first, count)
        if not bool(torch.all(S[keep] > 0)):
            raise WorkerError("a detector with a measured signal received no photon; raise count")
        out.append(torch.sqrt(w[keep]) * (torch.log(S[keep]) - torch.log(m[keep])))
    return torch.cat(out)


def _whitener(r, x, free):
    """|r| V diag(lambda^-1/2) over the free coordinates, with
    J^T J = V diag(lambda) V^T and J = dr/dx[free] from the graph of r (one
    backward pass per residual, no transport).  In the coordinates z with
    dx[free] = |r| V diag(lambda^-1/2) dz the linearised loss is
    |r|^2 |z - z*|^2 with |z*| <= 1 and each coordinate along one principal
    axis of the valley, so Adam's per-coordinate steps of about lr neither
    stall along the weak axis nor throw it far past the optimum."""
    J = torch.stack([torch.autograd.grad(r[i], x, retain_graph=True)[0][free] for i in range(len(r))])
    ev, V = torch.linalg.eigh(J.T @ J)
    ev = torch.clamp(ev, min=1e-12 * float(ev.max()) + 1e-300)
    return float(torch.linalg.vector_norm(r.detach())) * (V @ torch.diag(ev.rsqrt()))


def _whitened_adam(resid, init, steps, lr, lr_end, free, max_step, rewhiten, betas, callback):
    """Adam on the loss |resid(x)|^2 over the free saturations, in the
    whitened coordinates that fit_spo2 describes; returns (x, history)."""
    fidx = torch.tensor([i for i in range(2) if free[i]], dtype=torch.long)
    if len(fidx) == 0:
        raise ValueError("no free saturation")
    k = len(fidx)
    x0 = torch.tensor(init, dtype=torch.float64)
    P = torch.eye(k, dtype=torch.float64)
    gamma = 1.0 if lr_end is None else (lr_end / lr) ** (1.0 / max(steps - 1, 1))
    z = opt = None
    hist = []

    def embed(zz):
        dx = torch.zeros(2, dtype=torch.float64)
        return dx.index_put((fidx,), P @ zz)

    for it in range(steps):
        if z is None:
            z = torch.zeros(k, dtype=torch.float64, requires_grad=True)
            opt = torch.optim.Adam([z], lr=lr, betas=betas)
        xl = (x0 + embed(z.detach())).requires_grad_(True)
        xe = xl + (torch.clamp(xl, 0.0, 1.0) - xl).detach()  # evaluated clamped, differentiated unclamped
        r = resid(xe)
        loss = (r * r).sum()
        if rewhiten is not None and (it == 0 or (rewhiten > 0 and it % rewhiten == 0)):
            P = _whitener(r, xl, fidx).detach()
            x0 = xl.detach().clone()
            z = torch.zeros(k, dtype=torch.float64, requires_grad=True)
            opt = torch.optim.Adam([z], lr=lr, betas=betas)
        (gx,) = torch.autograd.grad(loss, xl)
        for g in opt.param_groups:
            g["lr"] = lr * gamma ** it
        opt.zero_grad()
        z.grad = P.T @ gx[fidx]
        z_prev = z.detach().clone()
        opt.step()
        with torch.no_grad():
            dx = float(torch.max(torch.abs(P @ (z - z_prev))))
            if dx > max_step:
                z.copy_(z_prev + (z - z_prev) * (max_step / dx))
        hist.append((float(loss.detach()), xe.detach().numpy().copy()))
        if callback is not None:
            callback(it, hist[-1])
    with torch.no_grad():
        x = torch.clamp(x0 + embed(z.detach()), 0.0, 1.0)
    return x.numpy().copy(), hist


def fit_spo2(sim, optics, measured, init=(0.8, 0.7), steps=60, lr=0.3, lr_end=0.002, first=0,
             count=1_000_000, weights=None, free=(True, True), max_step=0.2, rewhiten=0, betas=(0.5, 0.9),
             callback=None):
    """Adam fit of (maternal, fetal) SpO2 to measured detector signals, one
    transport (forward + adjoint replay) per wavelength and step.

    optics: list of SpO2Optics (one per wavelength); measured: list of
    (ndet,) arrays, the signals per photon at each wavelength; weights: None
    or one (ndet,) array per wavelength.  The loss is the sum over
    wavelengths and detectors of w_d (log S_d - log M_d)^2 (w_d = 1 by
    default; detectors with M_d = 0 or w_d = 0 are left out).  Every
    evaluation runs photons [first, first + count) (common random numbers),
    so the loss is a deterministic function of SpO2.  free selects the
    fitted saturations (the others stay at init, e.g. a maternal SpO2 known
    from a pulse oximeter).

    The two saturations are seen with very different strength (the maternal
    layers lie on every detected path, the fetal head only on the far
    detectors' paths), so the loss is a long valley in SpO2 (the
    Gauss-Newton matrix has condition numbers from hundreds to 10^5).  Adam
    therefore steps in whitened coordinates z, SpO2 = x0 + |r| V L^-1/2 z,
    where J^T J = V L V^T is the Gauss-Newton matrix and r the residuals at
    x0 (J from the adjoint gradients already computed for the step).  Every
    `rewhiten` steps (0: only at the start; None: never, plain Adam on SpO2)
    x0 and the whitening are taken afresh and Adam restarts there.  A step
    that would change a saturation by more than max_step is shortened to
    max_step.  The learning rate decays exponentially from lr to lr_end
    (constant when lr_end is None); the transport sees SpO2 clamped to
    [0, 1], and the gradient passes the clamp unchanged (a coordinate past a
    bound is pulled back when the optimum lies inside).  Returns (SpO2 numpy
    (2,), history list of (loss, SpO2) at each step's evaluation point).

    Common random numbers keep the photons, not their paths: a change of a
    tissue's mua moves float32 step lengths, and the paths of the photons
    that cross that tissue diverge, so this loss is rough in the weakly seen
    fetal saturation.  fit_spo2_paths fits on a fixed set of replayed paths
    instead.
    """
    return _whitened_adam(lambda x: _residuals(x, sim, optics, measured, weights, first, count), init, steps, lr,
                          lr_end, free, max_step, rewhiten, betas, callback)


class PathRecords:
    """The detected paths of photons [first, first + count) at the optics
    nmuasg, from one lp_adjoint with path records: per (photon, detector)
    pair the R exit weight w, and per tissue the path length L_t and the
    interaction count n_t.  At other optics (mua', mus') the same paths
    carry the weights (perturbation Monte Carlo)
        w' = w prod_t (mus'_t / mus_t)^n_t exp(-(mua'_t + mus'_t - mua_t - mus_t) L_t),
    the ratio of the probability densities of the path under the two sets of
    optics times the ratio of the interaction albedos: an unbiased estimate
    of the signals at (mua', mus') from photons that ran at (mua, mus)
    (roulette and Fresnel decisions do not depend on the optics).  Its
    variance grows with the distance from the optics the paths ran at."""

    def __init__(self, rec, nmuasg, count, ndet):
        self.det = torch.from_numpy(np.asarray(rec["det"], np.int64))
        self.w = torch.from_numpy(np.asarray(rec["w"], np.float64))
        self.L = torch.from_numpy(np.asarray(rec["length"], np.float64))
        self.n = torch.from_numpy(np.asarray(rec["interactions"], np.float64))
        nm = np.asarray(nmuasg, np.float32).astype(np.float64)
        self.mua0 = torch.from_numpy(nm[:, 1].copy())
        self.mus0 = torch.from_numpy(nm[:, 2].copy())
        self.count = int(count)
        self.ndet = int(ndet)
        if bool(torch.any((self.mus0 == 0) & (self.n.sum(0) > 0))):
            raise ValueError("interactions recorded in a tissue with mus 0")

    def weights(self, mua, mus):
        """w' of every pair at (mua, mus) (float64 tensors (ntissue,))."""
        ok = self.mus0 > 0
        ratio = torch.where(ok, mus / torch.where(ok, self.mus0, 1.0), 1.0)
        logr = self.n @ torch.log(ratio) - self.L @ ((mua - self.mua0) + (mus - self.mus0))
        return self.w * torch.exp(logr)

    def signals(self, mua, mus):
        return ReweightedSignals.apply(mua, mus, self)

    def relative_variance(self, mua=None, mus=None):
        """var(S_d) / S_d^2 of the signal means, per detector, from the pairs'
        weights (each photon counts once per detector)."""
        w = self.w if mua is None else self.weights(mua, mus)
        s1 = torch.zeros(self.ndet, dtype=torch.float64).index_add_(0, self.det, w)
        s2 = torch.zeros(self.ndet, dtype=torch.float64).index_add_(0, self.det, w * w)
        N = float(self.count)
        var = (s2 / N - (s1 / N) ** 2) / N
        return torch.where(s1 > 0, var / torch.clamp((s1 / N) ** 2, min=1e-300), torch.inf)

    def pairs_per_detector(self):
        return torch.bincount(self.det, minlength=self.ndet)


class ReweightedSignals(torch.autograd.Function):
    """S(mua, mus) from path records: the mean reweighted detector weight per
    photon, shape (ndet,).  The backward pass is the path-length estimator
    on the reweighted paths, dw'/dmua_t = -L_t w', dw'/dmus_t = (n_t/mus_t -
    L_t) w' (at the records' own optics it equals lp_adjoint's gradient)."""

    @staticmethod
    def forward(ctx, mua, mus, paths):
        w = paths.weights(mua.detach(), mus.detach())
        ctx.paths = paths
        ctx.save_for_backward(w, mus.detach())
        S = torch.zeros(paths.ndet, dtype=torch.float64).index_add_(0, paths.det, w) / paths.count
        return S.to(mua.dtype)

    @staticmethod
    def backward(ctx, g):
        w, mus = ctx.saved_tensors
        p = ctx.paths
        gw = g.to(torch.float64)[p.det] * w / p.count
        gL = gw @ p.L
        ga = -gL if ctx.needs_input_grad[0] else None
        gs = None
        if ctx.needs_input_grad[1]:
            gn = gw @ p.n
            gs = torch.where(mus > 0, gn / torch.where(mus > 0, mus, 1.0), 0.0) - gL
        return ga, gs, None


def fit_spo2_paths(sim, optics, measured, init=(0.8, 0.7), count=50_000_000, first=0, rounds=3,
                   measured_count=None, min_pairs=200, steps=150, lr=0.3, lr_end=0.002, free=(True, True),
                   max_step=0.2, betas=(0.5, 0.9), callback=None):
    """Adam fit of (maternal, fetal) SpO2 on replayed paths (perturbation
    Monte Carlo): each round runs photons [first, first + count) at the
    current estimate with path records (one lp_adjoint per wavelength;
    count: one number, or one per round) and
    fits on those fixed paths, reweighted to every SpO2 the fit visits
    (PathRecords), with the whitened Adam of fit_spo2; the next round runs
    the photons afresh at the round's result, where the reweighting is
    nearly the identity.  The loss is smooth in SpO2 (the paths do not
    change with it), so the weakly seen fetal saturation is resolved down
    to the Monte Carlo noise of the signals.

    Each residual is (log S_d - log M_d) / sigma_d with sigma_d^2 the
    relative variance of the model's signal (from the records) plus that of
    the data (the same per-photon spread over measured_count photons; None:
    data without noise); detectors with fewer than min_pairs detected pairs
    are left out.  The loss at the optimum is then a chi-square with (used
    detectors - free saturations) degrees of freedom.  Returns (SpO2 numpy
    (2,), info) with info: rounds (per round: reference SpO2, result, loss
    first and last, detectors used, seconds of transport), cov (the
    Gauss-Newton covariance of the free saturations at the result, 2 x 2
    with zeros for fixed ones), chi2, dof."""
    import time
    counts = [int(count)] * rounds if np.ndim(count) == 0 else [int(c) for c in count]
    if len(counts) != rounds or rounds < 1:
        raise ValueError("one photon count per round")
    x = np.clip(np.asarray(init, np.float64), 0.0, 1.0)
    info = dict(rounds=[])
    for rnd, count in enumerate(counts):
        t0 = time.perf_counter()
        recs = []
        for o in optics:
            nm = np.array(o.nmuasg, np.float32)
            nm[:, 1] = o.mua(torch.tensor(x)).numpy()
            r = sim.adjoint(nm, first, count, paths=True)
            recs.append(PathRecords(r["paths"], nm, count, sim.ndet))
            del r
        t_sim = time.perf_counter() - t0
        sig = []
        keep = []
        for rec, m in zip(recs, measured):
            rv = rec.relative_variance()
            if measured_count:
                rv = rv * (1.0 + float(count) / float(measured_count))
            k = (rec.pairs_per_detector() >= min_pairs) & (torch.as_tensor(np.asarray(m)) > 0)
            sig.append(torch.sqrt(rv))
            keep.append(k)

        def resid(xx, recs=recs, sig=sig, keep=keep):
            out = []
            for o, rec, m, s, k in zip(optics, recs, measured, sig, keep):
                S = rec.signals(o.mua(xx), o.mus.clone())
                mm = torch.as_tensor(np.asarray(m, np.float64))
                if not bool(torch.all(S[k] > 0)):
                    raise WorkerError("a detector lost its weight under reweighting; rounds from nearer SpO2")
                out.append((torch.log(S[k]) - torch.log(mm[k])) / s[k])
            return torch.cat(out)

        xr, hist = _whitened_adam(resid, tuple(x), steps, lr, lr_end, free, max_step, 0, betas, callback)
        info["rounds"].append(dict(reference=x.tolist(), result=xr.tolist(), loss_first=hist[0][0],
                                   loss_last=hist[-1][0], detectors=int(sum(int(k.sum()) for k in keep)),
                                   photons=count, pairs=[int(rec.det.numel()) for rec in recs], transport_s=t_sim))
        x = xr
    # Gauss-Newton covariance and chi-square at the result (last round's paths)
    xt = torch.tensor(x, dtype=torch.float64, requires_grad=True)
    r = resid(xt)
    fidx = [i for i in range(2) if free[i]]
    J = torch.stack([torch.autograd.grad(r[i], xt, retain_graph=True)[0] for i in range(len(r))])[:, fidx]
    cov = np.zeros((2, 2))
    cov[np.ix_(fidx, fidx)] = torch.linalg.inv(J.T @ J).detach().numpy()
    info.update(cov=cov, chi2=float((r.detach() ** 2).sum()), dof=int(len(r) - len(fidx)))
    return x, info

"""The torch side of the adjoint binding (lightpropagation_amd.torch_adjoint)
on the CPU: the worker's wire protocol, the autograd Function's backward
against a linear stand-in for the simulator, the SpO2 optics mapping against
lpm_mf_optics, and the Adam fit recovering the saturations of a
Beer-Lambert stand-in exactly.  The GPU path is tests/test_gpu_torch_adjoint.py.
"""
import os
import subprocess
import sys

import numpy as np
import pytest
import torch

from lightpropagation_amd import scenes
from lightpropagation_amd import torch_adjoint as TA
from lightpropagation_amd import worker as W

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_worker_protocol():
    r1, w1 = os.pipe()
    r2, w2 = os.pipe()
    env = dict(os.environ, PYTHONPATH=ROOT)
    p = subprocess.Popen([sys.executable, "-m", "lightpropagation_amd.worker", str(r1), str(w2)],
                         pass_fds=(r1, w2), env=env)
    os.close(r1)
    os.close(w2)
    with os.fdopen(w1, "wb") as out, os.fdopen(r2, "rb") as inp:
        W.send(out, ("ping", {}))
        st, v = W.recv(inp)
        assert st == "ok" and v["pid"] == p.pid and v["torch_loaded"] is False
        W.send(out, ("no_such_command", {}))
        st, v = W.recv(inp)
        assert st == "err" and "unknown command" in v
        W.send(out, ("materials", {"nmuasg": np.zeros((2, 4), np.float32)}))  # before any scene
        st, v = W.recv(inp)
        assert st == "err"
        W.send(out, ("close", {}))
        assert W.recv(inp) == ("ok", None)
    assert p.wait(timeout=60) == 0


def test_worker_reply_deadline():
    # a worker that never answers: the read gives up after the timeout and
    # ends the worker, so no late reply can answer a later request
    r, w = os.pipe()
    p = subprocess.Popen([sys.executable, "-c", "import time; time.sleep(60)"])
    sim = object.__new__(TA.Simulator)
    sim._in, sim._out, sim.timeout, sim.proc = r, os.fdopen(w, "wb"), 0.3, p
    with pytest.raises(TA.WorkerError, match="no reply"):
        sim._read(8)
    assert p.poll() is not None and sim.proc is None
    with pytest.raises(TA.WorkerError, match="closed"):
        sim.call("ping")
    sim._out.close()
    os.close(r)


class LinearSim:
    """S(mua, mus) = S0 + Ga (mua - mua0) + Gs (mus - mus0) on the float32
    values it receives, as the transport receives them."""

    def __init__(self, nt=5, nd=3, seed=0):
        rng = np.random.default_rng(seed)
        self.nmuasg = np.zeros((nt, 4), np.float32)
        self.nmuasg[:, 0] = 1.4
        self.nmuasg[:, 1] = rng.uniform(0.05, 0.5, nt)
        self.nmuasg[:, 2] = rng.uniform(10, 100, nt)
        self.Ga = rng.normal(size=(nd, nt))
        self.Gs = rng.normal(size=(nd, nt)) * 0.01
        self.S0 = rng.uniform(1, 2, nd)
        self.calls = []

    def adjoint(self, nm, first, count, jac=False):
        self.calls.append((first, count))
        d_a = nm[:, 1].astype(np.float64) - self.nmuasg[:, 1]
        d_s = nm[:, 2].astype(np.float64) - self.nmuasg[:, 2]
        S = self.S0 + self.Ga @ d_a + self.Gs @ d_s
        return dict(signals=S, grad=np.stack([self.Ga, self.Gs], -1), jac=None, pairs=1)


def test_autograd_backward_is_the_adjoint_jacobian():
    sim = LinearSim()
    mua = torch.tensor(sim.nmuasg[:, 1], dtype=torch.float64, requires_grad=True)
    mus = torch.tensor(sim.nmuasg[:, 2], dtype=torch.float64, requires_grad=True)
    w = torch.tensor([0.3, -1.2, 2.0], dtype=torch.float64)
    S = TA.detector_signals(mua, mus, sim, 7, 100)
    assert S.dtype == torch.float64 and S.shape == (3,)
    (w * S).sum().backward()
    assert np.allclose(mua.grad.numpy(), w.numpy() @ sim.Ga, rtol=1e-12, atol=0)
    assert np.allclose(mus.grad.numpy(), w.numpy() @ sim.Gs, rtol=1e-12, atol=0)
    assert sim.calls == [(7, 100)]
    # a nonlinear loss on top: chain rule through torch, against central differences
    # of the stand-in (linear, so exact up to the float32 rounding of the inputs)
    mua.grad = None
    loss = (torch.log(TA.detector_signals(mua, mus, sim, 0, 1)) ** 2).sum()
    loss.backward()
    h = 1e-2
    for t in range(len(mua)):
        e = torch.zeros_like(mua)
        e[t] = h
        with torch.no_grad():
            lp_ = (torch.log(TA.detector_signals(mua + e, mus, sim, 0, 1)) ** 2).sum()
            lm_ = (torch.log(TA.detector_signals(mua - e, mus, sim, 0, 1)) ** 2).sum()
        m0 = float(mua[t].detach())
        dh = float(np.float32(m0 + h)) - float(np.float32(m0 - h))
        fd = float(lp_ - lm_) / dh
        assert abs(float(mua.grad[t]) - fd) <= 2e-3 * abs(fd) + 1e-4, t


def test_inputs_without_grad_get_none():
    sim = LinearSim()
    mua = torch.tensor(sim.nmuasg[:, 1], dtype=torch.float64, requires_grad=True)
    mus = torch.tensor(sim.nmuasg[:, 2], dtype=torch.float64)
    TA.detector_signals(mua, mus, sim, 0, 1).sum().backward()
    assert mus.grad is None and np.allclose(mua.grad.numpy(), sim.Ga.sum(0))


class OpticsSim:
    def optics(self, wl, sm, sf):
        return scenes.mf_optics(wl, sm, sf)


@pytest.mark.parametrize("wl", [660.0, 940.0])
def test_spo2_optics_is_lpm_mf_optics(wl):
    o = TA.SpO2Optics(OpticsSim(), wl)
    # fetal tissues with blood: scalp, skull, grey and white matter, body (not the CSF)
    assert o.fetal.numpy().tolist() == [False] * 5 + [True, True, False, True, True, True]
    rng = np.random.default_rng(1)
    for sm, sf in rng.uniform(0, 1, (20, 2)):
        want = scenes.mf_optics(wl, sm, sf)
        got = o.mua(torch.tensor([sm, sf], dtype=torch.float64)).numpy()
        assert np.allclose(got, want[:, 1], rtol=3e-7, atol=1e-9), (sm, sf)
        assert np.array_equal(o.mus.numpy(), want[:, 2].astype(np.float64))
    # d mua / d SpO2 has the sign of (HbO2 - Hb) extinction: negative at 660 nm, positive at 940 nm
    b = o.b.numpy()[1:]
    blood = [0, 1, 2, 4, 5, 7, 8, 9]  # tissues 1-3, 5, 6, 8-10 (fluid and CSF hold no blood)
    assert np.all(b[blood] < 0) if wl == 660.0 else np.all(b[blood] > 0)
    assert np.all(b[[3, 6]] == 0)


class BeerLambertSim(OpticsSim):
    """S_d = exp(-sum_t C_dt mua_t): detector d sees tissue t over an
    effective path C_dt that grows with the detector distance, deeper
    (fetal) tissues only at the far detectors."""

    def __init__(self):
        d = np.linspace(2, 10, 8)
        self.nmuasg = scenes.mf_optics(660.0, 0.97, 0.5)
        C = np.zeros((8, len(self.nmuasg)))
        C[:, 1] = 0.8 + 0.1 * d
        C[:, 2] = 1.0 + 0.6 * d
        C[:, 3] = 0.3 * d
        C[:, 4] = 0.05 * d
        C[:, 5:11] = np.maximum(d - 3, 0)[:, None] * np.array([0.05, 0.1, 0.02, 0.2, 0.1, 0.05])
        self.C = C

    def adjoint(self, nm, first, count, jac=False):
        S = np.exp(-self.C @ nm[:, 1].astype(np.float64))
        G = np.zeros((8, len(nm), 2))
        G[..., 0] = -self.C * S[:, None]
        return dict(signals=S, grad=G, jac=None, pairs=1)


@pytest.mark.parametrize("init", [(0.85, 0.7), (0.6, 0.2), (0.99, 0.9), (0.5, 0.5), (0.95, 0.05)])
def test_fit_recovers_the_saturations(init):
    # log S is linear in SpO2 here, with a Gauss-Newton condition number of
    # about 230 (fetal tissue only on the far detectors' paths)
    sim = BeerLambertSim()
    opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
    meas = [sim.adjoint(sim.optics(o.wavelength, 0.97, 0.50), 0, 1)["signals"] for o in opts]
    x, hist = TA.fit_spo2(sim, opts, meas, init=init)
    assert hist[-1][0] < 1e-6 * hist[0][0]
    assert np.allclose(x, [0.97, 0.50], atol=2e-3), x


@pytest.mark.parametrize("init", [0.0, 0.2, 0.9])
def test_fit_of_the_fetal_saturation_alone(init):
    # maternal SpO2 known (a pulse oximeter): it stays at its value
    sim = BeerLambertSim()
    opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
    meas = [sim.adjoint(sim.optics(o.wavelength, 0.97, 0.50), 0, 1)["signals"] for o in opts]
    x, hist = TA.fit_spo2(sim, opts, meas, init=(0.97, init), steps=40, free=(False, True))
    assert x[0] == 0.97 and abs(x[1] - 0.50) < 2e-3, x
    assert all(h[1][0] == 0.97 for h in hist)


def test_fit_steps_are_capped():
    sim = BeerLambertSim()
    opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
    meas = [sim.adjoint(sim.optics(o.wavelength, 0.97, 0.50), 0, 1)["signals"] for o in opts]
    _, hist = TA.fit_spo2(sim, opts, meas, init=(0.6, 0.2), steps=10, max_step=0.05)
    xs = np.array([h[1] for h in hist])
    assert np.abs(np.diff(xs, axis=0)).max() <= 0.05 + 1e-12


def test_plain_adam_stalls_in_the_valley():
    # the reason for the whitening: 50 plain Adam steps on SpO2 end far off
    sim = BeerLambertSim()
    opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
    meas = [sim.adjoint(sim.optics(o.wavelength, 0.97, 0.50), 0, 1)["signals"] for o in opts]
    x, _ = TA.fit_spo2(sim, opts, meas, init=(0.85, 0.7), steps=50, lr=0.05, lr_end=0.001, rewhiten=None)
    assert np.abs(x - [0.97, 0.50]).max() > 0.05


def _records(seed=3, P=5000, nt=5, nd=3):
    rng = np.random.default_rng(seed)
    nm = np.zeros((nt, 4), np.float32)
    nm[:, 0] = 1.4
    nm[:, 1] = rng.uniform(0.05, 0.5, nt)
    nm[:, 2] = rng.uniform(10, 100, nt)
    nm[0, 1:3] = 0.0  # an ambient start tissue: no interactions, mus 0
    L = rng.exponential(0.5, (P, nt)).astype(np.float32)
    L[:, 0] = 0.002
    n = rng.poisson(L * nm[:, 2]).astype(np.uint32)
    rec = dict(det=rng.integers(0, nd, P).astype(np.int32), w=rng.uniform(0.01, 0.5, P).astype(np.float32),
               length=L, interactions=n)
    return TA.PathRecords(rec, nm, 20 * P, nd), nm


def test_reweighted_signals_and_their_gradient():
    paths, nm = _records()
    mua0 = torch.tensor(nm[:, 1], dtype=torch.float64)
    mus0 = torch.tensor(nm[:, 2], dtype=torch.float64)
    # at the records' own optics: the mean weight per photon, and the
    # path-length estimator's gradient
    mua, mus = mua0.clone().requires_grad_(True), mus0.clone().requires_grad_(True)
    S = paths.signals(mua, mus)
    want = np.bincount(paths.det.numpy(), paths.w.numpy(), 3) / paths.count
    assert np.allclose(S.detach().numpy(), want, rtol=1e-12)
    S[1].backward()
    sel = (paths.det == 1).numpy()
    w, L, n = paths.w.numpy()[sel], paths.L.numpy()[sel], paths.n.numpy()[sel]
    assert np.allclose(mua.grad.numpy(), -(w @ L) / paths.count, rtol=1e-12)
    gs = ((w @ n) / np.where(nm[:, 2] > 0, nm[:, 2], 1) - w @ L) / paths.count
    assert np.allclose(mus.grad.numpy()[1:], gs[1:], rtol=1e-10)
    # elsewhere: the backward pass is the derivative of the forward one
    f = lambda a, s: paths.signals(a, s)  # noqa: E731
    a1 = (mua0 * 1.3).requires_grad_(True)
    s1 = (mus0 * 0.9).requires_grad_(True)
    assert torch.autograd.gradcheck(f, (a1, s1), eps=1e-6, atol=1e-9, rtol=1e-6)
    # reweighting to other optics multiplies each weight by the density ratio
    w1 = paths.weights(a1.detach(), s1.detach()).numpy()
    ratio = np.where(nm[:, 2] > 0, 0.9, 1.0)
    r = np.prod(ratio ** paths.n.numpy(), 1) * np.exp(-paths.L.numpy() @ (a1.detach().numpy() - nm[:, 1] +
                                                                         s1.detach().numpy() - nm[:, 2]))
    assert np.allclose(w1, paths.w.numpy() * r, rtol=1e-10)


def test_relative_variance_of_the_means():
    paths, _ = _records(P=20000)
    rv = paths.relative_variance().numpy()
    w, d, N = paths.w.numpy(), paths.det.numpy(), paths.count
    for k in range(3):
        x = np.zeros(N)
        x[:int((d == k).sum())] = w[d == k]  # the photons of detector k, the rest 0
        assert np.isclose(rv[k], x.var() / N / x.mean() ** 2, rtol=1e-9)


class PathSim(OpticsSim):
    """Path records of a Beer-Lambert stand-in: detector d's pairs all have
    path lengths C_d. (so the reweighted signals are exactly exp(-C mua))
    and weights that scatter about that value; the measured signals are the
    exact ones."""

    def __init__(self, pairs=4000):
        self.bl = BeerLambertSim()
        self.pairs = pairs
        self.nmuasg = self.bl.nmuasg
        self.ndet = 8

    def adjoint(self, nm, first, count, jac=False, paths=False):
        C = self.bl.C
        rng = np.random.default_rng(first + 7)
        det = np.repeat(np.arange(8), self.pairs).astype(np.int32)
        S = np.exp(-C @ nm[:, 1].astype(np.float64))
        # weights scattered about the Beer-Lambert value, normalised so the
        # mean per photon is exactly S at the optics the records ran at
        u = rng.uniform(0.5, 1.5, det.size)
        for d in range(8):
            u[det == d] /= u[det == d].mean()
        w = (S[det] * u * count / self.pairs).astype(np.float64)
        rec = dict(det=det, w=w, length=C[det].astype(np.float32), interactions=np.zeros((det.size, len(nm)), np.uint32))
        return dict(signals=S, paths=rec)


@pytest.mark.parametrize("init", [(0.85, 0.7), (0.6, 0.2), (0.99, 0.9)])
def test_fit_on_path_records_recovers_the_saturations(init):
    sim = PathSim()
    opts = [TA.SpO2Optics(sim, wl) for wl in (660.0, 940.0)]
    meas = [sim.bl.adjoint(sim.optics(o.wavelength, 0.97, 0.50), 0, 1)["signals"] for o in opts]
    x, info = TA.fit_spo2_paths(sim, opts, meas, init=init, count=100_000, rounds=2, steps=150, min_pairs=100)
    # float32 path lengths: the reweighted model is Beer-Lambert to 1e-7
    assert np.allclose(x, [0.97, 0.50], atol=1e-3), x
    assert len(info["rounds"]) == 2 and info["rounds"][-1]["loss_last"] < 1e-6
    assert info["dof"] == 16 - 2 and info["cov"].shape == (2, 2) and np.all(np.diag(info["cov"]) > 0)
    # the maternal saturation known: only the fetal one moves
    x, info = TA.fit_spo2_paths(sim, opts, meas, init=(0.97, init[1]), count=100_000, rounds=2, free=(False, True))
    assert x[0] == 0.97 and abs(x[1] - 0.50) < 1e-3 and info["cov"][0, 0] == 0

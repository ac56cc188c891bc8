"""Pins the oracle's boundary directions (TransmitOrReflect,
src/Simulation/MCLightPropagation.cl:921-1018, with Transmit :506-529 and
Reflect :532-549) by what the physics fixes, independently of the formulas
the oracle restates:

* Snell (cos_in > 0, where the reference is physically correct):
  sin(theta_t) = (n1 / n2) sin(theta_i); the transmitted direction is a unit
  vector in the plane of (v, nv), on the far side of the face, with its
  tangential part along v's;
* mirror reflection: v' = v - 2 (v . n) n, for either sign of cos_in;
* total internal reflection: |cos_in| <= cos_crit reflects whatever the draw;
* a whole path at mismatched n: an oblique photon through a weakly
  scattering slab (ambient n 1.0 -> slab 1.5 -> back layer 1.2) leaves at
  x_hit + d tan(theta_t) with exit angle asin(sin(theta_i) / 1.2); photons
  reflected at entry leave at pi - theta_i; the entry reflection rate is
  the closed-form Fresnel reflectance.

Each check is a function that returns its failures, so that the guard tests
can show that the oracle built with a planted mistake (a flipped sin_t, a
flipped reflection tangent, swapped n1/n2; ORC_PLANT_* in mcml_oracle.c)
fails it.  The GPU kernel equals the oracle bit for bit on mismatched-index
scenes (tests/test_gpu_parity.py), so these pins carry over to it.
"""
import math

import numpy as np
import pytest

import _oracle as O
import scenes

F32 = np.float32
PAIRS = [(1.0, 1.4), (1.4, 1.0), (1.37, 1.44), (1.44, 1.37), (1.0, 1.5), (1.33, 1.0)]


def _rand_dirs(rng, n, cos_lo, cos_hi):
    """(nv, v) pairs: random unit normals and directions with cos_in = nv . v
    in [cos_lo, cos_hi]."""
    out = []
    while len(out) < n:
        nv = rng.normal(size=3)
        nv /= np.linalg.norm(nv)
        c = rng.uniform(cos_lo, cos_hi)
        t = rng.normal(size=3)
        t -= t.dot(nv) * nv
        t /= np.linalg.norm(t)
        v = c * nv + math.sqrt(max(0.0, 1 - c * c)) * t
        nv32, v32 = nv.astype(F32), v.astype(F32)
        cos32 = float(np.dot(nv32.astype(np.float64), v32.astype(np.float64)))
        if cos_lo <= cos32 <= cos_hi:
            out.append((nv32, v32))
    return out


def check_snell(n_cases=400, seed=1):
    bad = []
    rng = np.random.default_rng(seed)
    for n1, n2 in PAIRS:
        # below the critical angle when n1 > n2 (else TIR)
        lo = math.sqrt(1 - (n2 / n1) ** 2) + 0.02 if n1 > n2 else 0.05
        for nv, v in _rand_dirs(rng, n_cases // len(PAIRS), lo, 0.999):
            kind, w = O.transmit_or_reflect(nv, v, n1, n2, 1.0)  # x = 1 > r: transmit unless r = 1
            if kind != O.EV_TRANSMIT:
                bad.append(("not transmitted", n1, n2))
                continue
            N, Vi, Vt = nv.astype(float), v.astype(float), w.astype(float)
            if abs(np.linalg.norm(Vt) - 1) > 2e-6:
                bad.append(("|v'| != 1", np.linalg.norm(Vt)))
            if not Vt.dot(N) > 0:
                bad.append(("transmitted back", Vt.dot(N)))
            sin_i = np.linalg.norm(np.cross(Vi, N))
            sin_t = np.linalg.norm(np.cross(Vt, N))
            if abs(sin_t - n1 / n2 * sin_i) > 1e-5:
                bad.append(("Snell", n1, n2, sin_t, n1 / n2 * sin_i))
            if abs(Vt.dot(np.cross(N, Vi))) > 1e-5:
                bad.append(("out of the plane of incidence", Vt.dot(np.cross(N, Vi))))
            ti, tt = Vi - Vi.dot(N) * N, Vt - Vt.dot(N) * N
            if sin_i > 1e-3 and not ti.dot(tt) > 0:
                bad.append(("tangential part reversed", ti.dot(tt)))
    return bad


def check_mirror(n_cases=300, seed=2):
    bad = []
    rng = np.random.default_rng(seed)
    for sign in (1.0, -1.0):
        for nv, v in _rand_dirs(rng, n_cases // 2, 0.02, 0.999):
            v = (sign * v).astype(F32)
            kind, w = O.transmit_or_reflect(nv, v, 1.37, 1.44, 0.0)  # x = 0 <= r: reflect
            if kind != O.EV_REFLECT:
                bad.append(("not reflected", sign))
                continue
            N, Vi = nv.astype(float), v.astype(float)
            want = Vi - 2 * Vi.dot(N) * N
            if np.abs(w.astype(float) - want).max() > 3e-6:
                bad.append(("mirror", sign, w, want))
    return bad


def check_tir(seed=3):
    bad = []
    rng = np.random.default_rng(seed)
    for n1, n2 in [(1.4, 1.0), (1.44, 1.37), (1.33, 1.0)]:
        cc = math.sqrt(1 - (n2 / n1) ** 2)
        for nv, v in _rand_dirs(rng, 60, 0.01, cc - 1e-3):
            kind, w = O.transmit_or_reflect(nv, v, n1, n2, 1.0)
            if kind != O.EV_REFLECT:
                bad.append(("TIR transmitted", n1, n2))
        # just above the critical angle the draw x = 1 transmits
        for nv, v in _rand_dirs(rng, 60, cc + 2e-3, min(cc + 0.05, 0.999)):
            kind, w = O.transmit_or_reflect(nv, v, n1, n2, 1.0)
            if kind != O.EV_TRANSMIT:
                bad.append(("reflected above the critical angle", n1, n2))
    return bad


def test_snell_direction():
    assert check_snell() == []


def test_mirror_reflection():
    assert check_mirror() == []


def test_total_internal_reflection():
    assert check_tir() == []


# --- whole path at mismatched n ------------------------------------------
D = scenes.OBLIQUE_D
N_AMB, N_SLAB, N_BACK = scenes.OBLIQUE_N
THETA_I = scenes.OBLIQUE_THETA
X0, Z0 = scenes.OBLIQUE_X0, scenes.OBLIQUE_Z0


def check_oblique_path(n=20000):
    bad = []
    sc = scenes.oblique_slab()
    r = O.Run(sc, seed=17)
    ex, tr, tn = r.run(0, n, exit_rec=True, trace_max=4, nthreads=8)
    x_hit = X0 + (-Z0) * math.tan(THETA_I)
    theta_t = math.asin(N_AMB / N_SLAB * math.sin(THETA_I))
    theta_o = math.asin(N_SLAB / N_BACK * math.sin(theta_t))
    kinds0 = tr[:, 0, 0]
    # 1. straight through: two transmits, nothing else
    thru = (tn == 2) & (kinds0 == O.EV_TRANSMIT) & (tr[:, 1, 0] == O.EV_TRANSMIT)
    if thru.sum() < n // 2:
        bad.append(("too few straight paths", int(thru.sum())))
    else:
        xe = ex[thru, 0].astype(float)
        if np.abs(xe - (x_hit + D * math.tan(theta_t))).max() > 2e-5:
            bad.append(("exit point", xe[:3], x_hit + D * math.tan(theta_t)))
        ae = ex[thru, 6].astype(float)
        if np.abs(ae - theta_o).max() > 2e-5:
            bad.append(("exit angle", ae[:3], theta_o))
        if np.abs(ex[thru, 2].astype(float) - D).max() > 1e-5:
            bad.append(("exit depth", ex[thru, 2][:3]))
    # 2. reflected at entry: the photon leaves into the ambient at once
    refl = (tn == 1) & (kinds0 == O.EV_REFLECT)
    if refl.any():
        ae = ex[refl, 6].astype(float)
        if np.abs(ae - (math.pi - THETA_I)).max() > 2e-5:
            bad.append(("reflection angle", ae[:3], math.pi - THETA_I))
        if np.abs(ex[refl, 0].astype(float) - x_hit).max() > 1e-5:
            bad.append(("reflection point", ex[refl, 0][:3], x_hit))
    # 3. entry reflectance: unpolarised Fresnel at THETA_I from n 1.0 into 1.5
    ci, ct = math.cos(THETA_I), math.cos(theta_t)
    rs = ((N_AMB * ci - N_SLAB * ct) / (N_AMB * ci + N_SLAB * ct)) ** 2
    rp = ((N_AMB * ct - N_SLAB * ci) / (N_AMB * ct + N_SLAB * ci)) ** 2
    R = 0.5 * (rs + rp)
    frac = refl.mean()
    if abs(frac - R) > 4 * math.sqrt(R * (1 - R) / n):
        bad.append(("entry reflectance", frac, R))
    return bad


def test_oblique_path_through_mismatched_slab():
    assert check_oblique_path() == []


# --- guards: each planted mistake fails a check above ------------------------
@pytest.mark.parametrize("macro,checks", [
    ("ORC_PLANT_FLIP_SIN_T", (check_snell, check_oblique_path)),
    ("ORC_PLANT_SWAP_N", (check_snell, check_oblique_path, check_tir)),
    ("ORC_PLANT_FLIP_EY", (check_mirror,)),
])
def test_planted_mistake_is_caught(macro, checks, tmp_path):
    with O.planted(macro, tmp_path):
        for chk in checks:
            assert chk() != [], "%s passes %s" % (macro, chk.__name__)

"""Pins the oracle's own transcendentals against double-precision libm.

orc_logf / orc_cosf / orc_acosf are written from published series (named in
oracle/mcml_oracle.c); the product kernel implements the same algorithms and
must match them bit for bit (tests/test_gpu_parity.py).  Here the error bound
is checked, and the special cases the transport relies on.
"""
import math

import numpy as np
import pytest

import _oracle as O

f32 = np.float32


def ulp(x):
    x = abs(float(f32(x)))
    if x == 0:
        return float(np.finfo(f32).smallest_subnormal)
    return float(np.spacing(f32(x)))


def test_logf_error_bound():
    rng = np.random.default_rng(1)
    xs = np.concatenate([rng.random(20000), rng.random(2000) * 1e-6, np.exp(rng.uniform(-80, 80, 4000)),
                         [1.0, 0.5, 2.0, 1 - 2 ** -24, 2 ** -24, 1.41421354, 1.41421366, 0.70710677]]).astype(f32)
    xs = xs[xs > 0]
    worst = 0.0
    for x in xs:
        got = O.logf(x)
        ref = math.log(float(x))
        err = abs(got - ref) / max(ulp(ref), 2 ** -149)
        # near x = 1 log -> 0: measure against 1e-7 absolute there
        if abs(ref) < 1e-3:
            err = min(err, abs(got - ref) / 6e-8)
        worst = max(worst, err)
    assert worst <= 3.0, worst


def test_logf_special():
    assert O.logf(1.0) == 0.0
    assert O.logf(0.0) == -math.inf
    assert math.isnan(O.logf(-1.0))
    for x in (1e-45, 3e-40, 1.1e-38):  # subnormal inputs
        xf = float(f32(x))
        assert abs(O.logf(xf) - math.log(xf)) < 2e-5


def test_cosf_error_bound():
    rng = np.random.default_rng(2)
    # the kernel calls cos(2*pi*xi) with xi in [0, 1)
    xi = rng.random(30000).astype(f32)
    xs = np.concatenate([(f32(6.28318548) * xi).astype(f32), rng.uniform(-20, 20, 5000).astype(f32),
                         np.array([0, math.pi / 2, math.pi, 1.5 * math.pi, 2 * math.pi], f32)])
    worst = 0.0
    for x in xs:
        got = O.cosf(x)
        ref = math.cos(float(x))
        worst = max(worst, abs(got - ref))
    assert worst < 1.5e-7, worst
    assert O.cosf(0.0) == 1.0


def test_acosf_error_bound():
    rng = np.random.default_rng(3)
    xs = np.concatenate([rng.uniform(-1, 1, 30000), [-1, -0.5, 0, 0.5, 1, 1 - 2 ** -24, -1 + 2 ** -24]]).astype(f32)
    worst = 0.0
    for x in xs:
        got = O.acosf(x)
        ref = math.acos(float(x))
        worst = max(worst, abs(got - ref) / ulp(ref))
    assert worst <= 3.0, worst
    assert math.isnan(O.acosf(1.0000001))
    assert O.acosf(1.0) == 0.0


def test_sign_is_opencl_sign():
    assert O.lib().orc_sign(3.0) == 1.0 and O.lib().orc_sign(-2.0) == -1.0
    assert math.copysign(1, O.lib().orc_sign(-0.0)) == -1.0
    assert O.lib().orc_sign(float("nan")) == 0.0

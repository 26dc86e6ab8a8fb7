"""The FP32 propagator's small-angle polynomials (prop_f2 in
xflat_amd/csrc/xf_step_kernel.h) against the exact functions.

For x = lambda dl with x^2 < 1 the kernel forms cos x - 1 = t p(t) and
sin(x)/x - 1 = t q(t), t = x^2, with p and q cubics evaluated by FMA Horner in
FP32.  The constants are read from the kernel source and evaluated the way the
kernel does (each FMA rounded once to FP32); they must stay within FP32
rounding of the exact values over the whole range, so that the remainder
coefficients of the compensated error estimate keep their relative accuracy.
"""
import math
import os
import re

import numpy as np

from conftest import ROOT

SRC = os.path.join(ROOT, "xflat_amd", "csrc", "xf_step_kernel.h")
NUM = r"(-?[0-9.]+(?:e-?[0-9]+)?)f"


def _coefs(name):
    src = open(SRC).read()
    pat = (rf"const float {name} = fmaf\(x2, fmaf\(x2, fmaf\(x2, {NUM}, {NUM}\), {NUM}\), {NUM}\);")
    m = re.search(pat, src)
    assert m, f"polynomial {name} not found in {SRC}"
    c3, c2, c1, c0 = (np.float32(float(v)) for v in m.groups())
    return c0, c1, c2, c3


def _fma32(a, b, c):
    # a*b is exact in FP64 for FP32 operands; one rounding to FP32 after the add
    return np.float32(float(a) * float(b) + float(c))


def _horner32(c, t):
    c0, c1, c2, c3 = c
    return _fma32(t, _fma32(t, _fma32(t, c3, c2), c1), c0)


def _grid():
    t = np.concatenate([np.linspace(0.0, 1.0, 4001)[:-1], np.geomspace(1e-12, 0.999, 400)])
    return np.unique(t.astype(np.float32))


def _p_exact(t):
    # (cos sqrt(t) - 1) / t from its Taylor series (no cancellation)
    return math.fsum((-1) ** (k + 1) * t ** k / math.factorial(2 * k + 2) for k in range(16))


def _q_exact(t):
    # (sin sqrt(t) / sqrt(t) - 1) / t
    return math.fsum((-1) ** (k + 1) * t ** k / math.factorial(2 * k + 3) for k in range(16))


def test_taylor_helpers_match_libm():
    # the Taylor helpers against libm where cancellation is mild
    for t in (0.25, 0.5, 0.9):
        x = math.sqrt(t)
        assert abs(_p_exact(t) - (math.cos(x) - 1) / t) < 1e-14
        assert abs(_q_exact(t) - (math.sin(x) / x - 1) / t) < 1e-14


def test_cos_polynomial_within_fp32_rounding():
    c = _coefs("pc")
    assert c[0] == np.float32(-0.5)
    worst = max(abs(float(_horner32(c, t)) - _p_exact(float(t))) / abs(_p_exact(float(t))) for t in _grid())
    assert worst < 6e-8, worst


def test_sinc_polynomial_within_fp32_rounding():
    c = _coefs("ps")
    assert c[0] == np.float32(-1.0 / 6.0)
    worst = max(abs(float(_horner32(c, t)) - _q_exact(float(t))) / abs(_q_exact(float(t))) for t in _grid())
    assert worst < 1e-7, worst


def test_remainder_coefficient_relative_accuracy():
    # scr = dl * t * q(t) approximates sin(x)/lambda - dl with error relative to
    # its own size (x^2 dl / 6), the property the error estimate relies on
    cq = _coefs("ps")
    for lam, dl in ((1e3, 1e-4), (3e4, 2.2e-5), (2e5, 4e-6), (1e2, 9e-3)):
        x = lam * dl
        t = np.float32(np.float32(lam * lam) * np.float32(dl * dl))
        scr = float(np.float32(np.float32(dl) * np.float32(t * _horner32(cq, t))))
        exact = dl * float(t) * _q_exact(float(t))
        assert x < 1.0
        assert abs(scr - exact) <= 3e-7 * abs(exact), (lam, dl, scr, exact)

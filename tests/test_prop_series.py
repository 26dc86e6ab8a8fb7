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
import pytest

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


# ---- the FP32 wave-uniform small-angle tier and the FP64 series tiers ----
from fractions import Fraction as Fr  # noqa: E402


def _block(src, start, end):
    i = src.index(start)
    return src[i:src.index(end, i)]


def test_fp32_linear_tier_within_its_stated_bound():
    """When every lane of the wave has t = x^2 < 2^-10, prop_f2s takes the
    linear truncations (cos x - 1)/t = -1/2 + t/24 and (sin x/x - 1)/t =
    -1/6 + t/120: their omitted t^2/720 and t^2/5040 stay below 2.7e-9 and
    1.2e-9 relative, under FP32 rounding (evaluated as the kernel does, one
    FMA rounded once)."""
    src = open(SRC).read()
    blk = _block(src, "if (__all(x2 < 0x1p-10f))", "} else if (x2 < 1.f)")
    m = re.findall(r"const float (pc|ps) = fmaf\(x2, " + NUM + ", " + NUM + r"\);", blk)
    assert [x[0] for x in m] == ["pc", "ps"], m
    coef = {k: (np.float32(float(c0)), np.float32(float(c1))) for k, c1, c0 in m}
    assert coef["pc"][0] == np.float32(-0.5) and coef["ps"][0] == np.float32(-1.0 / 6.0)
    assert coef["pc"][1] == np.float32(1.0 / 24.0) and coef["ps"][1] == np.float32(1.0 / 120.0)
    ts = np.geomspace(1e-12, 2.0 ** -10, 300).astype(np.float32)
    # FP32 rounding bounds as for the cubics: -1/6 itself is 3e-8 off in FP32
    for k, exact, trunc, rnd in (("pc", _p_exact, 2.7e-9, 6e-8), ("ps", _q_exact, 1.2e-9, 1e-7)):
        c0, c1 = coef[k]
        lin = (lambda t: -0.5 + t / 24) if k == "pc" else (lambda t: -1 / 6 + t / 120)
        worst_trunc = max(abs(lin(float(t)) - exact(float(t))) / abs(exact(float(t))) for t in ts)
        assert worst_trunc < trunc, (k, worst_trunc)
        worst = max(abs(float(_fma32(t, c1, c0)) - exact(float(t))) / abs(exact(float(t))) for t in ts)
        assert worst < rnd + trunc, (k, worst)


def _fma64(a, b, c):
    # a*b + c rounded once to FP64 (exact rational arithmetic)
    return float(Fr(a) * Fr(b) + Fr(c))


def _chains(blk):
    """The two Horner chains (sn, cs) of one tier of prop_ds, as written:
    'double X = C;' then 'X = fma(u, [-]X, C);' with C = [-]1.0 / N or 0.5."""
    const = r"(-?1\.0 / [0-9.]+|0\.5)"
    out = {"sn": [], "cs": []}
    for m in re.finditer(r"double (sn|cs) = " + const + r";|(sn|cs) = fma\(u, (-?)(sn|cs), " + const + r"\)\s*;", blk):
        if m.group(1):
            out[m.group(1)].append(("init", m.group(2)))
        else:
            assert m.group(3) == m.group(5)
            out[m.group(3)].append((m.group(4), m.group(6)))

    def val(c):
        if c == "0.5":
            return Fr(1, 2)
        sign = -1 if c.startswith("-") else 1
        return sign * Fr(1) / Fr(c.split("/")[1].strip())
    return {k: [(op, val(c)) for op, c in v] for k, v in out.items()}


def _eval_chain(ch, u, exact):
    """exact=True: rational arithmetic (the truncated series itself);
    False: every FMA rounded once to FP64, constants rounded to FP64 as
    the compiler folds 1.0 / N."""
    x = None
    for op, c in ch:
        cc = c if exact else Fr(float(c))
        if op == "init":
            x = cc
        else:
            xx = -x if op == "-" else x
            x = (Fr(u) * xx + cc) if exact else Fr(_fma64(float(u), float(xx), float(cc)))
    return x


def _cos_sinc_exact(u, terms=30):
    # cos x and sin(x)/x for x^2 = u, exact rationals to far below FP64
    c = sum(Fr((-1) ** k) * Fr(u) ** k / math.factorial(2 * k) for k in range(terms))
    s = sum(Fr((-1) ** k) * Fr(u) ** k / math.factorial(2 * k + 1) for k in range(terms))
    return c, s


@pytest.mark.parametrize("tier,umax,trunc", [(0, 2.0 ** -10, 2.5e-22), (1, 0.0625, 1e-23)])
def test_fp64_series_tiers(tier, umax, trunc):
    """prop_ds below u = x^2 = 2^-10 (to x^8 / x^9) and below 1/16 (to x^16 /
    x^17): the truncated series differ from cos x and sin(x)/x by less than
    the stated remainders, and the FP64 evaluation (FMA Horner, as written)
    is within 2 ulp of them."""
    src = open(SRC).read()
    body = _block(src, "__device__ __forceinline__ void prop_ds(", "\n}\n")
    blocks = [_block(body, "if (u < 0x1p-10) {", "} else if (u < 0.0625)"),
              _block(body, "} else if (u < 0.0625) {", "} else {")]
    ch = _chains(blocks[tier])
    assert len(ch["sn"]) == (4 if tier == 0 else 8) and len(ch["cs"]) == (4 if tier == 0 else 8), ch
    assert "c = fma(-u, cs, 1.0);" in blocks[tier] and "sc = dl * fma(-u, sn, 1.0);" in blocks[tier]
    for u in [umax * f for f in (1e-6, 1e-3, 0.1, 0.5, 0.999)]:
        u = float(u)
        ce, se = _cos_sinc_exact(u)
        cs_x, sn_x = _eval_chain(ch["cs"], u, True), _eval_chain(ch["sn"], u, True)
        c_poly, s_poly = 1 - Fr(u) * cs_x, 1 - Fr(u) * sn_x
        assert abs(float(c_poly - ce)) < trunc and abs(float(s_poly - se)) < trunc, (u, float(c_poly - ce))
        cs_f, sn_f = _eval_chain(ch["cs"], u, False), _eval_chain(ch["sn"], u, False)
        c64 = _fma64(-u, float(cs_f), 1.0)
        s64 = _fma64(-u, float(sn_f), 1.0)
        assert abs(c64 - float(ce)) <= 2 * math.ulp(float(ce)), (u, c64, float(ce))
        assert abs(s64 - float(se)) <= 2 * math.ulp(float(se)), (u, s64, float(se))

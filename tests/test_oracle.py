"""Pins the CPU oracle (oracle/xflat_oracle.c) before it is trusted.

1. Against the reference itself: golden fixtures written by
   tests/golden/make_golden.py from the reference's own evolution path
   (oracle/_ref build of src/Nmr.cpp, NBeam.cpp, Phy_MA.cpp, Phy_MAA.cpp...):
   the accepted-step sequence, r, dr and psi must agree.
2. Against the mathematics, for what a trajectory comparison leaves open:
   U = expm(-i H dl) (NBeam.cpp:360-393), the angular moment factorisation vs
   the brute-force (1 - v.v') double sum (Phy_MA.cpp:372-446,
   Phy_MAA.cpp:445-562), the SU(2)-partner invariance of the nu_x columns,
   the two-flavour vacuum oscillation along the bulb's straight rays, and the
   single-angle H_vv (Phy_SA.cpp:250-283) as the continuum limit of the
   multi-angle integral on the radial trajectory.
"""
import os

import numpy as np
import pytest
import scipy.linalg

from conftest import GOLDEN, golden_files
from oracle import oracle as O
from xflat_amd.config import parse_config_text, bulb_params, XF_MA, XF_MAA, XF_SA
from xflat_amd.seed import seeded_psi


def _load(name):
    g = np.load(os.path.join(GOLDEN, name))
    p = parse_config_text(str(g["config"]), module=int(g["module"]))
    return g, p


@pytest.mark.parametrize("name", golden_files())
def test_oracle_matches_reference_golden(name):
    g, p = _load(name)
    o = O.Oracle(p)
    o.set_state(g["psi0"], float(g["r0"]), float(g["dr0"]))
    its = g["it"]
    kept = {int(its[i]): k for k, i in enumerate(g["psi_idx"])}
    done = 0
    acc_all = []
    for it_target in sorted(kept):
        n, rec, _ = o.run(it_target + 1 - done)
        assert n == it_target + 1 - done
        acc_all.extend((done + np.nonzero(rec["accepted"])[0]).tolist())
        # rec holds attempts done..it_target; dr/r of the accepted ones
        done = it_target + 1
        psi, r, dr, _ = o.state()
        ref = g["psi"][kept[it_target]]
        assert np.abs(psi - ref).max() <= 1e-13, f"psi differs at attempt {it_target}"
    n, rec, _ = o.run(int(g["total"]) - done)
    acc_all.extend((done + np.nonzero(rec["accepted"])[0]).tolist())
    # same accepted-step sequence as the reference
    assert np.array_equal(np.array(acc_all), its)
    _, r, dr, (tot, acc) = o.state()
    assert tot == int(g["total"]) and acc == int(g["accepted"])
    assert abs(r - g["r"][-1]) <= 1e-12 * abs(r)


def test_oracle_r_dr_sequence_bitwise():
    g, p = _load("ma_r10_ih.npz")
    o = O.Oracle(p)
    o.set_state(g["psi0"], float(g["r0"]), float(g["dr0"]))
    n, rec, _ = o.run(int(g["total"]))
    acc = rec["accepted"].astype(bool)
    assert np.array_equal((rec["r"] + rec["dr"])[acc], g["r"])
    assert np.array_equal(rec["dr"][acc], g["dr"])
    # the very first attempt (dr = 1e-3 km at R_nu) is rejected
    assert not acc[0] and acc[1:].all()


def _H(h_r0, h_r1, h_i1):
    return np.array([[h_r0, h_r1 + 1j * h_i1], [h_r1 - 1j * h_i1, -h_r0]])


@pytest.mark.parametrize("seed", range(6))
def test_U_is_expm(seed):
    rng = np.random.default_rng(seed)
    h = rng.normal(size=3) * 10.0 ** rng.uniform(-2, 4)
    dl = 10.0 ** rng.uniform(-6, -1)
    psi = rng.normal(size=4)
    out = O.U(dl, *h, psi)
    Um = scipy.linalg.expm(-1j * _H(*h) * dl)
    v = Um @ np.array([psi[0] + 1j * psi[1], psi[2] + 1j * psi[3]])
    np.testing.assert_allclose(out, [v[0].real, v[0].imag, v[1].real, v[1].imag], atol=1e-13 * np.abs(psi).max())


def test_U_unitary_and_composition():
    h = (0.7, -0.3, 0.2)
    psi = np.array([0.6, 0.0, 0.0, 0.8])
    out = O.U(0.37, *h, psi)
    assert abs(np.sum(out ** 2) - 1.0) < 1e-15
    twice = O.U(0.37, *h, O.U(0.37, *h, psi))
    np.testing.assert_allclose(twice, O.U(0.74, *h, psi), atol=1e-15)


def test_density():
    a, b = 0.6 + 0.1j, 0.3 - 0.7j
    d = O.density([a.real, a.imag, b.real, b.imag])
    ab = a * np.conj(b)
    np.testing.assert_allclose(d, [0.5 * (abs(a) ** 2 - abs(b) ** 2), 0.0, ab.real, ab.imag], atol=1e-16)


def _brute_hvv(o, p, k_set, pt, r):
    """H_vv(i) = coef * sum_j dcos_j X_j (1 - v_i . v_j), direct O(N^2)."""
    t = o.tables()
    L = t["L_E"]
    hs = o.get_hsum(k_set)                 # [beam][prtcl][4]
    X = (hs[:, 0] * L[0] - hs[:, 1] * L[1]) + (hs[:, 2] * L[2] - hs[:, 3] * L[3])
    X[:, 3] = (hs[:, 0, 3] * L[0] + hs[:, 1, 3] * L[1]) + (hs[:, 2, 3] * L[2] + hs[:, 3, 3] * L[3])
    nth, nph = p.abins, (p.pbins if p.module == XF_MAA else 1)
    x = t["cos2R"]
    ra = p.Rv / r
    c = np.sqrt(1 - ra * ra * (1 - x))
    s = ra * np.sqrt(1 - x)
    w = 0.5 / c * ra * ra * (1.0 / nth)
    if p.module == XF_MAA:
        phi = (np.arange(nph) + 0.5) * (2 * 3.14159265359 / nph)
        vx = (s[:, None] * np.cos(phi)[None, :]).ravel()
        vy = (s[:, None] * np.sin(phi)[None, :]).ravel()
        vz = np.repeat(c, nph)
        wj = np.repeat(w, nph)
    else:
        vx = np.zeros(nth); vy = np.zeros(nth); vz = c; wj = w
    dot = vx[:, None] * vx[None, :] + vy[:, None] * vy[None, :] + vz[:, None] * vz[None, :]
    return t["coef"] * ((1 - dot) * wj[None, :]) @ X


@pytest.mark.parametrize("module", [XF_MA, XF_MAA])
def test_hvv_moments_equal_bruteforce_pair_sum(module):
    over = dict(Abins=12, Ebins=8, R0=20)
    if module == XF_MAA:
        over["Pbins"] = 5
    p = bulb_params(module=module, **over)
    o = O.Oracle(p)
    nbeam = o.nbeam
    rng = np.random.default_rng(5)
    psi = rng.normal(size=(nbeam, 4, 4, p.ebins))       # generic, symmetry-breaking
    o.set_psi(0, psi)
    o.calc_esum(0)
    r = 23.0
    o.calc_angle_bins(r, 1)
    o.calc_hvv(0, 1, 2)
    got = o.get_hvv(2)
    want = _brute_hvv(o, p, 0, 1, r)
    scale = np.abs(want).max()
    assert scale > 0
    np.testing.assert_allclose(got[:, [0, 2, 3]], want[:, [0, 2, 3]], atol=1e-12 * scale)


def test_partner_columns_stay_partners():
    p = bulb_params(Abins=16, Ebins=12, R0=10, Ts=100)
    o = O.Oracle(p)
    psi0 = seeded_psi(16, 12, seed=3, amp=1e-3)
    o.set_state(psi0, 10.0, 1e-3)
    o.run(60)
    psi, *_ = o.state()
    for s in (0, 1):
        e, x = psi[:, s], psi[:, s + 2]
        # x = (-conj(b), conj(a))
        assert np.abs(x[:, 0] + e[:, 2]).max() < 1e-14
        assert np.abs(x[:, 1] - e[:, 3]).max() < 1e-14
        assert np.abs(x[:, 2] - e[:, 0]).max() < 1e-14
        assert np.abs(x[:, 3] + e[:, 1]).max() < 1e-14


def test_vacuum_oscillation_closed_form():
    """No matter, no nu-nu term: along ray theta the path length from R0 to R
    is sqrt(R^2 - b^2) - sqrt(R0^2 - b^2) with b^2 = Rv^2 (1 - x_theta), and
    P_ee = 1 - sin^2(2 theta_v) sin^2(delta L / 2), delta = dm2 / (2E)."""
    p = bulb_params(Abins=8, Ebins=10, R0=20, Rn=250, hasMatter=0,
                    Lve=0, Lv_e=0, Lvt=0, Lv_t=0, dr=".01", eps0="1.E-9")
    o = O.Oracle(p)
    o.init_state_pure()
    o.run(3000)
    psi, r, _, _ = o.state()
    t = o.tables()
    ne = p.ebins
    E = p.E0 + (np.arange(ne) + 0.5) * (p.E1 - p.E0) / ne
    delta = p.dm2 / (2 * E)
    b2 = p.Rv ** 2 * (1 - t["cos2R"])
    Lpath = np.sqrt(r * r - b2) - np.sqrt(p.R0 ** 2 - b2)
    pee = 1 - np.sin(p.theta2) ** 2 * np.sin(0.5 * delta[None, :] * Lpath[:, None]) ** 2
    got = psi[:, 0, 0] ** 2 + psi[:, 0, 1] ** 2
    assert r > 21.0
    np.testing.assert_allclose(got, pee, atol=2e-6)


def _radial_beam_hvv(state, nth, r):
    ne = state.shape[-1]
    pm = bulb_params(module=XF_MA, Abins=nth, Ebins=ne, R0=10)
    om = O.Oracle(pm)
    om.set_psi(0, np.broadcast_to(state, (nth, 4, 4, ne)).copy())
    om.calc_esum(0)
    om.calc_angle_bins(r, 1)
    om.calc_hvv(0, 1, 2)
    return om.get_hvv(2)[0]


@pytest.mark.parametrize("r", [10.0, 14.0, 40.0])
def test_sa_hvv_is_radial_limit_of_multi_angle(r):
    """With the same source X on every angle bin, the MA H_vv of the beam
    nearest the radial direction is coef * sum_j w_j X (1 - c_j c_0), whose
    continuum limit (c_0 -> 1, sum -> integral over c from sqrt(1 - (Rv/r)^2)
    to 1 of (1 - c) dc) is coef * D(r) * X with D = (1 - sqrt(1 - (Rv/r)^2))^2 / 2:
    the single-angle H_vv.  The MA value must converge to the SA one as the
    angle grid is refined -- first order in 1/Abins for r > Rv, half order at
    r = Rv where the weights 1/c_j are singular at the grid's edge -- which
    checks D(r), the SA coefficient and the sign conventions against an
    independent (multi-angle) code path."""
    ne = 6
    rng = np.random.default_rng(17)
    state = rng.normal(size=(4, 4, ne))               # one beam's 4 NBeams
    pa = bulb_params(module=XF_SA, Abins=1, Ebins=ne, R0=10)
    oa = O.Oracle(pa)
    assert oa.nbeam == 1
    oa.set_psi(0, state[None])
    oa.calc_esum(0)
    oa.calc_angle_bins(r, 1)
    oa.calc_hvv(0, 1, 2)
    sa = oa.get_hvv(2)[0][[0, 2, 3]]
    D = oa.L.oracle_sa_D(pa.Rv, r)
    assert abs(D - 0.5 * (1 - np.sqrt(1 - (pa.Rv / r) ** 2)) ** 2) < 1e-15
    scale = np.abs(sa).max()
    assert scale > 0
    err = [np.abs(_radial_beam_hvv(state, n, r)[[0, 2, 3]] - sa).max() / scale for n in (5000, 20000)]
    if r > pa.Rv:
        assert err[1] < 1e-4 and err[1] < 0.3 * err[0], err
    else:
        assert err[1] < 5e-3 and err[1] < 0.6 * err[0], err

"""Seeded initial states (host side).

The reference starts every beam in a pure flavour state (NBeam.cpp:246-276).
The benchmark configurations add a seeded off-diagonal perturbation of
amplitude `amp`: for every beam, antineutrino flag sp and energy bin e,

    eps = amp * ((2 u0 - 1) + i (2 u1 - 1)),   u_k uniform in [0, 1)
    psi_nu_e = (sqrt(1 - |eps|^2), eps),  psi_nu_x = (-conj(eps), sqrt(1 - |eps|^2))

so nu_x (anti-nu_x) is the SU(2) partner (-conj(b), conj(a)) of nu_e
(anti-nu_e).  u_k is the (n+1)-th splitmix64 output of state `seed` with
n = (((inst*2 + sp)*nbeams + beam)*NE + e)*2 + k, mapped by (z >> 11) 2^-53.
"""
from __future__ import annotations

import numpy as np

_GOLDEN = np.uint64(0x9E3779B97F4A7C15)
_M1 = np.uint64(0xBF58476D1CE4E5B9)
_M2 = np.uint64(0x94D049BB133111EB)


def splitmix64(seed: int, n: np.ndarray) -> np.ndarray:
    """(n+1)-th output of splitmix64 started at `seed` (vectorised)."""
    with np.errstate(over="ignore"):
        z = np.uint64(seed & 0xFFFFFFFFFFFFFFFF) + (n.astype(np.uint64) + np.uint64(1)) * _GOLDEN
        z = (z ^ (z >> np.uint64(30))) * _M1
        z = (z ^ (z >> np.uint64(27))) * _M2
        z = z ^ (z >> np.uint64(31))
    return z


def uniform01(seed: int, n: np.ndarray) -> np.ndarray:
    return (splitmix64(seed, n) >> np.uint64(11)).astype(np.float64) * (2.0 ** -53)


def seeded_psi(nbeams: int, ne: int, seed: int, amp: float, inst: int = 0) -> np.ndarray:
    """psi[beam][prtcl][comp][ebin] with the seeded perturbation above."""
    psi = np.zeros((nbeams, 4, 4, ne))
    beam = np.arange(nbeams)[:, None]
    e = np.arange(ne)[None, :]
    for sp in (0, 1):
        base = (((inst * 2 + sp) * nbeams + beam) * ne + e) * 2
        u0 = uniform01(seed, base)
        u1 = uniform01(seed, base + 1)
        er = amp * (2.0 * u0 - 1.0)
        ei = amp * (2.0 * u1 - 1.0)
        a = np.sqrt(1.0 - (er * er + ei * ei))
        psi[:, sp, 0] = a
        psi[:, sp, 1] = 0.0
        psi[:, sp, 2] = er
        psi[:, sp, 3] = ei
        psi[:, sp + 2, 0] = -er
        psi[:, sp + 2, 1] = ei
        psi[:, sp + 2, 2] = a
        psi[:, sp + 2, 3] = 0.0
    return psi


def pure_psi(nbeams: int, ne: int) -> np.ndarray:
    """Pure flavour states of NBeam::NBeam(int) (NBeam.cpp:246-276)."""
    psi = np.zeros((nbeams, 4, 4, ne))
    psi[:, 0:2, 0] = 1.0
    psi[:, 2:4, 2] = 1.0
    return psi


def azimuthal_psi(nth: int, nph: int, ne: int, seed: int, amp: float = 1e-7, mmax: int = 4,
                  form: str = "cos") -> np.ndarray:
    """MAA initial state with an azimuthal-symmetry-breaking perturbation of
    the multi-azimuthal workload: eps(phi) = amp * sum_{m=1..mmax}
    cos(m phi + phi_m), real-valued (form "cos"), or
    amp * sum_m exp(i (m phi + phi_m)) (form "exp"); phi_m = 2 pi u_m (u_m
    the m-th output of the splitmix64 stream `seed`, separately for the
    neutrinos and the antineutrinos), phi = (p + 1/2) 2 pi / Pbins the
    reference's azimuth grid (Phy_MAA.cpp:177-185); the same for every polar
    bin and energy.  Partner columns as in seeded_psi."""
    phi = (np.arange(nph) + 0.5) * 2.0 * np.pi / nph
    psi = np.zeros((nth * nph, 4, 4, ne))
    for sp in (0, 1):
        ph = 2.0 * np.pi * uniform01(seed, np.arange(mmax) + sp * mmax)
        if form == "cos":
            eps = amp * sum(np.cos(m * phi + ph[m - 1]) for m in range(1, mmax + 1)) + 0j
        elif form == "exp":
            eps = amp * sum(np.exp(1j * (m * phi + ph[m - 1])) for m in range(1, mmax + 1))
        else:
            raise ValueError(form)
        er = np.repeat(np.tile(eps.real, nth)[:, None], ne, axis=1)
        ei = np.repeat(np.tile(eps.imag, nth)[:, None], ne, axis=1)
        a = np.sqrt(1.0 - (er * er + ei * ei))
        psi[:, sp, 0] = a
        psi[:, sp, 2] = er
        psi[:, sp, 3] = ei
        psi[:, sp + 2, 0] = -er
        psi[:, sp + 2, 1] = ei
        psi[:, sp + 2, 2] = a
    return psi


def ensemble_configs(n: int = 96, seed: int = 4242, abins=(64, 1500), ebins=(32, 400)) -> list[dict]:
    """Per-instance config overrides of the batched-ensemble workload: angle
    bins uniform in abins, energy bins uniform in ebins, each flavour's
    luminosity uniform in 0.5-3e51 erg/s, theta_v uniform in [0.01, 0.6],
    hierarchy +/- with equal odds, and an electron-density prefactor
    log-uniform in 1e31-1e33 cm^-3 at R_nu (the bulb's far term scales with gs:
    gs = 5.5 gives 1e32).  Draw k of instance i is output 16 i + k of the
    splitmix64 stream `seed`."""
    out = []
    for i in range(n):
        u = uniform01(seed, 16 * i + np.arange(16))
        nth = int(abins[0] + np.floor(u[0] * (abins[1] - abins[0] + 1)))
        ne = int(ebins[0] + np.floor(u[1] * (ebins[1] - ebins[0] + 1)))
        L = [0.5e51 + u[2 + k] * 2.5e51 for k in range(4)]
        theta = 0.01 + u[6] * 0.59
        dm2 = 2.4e-3 if u[7] < 0.5 else -2.4e-3
        ne0 = 10.0 ** (31.0 + 2.0 * u[8])
        out.append(dict(Abins=nth, Ebins=ne, Lve=f"{L[0]:.6e}", Lv_e=f"{L[1]:.6e}", Lvt=f"{L[2]:.6e}",
                        Lv_t=f"{L[3]:.6e}", theta=f"{theta:.6f}", dm2=f"{dm2:.1e}", gs=f"{5.5 * ne0 / 1e32:.6e}"))
    return out

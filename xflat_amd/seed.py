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

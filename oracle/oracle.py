"""ctypes wrapper of the CPU FP64 oracle (oracle/xflat_oracle.c).

TEST INFRASTRUCTURE ONLY -- imported by tests/, __graft_entry__.smoke() and
the cpu_baseline leg of bench.py, always as the checker, never as the thing
measured or shipped.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

from xflat_amd.config import XfParams

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None


class XoRec(ctypes.Structure):
    _fields_ = [("r", ctypes.c_double), ("dr", ctypes.c_double), ("err", ctypes.c_double),
                ("accepted", ctypes.c_int32), ("pad", ctypes.c_int32)]


REC_DTYPE = np.dtype([("r", "<f8"), ("dr", "<f8"), ("err", "<f8"), ("accepted", "<i4"), ("pad", "<i4")])


def _cpu_has_avx512() -> bool:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("flags"):
                    return " avx512f " in (line + " ")
    except OSError:
        pass
    return False


def lib_path(isa: str | None = None) -> str:
    """Path of the built oracle; isa 'v4' (AVX-512) or 'v3' (AVX2)."""
    if isa is None:
        isa = "v4" if _cpu_has_avx512() else "v3"
    name = "libxflat_oracle_v4.so" if isa == "v4" else "libxflat_oracle.so"
    return os.path.join(_HERE, "build", name)


def load(isa: str = "v3"):
    """Loads the oracle (the v3 build by default: results are then the same
    bits on every x86-64 host)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = lib_path(isa)
    if not os.path.exists(path):
        raise RuntimeError(f"oracle not built: {path} (run __graft_entry__.build())")
    L = ctypes.CDLL(path)
    P = ctypes.c_void_p
    D = ctypes.c_double
    I = ctypes.c_int
    dp = ctypes.POINTER(ctypes.c_double)
    L.oracle_create.restype = P
    L.oracle_create.argtypes = [ctypes.POINTER(XfParams)]
    L.oracle_destroy.argtypes = [P]
    L.oracle_ne.argtypes = [P]
    L.oracle_nbeam.argtypes = [P]
    L.oracle_get_tables.argtypes = [P, dp, dp, dp, dp, dp, dp]
    L.oracle_get_hm.restype = D
    L.oracle_get_hm.argtypes = [P, D]
    L.oracle_init_pure.argtypes = [P, I]
    L.oracle_set_psi.argtypes = [P, I, dp]
    L.oracle_get_psi.argtypes = [P, I, dp]
    L.oracle_get_hsum.argtypes = [P, I, dp]
    L.oracle_get_hvv.argtypes = [P, I, dp]
    L.oracle_U.argtypes = [D, D, D, D, dp, dp]
    L.oracle_density.argtypes = [dp, dp]
    L.oracle_calc_angle_bins.argtypes = [P, D, I]
    L.oracle_calc_delta_ls.argtypes = [P, D, I, I, I]
    L.oracle_calc_hvv.argtypes = [P, I, I, I]
    L.oracle_evolve.argtypes = [P, I, I, I, D, I]
    L.oracle_evolve_hvv.argtypes = [P, I, I, I, D, I]
    L.oracle_evolve_avg.argtypes = [P, I, I, I, D, I, I]
    L.oracle_evolve_avg_inplace.argtypes = [P, I, I, D, I, I]
    L.oracle_avg_beam.argtypes = [P, I, I]
    L.oracle_calc_esum.argtypes = [P, I]
    L.oracle_calc_err_hvv.restype = D
    L.oracle_calc_err_hvv.argtypes = [P, I, I]
    L.oracle_set_state.argtypes = [P, dp, D, D]
    L.oracle_state_init_pure.argtypes = [P]
    L.oracle_get_state.argtypes = [P, dp, dp, dp, ctypes.POINTER(ctypes.c_int64)]
    L.oracle_run.restype = ctypes.c_int64
    L.oracle_run.argtypes = [P, ctypes.c_int64, ctypes.c_void_p, dp]
    L.oracle_max_threads.restype = I
    L.oracle_set_threads.argtypes = [I]
    L.oracle_sa_D.argtypes = [D, D]
    L.oracle_sa_D.restype = D
    _LIB = L
    return L


def _dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(ctypes.POINTER(ctypes.c_double))


class Oracle:
    """One oracle instance for an XfParams configuration."""

    def __init__(self, params: XfParams):
        self.L = load()
        self.p = params
        self.h = self.L.oracle_create(ctypes.byref(params))
        if not self.h:
            raise ValueError("oracle_create rejected the parameters")
        self.ne = self.L.oracle_ne(self.h)
        self.nbeam = self.L.oracle_nbeam(self.h)
        self.shape = (self.nbeam, 4, 4, self.ne)

    def close(self):
        if self.h:
            self.L.oracle_destroy(self.h)
            self.h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # ---- set-up tables
    def tables(self):
        fv = np.zeros((4, self.ne)); hv = np.zeros(2 * self.ne)
        le = np.zeros(4); ev = np.zeros(4)
        c2 = np.zeros(max(1, self.nbeam if self.p.module == 2 else self.p.abins)); coef = np.zeros(1)
        self.L.oracle_get_tables(self.h, _dp(fv), _dp(hv), _dp(le), _dp(ev), _dp(c2), _dp(coef))
        return dict(fv=fv, hvac=hv, L_E=le, Ev=ev, cos2R=c2, coef=coef[0])

    def hm(self, r: float) -> float:
        return self.L.oracle_get_hm(self.h, r)

    # ---- fine-grained beam sets
    def set_psi(self, k: int, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.float64).reshape(-1)
        assert a.size == int(np.prod(self.shape))
        self.L.oracle_set_psi(self.h, k, _dp(a))

    def get_psi(self, k: int) -> np.ndarray:
        a = np.zeros(self.shape)
        self.L.oracle_get_psi(self.h, k, _dp(a))
        return a

    def get_hsum(self, k: int) -> np.ndarray:
        a = np.zeros((self.nbeam, 4, 4))
        self.L.oracle_get_hsum(self.h, k, _dp(a))
        return a

    def get_hvv(self, k: int) -> np.ndarray:
        a = np.zeros((self.nbeam, 4))
        self.L.oracle_get_hvv(self.h, k, _dp(a))
        return a

    def init_pure(self, k): self.L.oracle_init_pure(self.h, k)
    def calc_esum(self, k): self.L.oracle_calc_esum(self.h, k)
    def calc_angle_bins(self, r, pt): self.L.oracle_calc_angle_bins(self.h, r, pt)
    def calc_delta_ls(self, dr, cur, s, e): self.L.oracle_calc_delta_ls(self.h, dr, cur, s, e)
    def calc_hvv(self, k, pt, hv): self.L.oracle_calc_hvv(self.h, k, pt, hv)
    def evolve(self, i, pt, hv, hm, o): self.L.oracle_evolve(self.h, i, pt, hv, hm, o)
    def evolve_hvv(self, i, pt, hv, hm, o): self.L.oracle_evolve_hvv(self.h, i, pt, hv, hm, o)
    def evolve_avg(self, i, pt, hv, hm, o, a): self.L.oracle_evolve_avg(self.h, i, pt, hv, hm, o, a)
    def evolve_avg_inplace(self, pt, hv, hm, io, a): self.L.oracle_evolve_avg_inplace(self.h, pt, hv, hm, io, a)
    def avg_beam(self, i, o): self.L.oracle_avg_beam(self.h, i, o)
    def calc_err_hvv(self, a, b) -> float: return self.L.oracle_calc_err_hvv(self.h, a, b)

    # ---- stepper
    def set_state(self, psi: np.ndarray, r: float, dr: float):
        a = np.ascontiguousarray(psi, dtype=np.float64).reshape(-1)
        self.L.oracle_set_state(self.h, _dp(a), r, dr)

    def init_state_pure(self):
        self.L.oracle_state_init_pure(self.h)

    def state(self):
        psi = np.zeros(self.shape)
        r = ctypes.c_double(); dr = ctypes.c_double()
        cnt = (ctypes.c_int64 * 2)()
        self.L.oracle_get_state(self.h, _dp(psi), ctypes.byref(r), ctypes.byref(dr), cnt)
        return psi, r.value, dr.value, (cnt[0], cnt[1])

    def run(self, max_steps: int, log: bool = True):
        """Runs up to max_steps attempts; returns (n, records, loop_seconds)."""
        rec = np.zeros(max(max_steps, 1), dtype=REC_DTYPE) if log else None
        secs = ctypes.c_double()
        n = self.L.oracle_run(self.h, max_steps, rec.ctypes.data if log else None, ctypes.byref(secs))
        return n, (rec[:n] if log else None), secs.value


def U(dl, h_r0, h_r1, h_i1, psi4):
    L = load()
    a = np.ascontiguousarray(psi4, dtype=np.float64)
    o = np.zeros(4)
    L.oracle_U(dl, h_r0, h_r1, h_i1, _dp(a), _dp(o))
    return o


def density(psi4):
    L = load()
    a = np.ascontiguousarray(psi4, dtype=np.float64)
    o = np.zeros(4)
    L.oracle_density(_dp(a), _dp(o))
    return o


def set_threads(n: int):
    load().oracle_set_threads(n)


def max_threads() -> int:
    return load().oracle_max_threads()

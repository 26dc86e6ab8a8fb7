"""ctypes binding of the C-ABI in include/xflat.h (libxflat.so, built in-tree).

The product path is the HIP library: if it is missing this module raises,
there is no CPU fallback.
"""
from __future__ import annotations

import atexit
import ctypes
import os
import sys
import weakref

import numpy as np

from .config import XfParams, XF_FP64, XF_MIXED, XF_FP32, n_beams  # noqa: F401

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libxflat.so")
_LIB = None

XF_OK = 0
ERRORS = {-1: "XF_EINVAL", -2: "XF_EHIP", -3: "XF_ENOMEM", -4: "XF_ESTATE", -5: "XF_EIO",
          -6: "XF_EBARRIER", -7: "XF_ENODEV"}


class XfError(RuntimeError):
    def __init__(self, code, msg):
        super().__init__(f"{ERRORS.get(code, code)}: {msg}")
        self.code = code


class XfOpts(ctypes.Structure):
    _fields_ = [("precision", ctypes.c_int), ("species", ctypes.c_int), ("renorm", ctypes.c_int),
                ("block", ctypes.c_int), ("cpt", ctypes.c_int), ("resident", ctypes.c_int),
                ("grid", ctypes.c_int), ("mode", ctypes.c_int)]


class XfStepRec(ctypes.Structure):
    _fields_ = [("r", ctypes.c_double), ("dr", ctypes.c_double), ("err", ctypes.c_double),
                ("accepted", ctypes.c_int32), ("pad", ctypes.c_int32)]


STEP_DTYPE = np.dtype([("r", "<f8"), ("dr", "<f8"), ("err", "<f8"), ("accepted", "<i4"), ("pad", "<i4")])


class XfRunInfo(ctypes.Structure):
    _fields_ = [("steps", ctypes.c_int64), ("accepted", ctypes.c_int64), ("total_steps", ctypes.c_int64),
                ("total_accepted", ctypes.c_int64), ("r", ctypes.c_double), ("dr", ctypes.c_double),
                ("last_err", ctypes.c_double), ("last_dr", ctypes.c_double), ("kernel_ms", ctypes.c_double),
                ("stop_reason", ctypes.c_int32), ("pad", ctypes.c_int32)]


class XfStop(ctypes.Structure):
    _fields_ = [("dump_rstep", ctypes.c_double), ("dump_last_r", ctypes.c_double),
                ("dump_itrstep", ctypes.c_int), ("pad", ctypes.c_int)]


# name -> (restype, argtypes)
_P, _I, _D = ctypes.c_void_p, ctypes.c_int, ctypes.c_double
_DP = ctypes.POINTER(ctypes.c_double)
_PP = ctypes.POINTER(XfParams)
_OP = ctypes.POINTER(XfOpts)
SIGNATURES = {
    "xf_last_error": (ctypes.c_char_p, []),
    "xf_version": (_I, []),
    "xf_device_count": (_I, [ctypes.POINTER(_I)]),
    "xf_parse_config": (_I, [ctypes.c_char_p, _PP]),
    "xf_parse_config_text": (_I, [ctypes.c_char_p, _PP]),
    "xf_create": (_I, [_PP, _OP, ctypes.POINTER(_P)]),
    "xf_destroy": (_I, [_P]),
    "xf_beam_len": (_I, [_P]),
    "xf_get_dim": (_I, [_P]),
    "xf_count_dim": (_I, [_P, ctypes.POINTER(ctypes.c_size_t)]),
    "xf_get_spectra": (_I, [_P, _DP, _DP, _DP, _DP]),
    "xf_get_geometry": (_I, [_P, _DP, _DP]),
    "xf_get_hm": (_I, [_P, _D, _DP]),
    "xf_beams_upload": (_I, [_P, _I, _DP]),
    "xf_beams_download": (_I, [_P, _I, _DP]),
    "xf_beams_init_pure": (_I, [_P, _I]),
    "xf_calc_esum": (_I, [_P, _I]),
    "xf_get_esum": (_I, [_P, _I, _DP]),
    "xf_calc_angle_bins": (_I, [_P, _D, _I]),
    "xf_calc_delta_ls": (_I, [_P, _D, _I, _I, _I]),
    "xf_calc_hvv": (_I, [_P, _I, _I, _I]),
    "xf_get_hvv": (_I, [_P, _I, _DP]),
    "xf_evolve": (_I, [_P, _I, _I, _I, _D, _I]),
    "xf_evolve_hvv": (_I, [_P, _I, _I, _I, _D, _I]),
    "xf_evolve_avg": (_I, [_P, _I, _I, _I, _D, _I, _I]),
    "xf_evolve_avg_inplace": (_I, [_P, _I, _I, _D, _I, _I]),
    "xf_avg_beam": (_I, [_P, _I, _I]),
    "xf_calc_err_hvv": (_I, [_P, _I, _I, _DP]),
    "xf_state_upload": (_I, [_P, _DP, _D, _D]),
    "xf_state_init_pure": (_I, [_P]),
    "xf_state_download": (_I, [_P, _DP, _DP, _DP]),
    "xf_state_counters": (_I, [_P, ctypes.POINTER(ctypes.c_int64), ctypes.POINTER(ctypes.c_int64)]),
    "xf_state_species": (_I, [_P]),
    "xf_set_rn": (_I, [_P, _D]),
    "xf_run_steps": (_I, [_P, _I, ctypes.POINTER(XfStop), _P, ctypes.POINTER(XfRunInfo)]),
    "xf_run_steps_noalloc": (_I, [_P, _I, ctypes.POINTER(XfStop), _P, ctypes.POINTER(XfRunInfo)]),
    "xf_run_graph": (_I, [_P, _I, _I, ctypes.POINTER(XfRunInfo)]),
    "xf_get_launch": (_I, [_P, _OP, ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "xf_set_launch": (_I, [_P, _I, _I, _I, _I]),
    "xf_launch_info": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    "xf_autotune": (_I, [_P, _I, _OP, _DP]),
    "xf_trace_next": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _I]),
    "xf_trace_fetch": (_I, [_P, ctypes.POINTER(ctypes.c_longlong), _I]),
    "xf_ensemble_create": (_I, [_PP, _I, _OP, ctypes.POINTER(_P)]),
    "xf_ensemble_destroy": (_I, [_P]),
    "xf_ensemble_upload": (_I, [_P, _I, _DP, _D, _D]),
    "xf_ensemble_init_pure": (_I, [_P, _I]),
    "xf_ensemble_download": (_I, [_P, _I, _DP, _DP, _DP]),
    "xf_ensemble_run_steps": (_I, [_P, _I, ctypes.POINTER(XfRunInfo)]),
    "xf_ensemble_run_graph": (_I, [_P, _I, _I, ctypes.POINTER(XfRunInfo)]),
    "xf_ensemble_autotune": (_I, [_P, _I, _OP, _DP]),
    "xf_ensemble_cells": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "xf_ensemble_set_rn": (_I, [_P, _I, _D]),
    "xf_ensemble_launch": (_I, [_P, ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I), ctypes.POINTER(_I)]),
    # include/xflat_io.h (host-side NetCDF output and resume)
    "xf_snapshot_create": (_I, [_PP, ctypes.c_char_p, ctypes.POINTER(_P)]),
    "xf_snapshot_append": (_I, [_P, ctypes.c_double, ctypes.c_double, _DP]),
    "xf_probability_create": (_I, [_PP, ctypes.c_char_p, ctypes.POINTER(_P)]),
    "xf_probability_append": (_I, [_P, ctypes.c_double, _DP, _DP]),
    "xf_nc_records": (_I, [_P, ctypes.POINTER(ctypes.c_int64)]),
    "xf_nc_sync": (_I, [_P]),
    "xf_nc_close": (_I, [_P]),
    "xf_snapshot_read_last": (_I, [ctypes.c_char_p, _PP, _DP, _DP, _DP]),
    "xf_output_name": (_I, [_PP, _I, _I, _I, ctypes.c_char_p, ctypes.c_size_t]),
}


# Handles still open: closed by an exit handler, before the interpreter
# finalises its modules and before the HIP runtime's own exit handlers run,
# so that no device memory is freed from a finaliser or a static destructor.
_LIVE = weakref.WeakSet()


def _close_all():
    for obj in list(_LIVE):
        try:
            obj.close()
        except Exception:
            pass


atexit.register(_close_all)


def load():
    """Loads libxflat.so (raises if it has not been built)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    if not os.path.exists(LIB_PATH):
        raise RuntimeError(f"{LIB_PATH} is missing: build it with __graft_entry__.build() "
                           "(python xflat_amd/build.py); there is no CPU fallback")
    L = ctypes.CDLL(LIB_PATH)
    for name, (res, args) in SIGNATURES.items():
        f = getattr(L, name)
        f.restype = res
        f.argtypes = args
    _LIB = L
    return L


def check(rc: int):
    if rc != XF_OK:
        msg = load().xf_last_error()
        raise XfError(rc, msg.decode() if msg else "")
    return rc


def dp(a: np.ndarray):
    assert a.dtype == np.float64 and a.flags.c_contiguous
    return a.ctypes.data_as(_DP)


def block_for(bpt: int, ne: int, cpt: int) -> int:
    """Threads of a workgroup whose tile is `bpt` beams of `ne` energy bins at
    `cpt` cells per lane (the tile's bpt * ne cells over the lanes, rounded up
    to whole waves): the block xf_set_launch needs for that tile."""
    return 64 * (((bpt * ne + cpt - 1) // cpt + 63) // 64)


def device_count() -> int:
    n = ctypes.c_int()
    load().xf_device_count(ctypes.byref(n))
    return n.value


def parse_config_text(text: str) -> XfParams:
    p = XfParams()
    check(load().xf_parse_config_text(text.encode(), ctypes.byref(p)))
    return p


def parse_config(path: str) -> XfParams:
    p = XfParams()
    check(load().xf_parse_config(path.encode(), ctypes.byref(p)))
    return p


def make_opts(precision=XF_FP64, species=0, renorm=-1, block=0, cpt=0, resident=-1, grid=0) -> XfOpts:
    return XfOpts(precision, species, renorm, block, cpt, resident, grid)


class Context:
    """One evolution-step context (xf_ctx): the fused stepper plus the Phy.h mirror."""

    def __init__(self, params: XfParams, opts: XfOpts | None = None, **kw):
        self.L = load()
        self.p = params
        self.opts = opts if opts is not None else make_opts(**kw)
        h = ctypes.c_void_p()
        check(self.L.xf_create(ctypes.byref(params), ctypes.byref(self.opts), ctypes.byref(h)))
        self.h = h
        _LIVE.add(self)
        self.nbeam = n_beams(params)
        self.ne = params.ebins
        self.shape = (self.nbeam, 4, 4, self.ne)

    def close(self):
        if self.h:
            self.L.xf_destroy(self.h)
            self.h = None
        _LIVE.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        if sys.is_finalizing():        # the exit handler has closed what it could
            return
        try:
            self.close()
        except Exception:
            pass

    # ---- fused stepper
    def upload(self, psi: np.ndarray, r: float, dr: float):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        assert a.shape == self.shape, (a.shape, self.shape)
        check(self.L.xf_state_upload(self.h, dp(a), r, dr))

    def init_pure(self):
        check(self.L.xf_state_init_pure(self.h))

    def download(self):
        a = np.zeros(self.shape)
        r = ctypes.c_double(); dr = ctypes.c_double()
        check(self.L.xf_state_download(self.h, dp(a), ctypes.byref(r), ctypes.byref(dr)))
        return a, r.value, dr.value

    def counters(self):
        t = ctypes.c_int64(); a = ctypes.c_int64()
        check(self.L.xf_state_counters(self.h, ctypes.byref(t), ctypes.byref(a)))
        return t.value, a.value

    def species(self) -> int:
        return self.L.xf_state_species(self.h)

    def set_rn(self, rn: float):
        """Moves the end radius Rn (runs stop exactly on it)."""
        check(self.L.xf_set_rn(self.h, rn))

    def run_to(self, rn: float, max_attempts: int = 1 << 20, chunk: int = 4000) -> int:
        """Runs until r == rn exactly (Rn moved to rn); returns the attempts."""
        self.set_rn(rn)
        done = 0
        while done < max_attempts:
            info, _ = self.run(min(chunk, max_attempts - done))
            done += info.steps
            if info.stop_reason == 1:
                return done
            if info.steps == 0:
                break
        raise RuntimeError(f"r = {rn} not reached in {done} attempts")

    def run(self, max_steps: int, log: bool = False, stop: XfStop | None = None):
        rec = np.zeros(max(max_steps, 1), dtype=STEP_DTYPE) if log else None
        info = XfRunInfo()
        check(self.L.xf_run_steps(self.h, max_steps, ctypes.byref(stop) if stop else None,
                                  rec.ctypes.data if log else None, ctypes.byref(info)))
        if log:
            rec = rec[:info.steps]
        return info, rec

    def run_noalloc(self, max_steps: int, log: bool = False, stop: XfStop | None = None):
        """run() through xf_run_steps_noalloc (no device allocation; the
        launch shape must have been chosen, e.g. by launch())."""
        rec = np.zeros(max(max_steps, 1), dtype=STEP_DTYPE) if log else None
        info = XfRunInfo()
        check(self.L.xf_run_steps_noalloc(self.h, max_steps, ctypes.byref(stop) if stop else None,
                                          rec.ctypes.data if log else None, ctypes.byref(info)))
        if log:
            rec = rec[:info.steps]
        return info, rec

    def run_graph(self, launches: int, max_steps: int):
        """`launches` launches of up to max_steps steps from one HIP graph."""
        info = XfRunInfo()
        check(self.L.xf_run_graph(self.h, launches, max_steps, ctypes.byref(info)))
        return info

    def launch(self):
        o = XfOpts(); t = ctypes.c_int(); b = ctypes.c_int()
        check(self.L.xf_get_launch(self.h, ctypes.byref(o), ctypes.byref(t), ctypes.byref(b)))
        return dict(precision=o.precision, species=o.species, block=o.block, cpt=o.cpt, mode=o.mode,
                    grid=o.grid, renorm=o.renorm, tiles=t.value, beams_per_tile=b.value)

    def set_launch(self, mode: int, cpt: int, block: int, bpt: int):
        check(self.L.xf_set_launch(self.h, mode, cpt, block, bpt))

    def launch_info(self):
        v = ctypes.c_int(); s = ctypes.c_int(); m = ctypes.c_int(); lds = ctypes.c_int()
        check(self.L.xf_launch_info(self.h, ctypes.byref(v), ctypes.byref(s), ctypes.byref(m), ctypes.byref(lds)))
        return dict(vgprs=v.value, scratch=s.value, max_blocks=m.value, lds=lds.value)

    def trace(self, steps: int):
        """Runs `steps` steps with per-phase timestamps; returns [steps, 16] in microseconds."""
        buf = np.zeros((steps, 16), dtype=np.int64)
        ptr = buf.ctypes.data_as(ctypes.POINTER(ctypes.c_longlong))
        check(self.L.xf_trace_next(self.h, ptr, steps))
        info, _ = self.run(steps)
        check(self.L.xf_trace_fetch(self.h, ptr, steps))
        return buf.astype(np.float64) * 0.01, info

    def autotune(self, steps: int = 20):
        o = XfOpts(); ms = ctypes.c_double()
        check(self.L.xf_autotune(self.h, steps, ctypes.byref(o), ctypes.byref(ms)))
        return self.launch(), ms.value

    # ---- Phy.h mirror
    def beams_upload(self, k, psi):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        check(self.L.xf_beams_upload(self.h, k, dp(a)))

    def beams_download(self, k):
        a = np.zeros(self.shape)
        check(self.L.xf_beams_download(self.h, k, dp(a)))
        return a

    def get_esum(self, k):
        a = np.zeros((self.nbeam, 4, 4))
        check(self.L.xf_get_esum(self.h, k, dp(a)))
        return a

    def get_hvv(self, k):
        a = np.zeros((self.nbeam, 4))
        check(self.L.xf_get_hvv(self.h, k, dp(a)))
        return a

    def hm(self, r):
        v = ctypes.c_double()
        check(self.L.xf_get_hm(self.h, r, ctypes.byref(v)))
        return v.value

    def calc_esum(self, k): check(self.L.xf_calc_esum(self.h, k))
    def calc_angle_bins(self, r, pt): check(self.L.xf_calc_angle_bins(self.h, r, pt))
    def calc_delta_ls(self, dr, cur, s, e): check(self.L.xf_calc_delta_ls(self.h, dr, cur, s, e))
    def calc_hvv(self, k, pt, hv): check(self.L.xf_calc_hvv(self.h, k, pt, hv))
    def evolve(self, i, pt, hv, hm, o): check(self.L.xf_evolve(self.h, i, pt, hv, hm, o))
    def evolve_hvv(self, i, pt, hv, hm, o): check(self.L.xf_evolve_hvv(self.h, i, pt, hv, hm, o))
    def evolve_avg(self, i, pt, hv, hm, o, a): check(self.L.xf_evolve_avg(self.h, i, pt, hv, hm, o, a))
    def evolve_avg_inplace(self, pt, hv, hm, io, a): check(self.L.xf_evolve_avg_inplace(self.h, pt, hv, hm, io, a))
    def avg_beam(self, i, o): check(self.L.xf_avg_beam(self.h, i, o))

    def calc_err_hvv(self, a, b) -> float:
        v = ctypes.c_double()
        check(self.L.xf_calc_err_hvv(self.h, a, b, ctypes.byref(v)))
        return v.value


class Ensemble:
    """Batched ensemble of independent instances advanced by one launch."""

    def __init__(self, params_list, opts: XfOpts | None = None, **kw):
        self.L = load()
        n = len(params_list)
        arr = (XfParams * n)(*params_list)
        self.params = list(params_list)
        self.opts = opts if opts is not None else make_opts(**kw)
        h = ctypes.c_void_p()
        check(self.L.xf_ensemble_create(arr, n, ctypes.byref(self.opts), ctypes.byref(h)))
        self.h = h
        self.n = n
        _LIVE.add(self)

    def close(self):
        if self.h:
            self.L.xf_ensemble_destroy(self.h)
            self.h = None
        _LIVE.discard(self)

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()

    def __del__(self):
        if sys.is_finalizing():
            return
        try:
            self.close()
        except Exception:
            pass

    def shape(self, i):
        p = self.params[i]
        nb = n_beams(p)
        return (nb, 4, 4, p.ebins)

    def upload(self, i, psi, r, dr):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        check(self.L.xf_ensemble_upload(self.h, i, dp(a), r, dr))

    def init_pure(self, i):
        check(self.L.xf_ensemble_init_pure(self.h, i))

    def download(self, i):
        a = np.zeros(self.shape(i))
        r = ctypes.c_double(); dr = ctypes.c_double()
        check(self.L.xf_ensemble_download(self.h, i, dp(a), ctypes.byref(r), ctypes.byref(dr)))
        return a, r.value, dr.value

    def run(self, max_steps):
        infos = (XfRunInfo * self.n)()
        check(self.L.xf_ensemble_run_steps(self.h, max_steps, infos))
        return list(infos)

    def run_graph(self, launches, max_steps):
        infos = (XfRunInfo * self.n)()
        check(self.L.xf_ensemble_run_graph(self.h, launches, max_steps, infos))
        return list(infos)

    def set_rn(self, i: int, rn: float):
        check(self.L.xf_ensemble_set_rn(self.h, i, rn))

    def run_to(self, radii, max_attempts: int = 1 << 20, chunk: int = 2000):
        """Every instance i runs until r == radii[i] exactly; returns the
        attempts of each."""
        for i, rn in enumerate(radii):
            self.set_rn(i, rn)
        done = [0] * self.n
        left = max_attempts
        while left > 0:
            infos = self.run(min(chunk, left))
            left -= max(x.steps for x in infos)
            for i, x in enumerate(infos):
                done[i] += x.steps
            if all(x.stop_reason == 1 for x in infos):
                return done
        raise RuntimeError("not every instance reached its radius")

    def cells(self) -> int:
        c = ctypes.c_int64()
        check(self.L.xf_ensemble_cells(self.h, ctypes.byref(c)))
        return c.value

    def autotune(self, steps=20):
        o = XfOpts()
        ms = ctypes.c_double()
        check(self.L.xf_ensemble_autotune(self.h, steps, ctypes.byref(o), ctypes.byref(ms)))
        return o, ms.value

    def launch(self):
        m = ctypes.c_int(); c = ctypes.c_int(); b = ctypes.c_int(); g = ctypes.c_int()
        check(self.L.xf_ensemble_launch(self.h, ctypes.byref(m), ctypes.byref(c), ctypes.byref(b), ctypes.byref(g)))
        return dict(mode=m.value, cpt=c.value, block=b.value, grid=g.value)


class NcFile:
    """Snapshot or Probability output file (include/xflat_io.h)."""

    def __init__(self, p: XfParams, path: str, probability: bool = False):
        self.L = load()
        self.p = p
        self.h = ctypes.c_void_p()
        fn = self.L.xf_probability_create if probability else self.L.xf_snapshot_create
        check(fn(ctypes.byref(p), path.encode(), ctypes.byref(self.h)))

    def append_snapshot(self, r: float, dr: float, psi: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        check(self.L.xf_snapshot_append(self.h, r, dr, dp(a)))

    def append_probability(self, r: float, psi: np.ndarray, fv: np.ndarray):
        a = np.ascontiguousarray(psi, dtype=np.float64)
        w = np.ascontiguousarray(fv, dtype=np.float64)
        check(self.L.xf_probability_append(self.h, r, dp(a), dp(w)))

    def records(self) -> int:
        n = ctypes.c_int64()
        check(self.L.xf_nc_records(self.h, ctypes.byref(n)))
        return n.value

    def sync(self):
        check(self.L.xf_nc_sync(self.h))

    def close(self):
        if self.h:
            check(self.L.xf_nc_close(self.h))
            self.h = ctypes.c_void_p()

    def __enter__(self):
        return self

    def __exit__(self, *a):
        self.close()


def snapshot_read_last(path: str, p: XfParams):
    """Last record (psi, r, dr) of a Snapshot file (the resume read)."""
    nb = n_beams(p)
    psi = np.zeros((nb, 4, 4, p.ebins))
    r = ctypes.c_double()
    dr = ctypes.c_double()
    check(load().xf_snapshot_read_last(path.encode(), ctypes.byref(p), dp(psi), ctypes.byref(r), ctypes.byref(dr)))
    return psi, r.value, dr.value


def output_name(p: XfParams, probability: bool, ctr: int = 0, rank: int = 0) -> str:
    buf = ctypes.create_string_buffer(1024)
    check(load().xf_output_name(ctypes.byref(p), int(probability), ctr, rank, buf, 1024))
    return buf.value.decode()

"""Throughput of the multi-angle evolution step on one MI355X.

Default workload (the headline): the bulb model (MA) at production size,
1000 angle bins x 400 energy bins (400k beams, each carrying nu and anti-nu),
from the neutrinosphere R_nu = 10 km with a seeded 1e-6 off-diagonal
perturbation, tolerance eps0 = 1e-9, inverted hierarchy.  A "step" is one
attempted adaptive radial step (S1..S11 + error + accept/reject);
beam-steps/s = beams x attempted steps / wall time of the timed call, with the
state resident on the GPU before the timed region.  The launch shape is
chosen by the project's autotuner before warm-up.

The headline `value` is the FP32-state stepper: FP32 wave functions and
propagator increments, FP64 angular moments, error norm, accept/reject and
unitarity renormalisation.  Reported beside it from the same run: the
FP64-state stepper ("mixed": FP64 state, FP32 increments), the plain FP64
stepper, the max |dP_ee| / |dPbar_ee| of each against the FP64 run (after
the timed steps, and at 20 output radii of an untimed run of the same
length), the same figures in the normal hierarchy, and the reference's own
CPU path (oracle/_ref build, OpenMP over this host's cores) on a bounded
sample of the same workload.

Other workloads (one JSON line each, not the headline):
  --workload maa256x64x200   MAA 256 polar x 64 azimuthal x 200 energies
                              (3.3M beams), azimuthal-symmetry-breaking seed
  --workload ensemble96       96 independent bulb instances of seeded sizes
                              in one launch (14.2M beams)

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|mixed|fp64]
"""
import argparse
import json
import os
import re
import subprocess
import sys
import tempfile
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

import numpy as np  # noqa: E402

HBM_PEAK_GBS = 8000.0
METRIC = "beam-steps/sec (angle x energy x dr) at stated P_ee error vs FP64; % HBM roofline"
R0, DR0 = 10.0, 1e-3
PROFILE_HBM = os.path.join(ROOT, "profiles", "k_step_fp32_ma_1000x400_hbm.md")
PROFILED = ("fp32", 1, 4, 4)         # precision, mode, cells per lane, beams per tile of that profile

WORKLOADS = {
    "ma1000x400": dict(
        desc="bulb MA 1000 angles x 400 energies, R0 = 10 km, eps0 = 1e-9, inverted hierarchy",
        params=dict(Abins=1000, Ebins=400, R0=R0, Rn=1000, eps0="1.E-9", dm2="-2.4E-3"),
        steps=4000, warmup=300),
    "maa256x64x200": dict(
        desc="MAA 256 polar x 64 azimuthal x 200 energies, R0 = 10 km, Lve = 1.25e51 erg/s, eps0 = 1e-8",
        params=dict(Abins=256, Pbins=64, Ebins=200, R0=R0, Rn=400, Lve="1.25E+51"),
        steps=200, warmup=20),
    "ensemble96": dict(
        desc="96 bulb MA instances, angle bins 64..1500, energy bins 32..400, luminosities 0.5-3e51 erg/s, "
             "theta_v 0.01-0.6, either hierarchy, n_e(R_nu) 1e31-1e33 cm^-3, seeded; R0 = 10 km, eps0 = 1e-8",
        steps=200, warmup=20),
}


def _cfg(**over):
    from xflat_amd.config import bulb_params
    return bulb_params(**over)


def _pee(s):
    return s[:, 0, 0] ** 2 + s[:, 0, 1] ** 2, s[:, 1, 0] ** 2 + s[:, 1, 1] ** 2


def _errors(psi, ref):
    pe, pb = _pee(psi)
    re_, rb = _pee(ref)
    return {"max_abs_dPee": float(np.abs(pe - re_).max()), "max_abs_dPbar_ee": float(np.abs(pb - rb).max()),
            "unitarity_drift": float(np.abs(np.sum(psi[:, 0] ** 2, axis=1) - 1).max())}


def _prec(name):
    from xflat_amd import capi
    return {"mixed": capi.XF_MIXED, "fp64": capi.XF_FP64, "fp32": capi.XF_FP32}[name]


def _state_bytes(prec_name):
    """Algorithmic bytes per beam-step: nu and anti-nu 2-spinors (8 reals) read
    and written once, in the state's precision."""
    return 64 if prec_name == "fp32" else 128


def time_ctx(params, prec_name, psi0, steps, warmup, tune_steps, launch=None):
    from xflat_amd import capi
    ctx = capi.Context(params, precision=_prec(prec_name))
    ctx.upload(psi0, params.R0, DR0)
    if launch:
        mode, cpt, bpt = launch
        ctx.set_launch(mode, cpt, 64 * (((bpt * params.ebins + cpt - 1) // cpt + 63) // 64), bpt)
    elif tune_steps > 0:
        ctx.autotune(tune_steps)
    ctx.upload(psi0, params.R0, DR0)
    if warmup > 0:
        ctx.run(warmup)
    t0 = time.perf_counter()
    info, _ = ctx.run(steps)          # synchronous: returns after the stream drained
    t1 = time.perf_counter()
    li = dict(ctx.launch(), **ctx.launch_info())
    psi, r, dr = ctx.download()
    ctx.close()
    return dict(wall_s=t1 - t0, kernel_ms=info.kernel_ms, steps=info.steps, accepted=info.accepted,
                r=r, dr=dr, psi=psi, launch=li)


def time_short_launches(params, prec_name, psi0, launch, launches=100, per=10):
    """A radial loop cut into short launches (a host that stops every few
    steps to test its stop rules or write output): `launches` eager
    xf_run_steps calls against one xf_run_graph replay of the same launches,
    each from the same state, after one untimed pass of each."""
    from xflat_amd import capi
    ctx = capi.Context(params, precision=_prec(prec_name))
    ctx.upload(psi0, params.R0, DR0)
    if launch:
        mode, cpt, bpt = launch
        ctx.set_launch(mode, cpt, 64 * (((bpt * params.ebins + cpt - 1) // cpt + 63) // 64), bpt)
    cells = params.abins * params.ebins
    out = {"launches": launches, "steps_per_launch": per}
    for name in ("eager", "graph"):
        for timed in (False, True):
            ctx.upload(psi0, params.R0, DR0)
            t0 = time.perf_counter()
            if name == "eager":
                steps = sum(ctx.run(per)[0].steps for _ in range(launches))
            else:
                steps = ctx.run_graph(launches, per).steps
            t1 = time.perf_counter()
        out[f"{name}_value"] = cells * steps / (t1 - t0)
        out[f"{name}_ms_per_step"] = 1e3 * (t1 - t0) / max(1, steps)
        out[f"{name}_radius_km"] = ctx.download()[1]
    ctx.close()
    return out


def accuracy_points(params, prec_name, psi0, total, npts=20):
    """Untimed accuracy leg: the stepper and the FP64 stepper advance from the
    same state in npts equal chunks of attempted steps; P_ee, anti-P_ee and the
    unitarity drift are compared at the end of every chunk (the output radii)."""
    from xflat_amd import capi
    a = capi.Context(params, precision=_prec(prec_name))
    b = capi.Context(params, precision=capi.XF_FP64)
    try:
        for c in (a, b):
            c.upload(psi0, params.R0, DR0)
        chunk = max(1, total // npts)
        pts = []
        done = 0
        for k in range(npts):
            n = chunk if k < npts - 1 else max(1, total - done)
            a.run(n)
            b.run(n)
            done += n
            pa, ra, _ = a.download()
            pb, rb, _ = b.download()
            pts.append(dict(_errors(pa, pb), attempts=done, radius_km=ra, radius_km_fp64=rb))
    finally:
        a.close()
        b.close()
    return {"points": npts, "note": "the stepper vs FP64 from the same state after every 1/%d of the attempted steps" % npts,
            "max_abs_dPee": max(x["max_abs_dPee"] for x in pts),
            "max_abs_dPbar_ee": max(x["max_abs_dPbar_ee"] for x in pts),
            "unitarity_drift": max(x["unitarity_drift"] for x in pts),
            "per_point": pts}


def time_ensemble(params, prec_name, inits, steps, warmup, tune_steps):
    from xflat_amd import capi
    ens = capi.Ensemble(params, precision=_prec(prec_name))
    for i, p in enumerate(params):
        ens.upload(i, inits(i), p.R0, DR0)
    if tune_steps > 0:
        ens.autotune(tune_steps)
        for i, p in enumerate(params):
            ens.upload(i, inits(i), p.R0, DR0)
    if warmup > 0:
        ens.run(warmup)
    t0 = time.perf_counter()
    infos = ens.run(steps)
    t1 = time.perf_counter()
    out = dict(wall_s=t1 - t0, kernel_ms=infos[0].kernel_ms, steps=[x.steps for x in infos],
               accepted=[x.accepted for x in infos], r=[x.r for x in infos], launch=ens.launch(),
               cells=ens.cells())
    return ens, out


def _host_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "?"


def _ref_exe(maa):
    isa = "v3"
    try:
        with open("/proc/cpuinfo") as f:
            if " avx512f" in f.read():
                isa = "v4"
    except OSError:
        pass
    return os.path.join(ROOT, "oracle", "_ref", f"xflat_ref_{'maa' if maa else 'ma'}_{isa}")


def _ncores():
    n = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else os.cpu_count()
    return max(1, min(16, n or 1))


def ref_run(over, psi0, steps, ncores):
    """The reference's evolution path (oracle/_ref) for `steps` attempts on a
    config; the initial state is injected through its resume path when given
    (MA only).  Returns (attempts, loop seconds) or None."""
    from xflat_amd.config import config_text
    maa = int(over.get("Pbins", 1)) > 1
    exe = _ref_exe(maa)
    if not os.path.exists(exe):
        return None
    with tempfile.TemporaryDirectory() as td:
        cfg = os.path.join(td, "config.txt")
        with open(cfg, "w") as f:
            f.write(config_text(**dict(over, Ts=steps, dumpMode=0)))
        env = dict(os.environ, OMP_NUM_THREADS=str(ncores), OMP_PROC_BIND="close")
        args = [exe, cfg]
        if psi0 is not None:
            init = os.path.join(td, "init.bin")
            with open(init, "wb") as f:
                np.array([psi0.shape[0] * 4, 4, psi0.shape[-1]], dtype=np.int64).tofile(f)
                np.array([float(over.get("R0", R0)), DR0]).tofile(f)
                psi0.astype(np.float64).tofile(f)
            env["XFLAT_REF_INIT"] = init
            args.append(init)
        r = subprocess.run(args, env=env, capture_output=True, text=True, cwd=td, timeout=900)
        m = re.search(r"total time:([0-9.eE+-]+)", r.stdout)
        tot = re.search(r"Total iterations:\s*(\d+)", r.stdout)
        if r.returncode == 0 and m and tot:
            return int(tot.group(1)), float(m.group(1))
    return None


def cpu_baseline(workload, psi0=None, ens_cfgs=None):
    """Reference evolution path on this host's cores, bounded sample (about
    10-30 s); the project's CPU restatement when the reference build is absent."""
    ncores = _ncores()
    model = _host_model()
    w = WORKLOADS[workload]
    if workload == "ensemble96":
        # the first three instances, each for a bounded number of attempts
        cells = steps = secs = 0.0
        for d in ens_cfgs[:3]:
            over = dict(d, R0=R0, Rn=1000)
            n_cells = d["Abins"] * d["Ebins"]
            n = max(8, int(1.2e9 / n_cells / 8))
            from xflat_amd.seed import seeded_psi
            res = ref_run(over, seeded_psi(d["Abins"], d["Ebins"], seed=7, amp=1e-6), n, ncores)
            if res is None:
                return None
            cells += n_cells * res[0]
            secs += res[1]
        return dict(value=cells / secs, unit="beam-steps/s", cores=ncores, kind="reference",
                    sample=f"the first 3 ensemble instances run one after another on the reference's own evolution "
                           f"path (oracle/_ref, OpenMP {ncores} threads), {secs:.2f} s loop time, host '{model}'")
    n_cells = np.prod([w["params"][k] for k in ("Abins", "Ebins")]) * int(w["params"].get("Pbins", 1))
    steps = max(8, int(3.2e8 * ncores / 16 / n_cells))
    maa = workload.startswith("maa")
    res = ref_run(w["params"], None if maa else psi0, steps, ncores)
    if res is not None:
        n, secs = res
        return dict(value=n_cells * n / secs, unit="beam-steps/s", cores=ncores, kind="reference",
                    sample=f"{w['desc']}, FP64, {n} attempted steps on the reference's own evolution path "
                           f"(oracle/_ref/{os.path.basename(_ref_exe(maa))}, OpenMP {ncores} threads, "
                           f"OMP_PROC_BIND=close, {secs:.2f} s loop time"
                           f"{', pure flavour initial states: the reference cannot load an MAA state' if maa else ''}"
                           f", host '{model}')")
    from oracle.oracle import Oracle, set_threads
    set_threads(ncores)
    o = Oracle(_cfg(**w["params"]))
    if maa:
        o.init_state_pure()
    else:
        o.set_state(psi0, R0, DR0)
    n, _, secs = o.run(steps, log=False)
    return dict(value=n_cells * n / secs, unit="beam-steps/s", cores=ncores, kind="port",
                sample=f"{w['desc']}, FP64, {n} attempted steps on the project's CPU restatement "
                       f"(OpenMP {ncores} threads, {secs:.2f} s, host '{model}')")


def profiled_traffic_per_step(prec_name, launch):
    """HBM bytes per attempted step measured by rocprofv3 FETCH_SIZE/WRITE_SIZE
    (scaled by a copy kernel) for the launch shape of the committed profile;
    None for any other precision or shape."""
    if (prec_name, launch["mode"], launch["cpt"], launch["beams_per_tile"]) != PROFILED:
        return None
    try:
        with open(PROFILE_HBM) as f:
            m = re.search(r"^traffic_bytes_per_step:\s*([0-9.eE+]+)", f.read(), re.M)
        return float(m.group(1)) if m else None
    except OSError:
        return None


def calibrate_copy(nbytes=1 << 30):
    """A plain device-to-device copy of known size (the HIP runtime's own copy
    kernel), for scaling rocprof FETCH_SIZE / WRITE_SIZE.  Calls the HIP runtime
    directly so that a counter-collection run loads no other GPU library."""
    import ctypes
    try:
        hip = ctypes.CDLL("libamdhip64.so")
    except OSError:
        hip = ctypes.CDLL("/opt/rocm/lib/libamdhip64.so")
    a, b = ctypes.c_void_p(), ctypes.c_void_p()
    for p in (a, b):
        if hip.hipMalloc(ctypes.byref(p), ctypes.c_size_t(nbytes)) != 0:
            raise RuntimeError("hipMalloc failed")
    try:
        if hip.hipMemset(a, 1, ctypes.c_size_t(nbytes)) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemset failed")
        if hip.hipMemcpy(b, a, ctypes.c_size_t(nbytes), 3) != 0 or hip.hipDeviceSynchronize() != 0:
            raise RuntimeError("hipMemcpy failed")
    finally:
        hip.hipFree(a)
        hip.hipFree(b)
    return nbytes


def _line(workload, prec_name, value, steps, warmup, ms_per_step, cells, extra):
    return dict({
        "metric": METRIC, "value": value, "unit": "beam-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "fp32 (propagator) + fp64 (state, moments, error, renorm)", "fp64": "fp64",
                  "fp32": "fp32 (state, propagator) + fp64 (moments, error, renorm)"}[prec_name],
    }, **extra)


def _roofline(prec_name, cells, steps, kernel_ms, launch, traffic_per_step=None):
    b = _state_bytes(prec_name)
    achieved = b * cells * steps / (kernel_ms / 1e3) / 1e9
    tr = traffic_per_step * steps if traffic_per_step else None
    return {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS, "traffic": tr,
            "traffic_source": (os.path.relpath(PROFILE_HBM, ROOT) if tr else None),
            "note": (f"algorithmic {b} B per beam-step (nu + anti-nu 2-spinors read and written once) over the "
                     f"stepper kernel's HIP-event time (mode {launch['mode']})")}


def bench_ma(a, prec_name):
    from xflat_amd.seed import seeded_psi
    w = WORKLOADS["ma1000x400"]
    p = _cfg(**w["params"])
    cells = p.abins * p.ebins
    psi0 = seeded_psi(p.abins, p.ebins, seed=2024, amp=1e-6)
    launch = tuple(int(x) for x in a.launch.split(",")) if a.launch else None
    print(f"[bench] {prec_name} MA 1000x400: autotune, warmup {a.warmup}, timing {a.steps} steps",
          file=sys.stderr, flush=True)
    res = time_ctx(p, prec_name, psi0, a.steps, a.warmup, a.tune_steps, launch)
    value = cells * res["steps"] / res["wall_s"]
    print(f"[bench] {prec_name}: {value:.3e} beam-steps/s, launch {res['launch']}", file=sys.stderr, flush=True)
    out = _line("ma1000x400", prec_name, value, res["steps"], a.warmup, 1e3 * res["wall_s"] / max(1, res["steps"]),
                cells, {
                    "data": "synthetic: pure flavour states with a seeded 1e-6 off-diagonal perturbation",
                    "config": {"workload": w["desc"], "beams": cells, "precision": prec_name,
                               "launch": res["launch"]},
                    "accepted_steps": res["accepted"], "radius_km": res["r"],
                    "kernel_ms_per_step": res["kernel_ms"] / max(1, res["steps"]),
                    "roofline": _roofline(prec_name, cells, res["steps"], res["kernel_ms"], res["launch"],
                                          profiled_traffic_per_step(prec_name, res["launch"])),
                })
    if not a.no_fp64 and prec_name != "fp64":
        ref = time_ctx(p, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
        out["fp64"] = {"value": cells * ref["steps"] / ref["wall_s"],
                       "ms_per_step": 1e3 * ref["wall_s"] / max(1, ref["steps"]),
                       "accepted_steps": ref["accepted"], "radius_km": ref["r"], "launch": ref["launch"]}
        out["radius_advance_vs_fp64"] = (res["r"] - R0) / max(ref["r"] - R0, 1e-300)
        out["pee_error_vs_fp64"] = dict(_errors(res["psi"], ref["psi"]), note=(
            "after the same number of attempted steps from the same state"), radius_km=[res["r"], ref["r"]],
            same_accepted_count=bool(ref["accepted"] == res["accepted"]))
        print(f"[bench] fp64: {out['fp64']['value']:.3e} beam-steps/s", file=sys.stderr, flush=True)
        out["pee_error_vs_fp64_at_20_radii"] = accuracy_points(p, prec_name, psi0, a.warmup + a.steps)
        print(f"[bench] max |dP_ee| vs fp64 over 20 radii: {out['pee_error_vs_fp64_at_20_radii']['max_abs_dPee']:.2e}",
              file=sys.stderr, flush=True)
        if not a.no_extra:
            # the same workload in the normal hierarchy (configs run both)
            pn = _cfg(**dict(w["params"], dm2="2.4E-3"))
            rn = time_ctx(pn, prec_name, psi0, a.steps, a.warmup, a.tune_steps, launch)
            fn = time_ctx(pn, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
            out["normal_hierarchy"] = dict(
                {"value": cells * rn["steps"] / rn["wall_s"], "ms_per_step": 1e3 * rn["wall_s"] / max(1, rn["steps"]),
                 "accepted_steps": rn["accepted"], "radius_km": rn["r"],
                 "fp64_value": cells * fn["steps"] / fn["wall_s"], "fp64_accepted_steps": fn["accepted"],
                 "fp64_radius_km": fn["r"]}, **_errors(rn["psi"], fn["psi"]))
            print(f"[bench] normal hierarchy: {prec_name} {out['normal_hierarchy']['value']:.3e}, fp64 "
                  f"{out['normal_hierarchy']['fp64_value']:.3e} beam-steps/s", file=sys.stderr, flush=True)
        if not a.no_extra:
            li = res["launch"]                        # the headline's (tuned) shape
            out["short_launches"] = time_short_launches(p, prec_name, psi0,
                                                        (li["mode"], li["cpt"], li["beams_per_tile"]))
            s = out["short_launches"]
            print(f"[bench] {s['launches']} launches x {s['steps_per_launch']} steps: eager {s['eager_value']:.3e}, "
                  f"graph {s['graph_value']:.3e} beam-steps/s", file=sys.stderr, flush=True)
        if prec_name == "fp32" and not a.no_extra:
            mx = time_ctx(p, "mixed", psi0, a.steps, a.warmup, a.tune_steps)
            out["fp64_state_mixed"] = dict(
                {"value": cells * mx["steps"] / mx["wall_s"], "ms_per_step": 1e3 * mx["wall_s"] / max(1, mx["steps"]),
                 "accepted_steps": mx["accepted"], "radius_km": mx["r"], "launch": mx["launch"]},
                **_errors(mx["psi"], ref["psi"]))
            print(f"[bench] mixed (fp64 state): {out['fp64_state_mixed']['value']:.3e} beam-steps/s",
                  file=sys.stderr, flush=True)
    if not a.no_cpu:
        print("[bench] cpu baseline ...", file=sys.stderr, flush=True)
        out["cpu_baseline"] = cpu_baseline("ma1000x400", psi0=psi0)
    return out


def bench_maa(a, prec_name):
    from xflat_amd.seed import azimuthal_psi
    w = WORKLOADS["maa256x64x200"]
    p = _cfg(**w["params"])
    cells = p.abins * p.pbins * p.ebins
    psi0 = azimuthal_psi(p.abins, p.pbins, p.ebins, seed=31, amp=1e-7)
    print(f"[bench] {prec_name} MAA 256x64x200: autotune, warmup {a.warmup}, timing {a.steps} steps",
          file=sys.stderr, flush=True)
    res = time_ctx(p, prec_name, psi0, a.steps, a.warmup, a.tune_steps)
    value = cells * res["steps"] / res["wall_s"]
    out = _line("maa256x64x200", prec_name, value, res["steps"], a.warmup, 1e3 * res["wall_s"] / max(1, res["steps"]),
                cells, {
                    "data": "synthetic: pure flavour states with a seeded azimuthal perturbation "
                            "1e-7 sum_{m=1..4} exp(i(m phi + phi_m))",
                    "config": {"workload": w["desc"], "beams": cells, "precision": prec_name, "launch": res["launch"]},
                    "accepted_steps": res["accepted"], "radius_km": res["r"],
                    "kernel_ms_per_step": res["kernel_ms"] / max(1, res["steps"]),
                    "roofline": _roofline(prec_name, cells, res["steps"], res["kernel_ms"], res["launch"]),
                })
    print(f"[bench] {prec_name}: {value:.3e} beam-steps/s, launch {res['launch']}", file=sys.stderr, flush=True)
    if not a.no_fp64 and prec_name != "fp64":
        ref = time_ctx(p, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
        out["fp64"] = {"value": cells * ref["steps"] / ref["wall_s"], "accepted_steps": ref["accepted"],
                       "radius_km": ref["r"], "launch": ref["launch"]}
        out["pee_error_vs_fp64"] = dict(_errors(res["psi"], ref["psi"]), radius_km=[res["r"], ref["r"]])
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline("maa256x64x200")
    return out


def bench_ensemble(a, prec_name):
    from xflat_amd.seed import ensemble_configs, seeded_psi
    w = WORKLOADS["ensemble96"]
    cfgs = ensemble_configs(96)
    params = [_cfg(**dict(d, R0=R0, Rn=1000)) for d in cfgs]
    inits = lambda i: seeded_psi(params[i].abins, params[i].ebins, seed=5000 + i, amp=1e-6)  # noqa: E731
    print(f"[bench] {prec_name} ensemble of 96: autotune, warmup {a.warmup}, timing {a.steps} steps",
          file=sys.stderr, flush=True)
    ens, res = time_ensemble(params, prec_name, inits, a.steps, a.warmup, a.tune_steps)
    cells = res["cells"]
    done = sum(p.abins * p.ebins * n for p, n in zip(params, res["steps"]))
    value = done / res["wall_s"]
    out = _line("ensemble96", prec_name, value, int(max(res["steps"])), a.warmup,
                1e3 * res["wall_s"] / max(1, max(res["steps"])), cells, {
                    "data": "synthetic: seeded instance parameters; pure flavour states with a seeded 1e-6 "
                            "off-diagonal perturbation",
                    "config": {"workload": w["desc"], "instances": 96, "beams": cells, "precision": prec_name,
                               "launch": res["launch"]},
                    "accepted_steps_total": int(sum(res["accepted"])),
                    "roofline": _roofline(prec_name, cells, max(res["steps"]), res["kernel_ms"],
                                          dict(res["launch"])),
                })
    print(f"[bench] {prec_name}: {value:.3e} beam-steps/s, launch {res['launch']}", file=sys.stderr, flush=True)
    if not a.no_fp64 and prec_name != "fp64":
        ens64, ref = time_ensemble(params, "fp64", inits, a.steps, a.warmup, a.tune_steps)
        worst = []
        for i in range(96):
            psi, _, _ = ens.download(i)
            psi64, _, _ = ens64.download(i)
            worst.append(_errors(psi, psi64))
        ens64.close()
        out["fp64"] = {"value": sum(p.abins * p.ebins * n for p, n in zip(params, ref["steps"])) / ref["wall_s"],
                       "accepted_steps_total": int(sum(ref["accepted"]))}
        out["pee_error_vs_fp64"] = {
            "per_instance_max_abs_dPee": [x["max_abs_dPee"] for x in worst],
            "max_abs_dPee": max(x["max_abs_dPee"] for x in worst),
            "max_abs_dPbar_ee": max(x["max_abs_dPbar_ee"] for x in worst),
            "unitarity_drift": max(x["unitarity_drift"] for x in worst),
            "instances_with_same_accepted_count": int(sum(int(x == y) for x, y in zip(res["accepted"],
                                                                                      ref["accepted"])))}
    ens.close()
    if not a.no_cpu:
        out["cpu_baseline"] = cpu_baseline("ensemble96", ens_cfgs=cfgs)
    return out


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default="ma1000x400", choices=sorted(WORKLOADS))
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--tune-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extra", action="store_true", help="skip the FP64-state mixed leg")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--launch", default=None, help="mode,cpt,bpt: fixed launch shape instead of autotuning")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (several GPUs would shard the "
              "angle bins, one all-reduce of the moments per exchange; that path is not built).", file=sys.stderr)
        sys.exit(2)
    w = WORKLOADS[a.workload]
    a.steps = a.steps if a.steps is not None else w["steps"]
    a.warmup = a.warmup if a.warmup is not None else w["warmup"]
    fn = {"ma1000x400": bench_ma, "maa256x64x200": bench_maa, "ensemble96": bench_ensemble}[a.workload]
    out = fn(a, a.precision)
    if a.calibrate_copy:
        out["copy_calibration_bytes"] = calibrate_copy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

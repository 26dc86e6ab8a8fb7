"""Throughput of the multi-angle evolution step on one MI355X.

Headline workload: the bulb model (MA) at production size, 1000 angle bins x
400 energy bins (400k beams, each carrying nu and anti-nu), from the
neutrinosphere R_nu = 10 km with a seeded 1e-6 off-diagonal perturbation,
tolerance eps0 = 1e-9, inverted hierarchy.  A "step" is one attempted
adaptive radial step (S1..S11 + error + accept/reject); beam-steps/s =
beams x attempted steps / wall time of the timed call, with the state
resident on the GPU before the timed region.  The launch shape is chosen by
the project's autotuner before warm-up.

The headline `value` is the FP32-state stepper: FP32 wave functions and
propagator increments, FP64 angular moments, error norm, accept/reject and
unitarity renormalisation.  Reported beside it from the same run:

- the plain FP64 stepper and the FP64-state mixed stepper, timed alike;
- the FP32 state against an FP64 run stopped on exactly the radius the
  FP32 run reached (max |dP_ee|, |dPbar_ee|, unitarity drift);
- the same at 20 output radii of a stated window where the flavour content
  moves (300 -> 340 km, inverted hierarchy: the antineutrinos convert),
  both steppers stopping exactly on each radius;
- the normal hierarchy; the radial loop cut into short launches, eager
  against one HIP-graph replay;
- the other workloads of the same path with the same --steps/--warmup
  (MAA 256 x 64 x 200 in a luminosity sweep, and the 96-instance batched
  ensemble), each with its FP64 figure and per-instance errors;
- the reference's own CPU path (oracle/_ref, OpenMP on this host's CPU
  share) on a bounded sample of each workload;
- the HBM roofline of the timed kernel (algorithmic bytes), its measured
  HBM traffic and its VALU issue rate from the committed rocprof profiles.

python bench.py [--gpus 1] [--steps K] [--warmup W] [--precision fp32|mixed|fp64]
                [--workload ma1000x400|maa256x64x200|ensemble96]
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
# VALU issue: 256 CUs x 4 SIMDs, one wave64 instruction per SIMD every 4
# cycles at 2.4 GHz (non-packed); FP32 and FP64 vector peaks of the part.
VALU_PEAK_WAVE_INSTR = 1024 * 2.4e9 / 4
FP32_PEAK_TFLOPS = 157.3
FP64_PEAK_TFLOPS = 78.6
METRIC = "beam-steps/sec (angle x energy x dr) at stated P_ee error vs FP64; % HBM roofline"
R0, DR0 = 10.0, 1e-3
PROF = os.path.join(ROOT, "profiles")
# committed rocprof summaries (HBM traffic file, SQ counter file) of the
# launch shapes they were taken at: (workload, precision) -> list of ((mode,
# cells per lane, beams per tile or None, block or None), HBM file, SQ file),
# one per shape the autotuner picks on some boxes
PROFILES = {
    ("ma1000x400", "fp32"): [((1, 4, 4, None), "k_step_fp32_ma_1000x400_hbm.md", "k_step_fp32_ma_1000x400_valu.md"),
                             ((0, 4, 4, None), "k_step_fp32_ma_1000x400_mode0.md", "k_step_fp32_ma_1000x400_mode0.md")],
    ("ma1000x400", "fp64"): [((4, 1, 1, None), "k_step_fp64_ma_1000x400.md", "k_step_fp64_ma_1000x400.md")],
    ("ma1000x400", "mixed"): [((1, 4, 4, None), "k_step_mixed_ma_1000x400.md", "k_step_mixed_ma_1000x400.md")],
    ("maa256x64x200", "fp32"): [((4, 2, 5, None), "k_step_fp32_maa_256x64x200.md", "k_step_fp32_maa_256x64x200.md")],
    ("ensemble96", "fp32"): [((4, 4, None, 512), "k_step_fp32_ensemble96.md", "k_step_fp32_ensemble96.md")],
}
ACC_WINDOW = dict(R0=300.0, R1=340.0, points=20)

WORKLOADS = {
    "ma1000x400": dict(
        desc="bulb MA 1000 angles x 400 energies, R0 = 10 km, eps0 = 1e-9, inverted hierarchy",
        params=dict(Abins=1000, Ebins=400, R0=R0, Rn=1000, eps0="1.E-9", dm2="-2.4E-3"),
        steps=4000, warmup=300),
    "maa256x64x200": dict(
        desc="MAA 256 polar x 64 azimuthal x 200 energies, R0 = 10 km, eps0 = 1e-8, Lve = 1.0/1.25/1.5e51 erg/s",
        params=dict(Abins=256, Pbins=64, Ebins=200, R0=R0, Rn=400),
        lve=("1.0E+51", "1.25E+51", "1.5E+51"),
        steps=200, warmup=20),
    "ensemble96": dict(
        desc="96 bulb MA instances, angle bins 64..1500, energy bins 32..400, luminosities 0.5-3e51 erg/s, "
             "theta_v 0.01-0.6, either hierarchy, n_e(R_nu) 1e31-1e33 cm^-3, seeded; R0 = 10 km, eps0 = 1e-8",
        steps=200, warmup=20),
}


def log(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


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


def _set_shape(ctx, params, launch):
    from xflat_amd import capi
    mode, cpt, bpt = launch
    ctx.set_launch(mode, cpt, capi.block_for(bpt, params.ebins, cpt), bpt)


def time_ctx(params, prec_name, psi0, steps, warmup, tune_steps, launch=None, r0=None):
    """Tunes (or sets) the launch shape on the state, re-uploads it, runs the
    warm-up and times one xf_run_steps call of `steps` attempts."""
    from xflat_amd import capi
    r0 = params.R0 if r0 is None else r0
    ctx = capi.Context(params, precision=_prec(prec_name))
    try:
        ctx.upload(psi0, r0, DR0)
        if launch:
            _set_shape(ctx, params, launch)
        elif tune_steps > 0:
            ctx.autotune(tune_steps)
        ctx.upload(psi0, r0, DR0)
        if warmup > 0:
            ctx.run(warmup)
        t0 = time.perf_counter()
        info, _ = ctx.run(steps)          # synchronous: returns after the stream drained
        t1 = time.perf_counter()
        li = dict(ctx.launch(), **ctx.launch_info())
        psi, r, dr = ctx.download()
    finally:
        ctx.close()
    return dict(wall_s=t1 - t0, kernel_ms=info.kernel_ms, steps=info.steps, accepted=info.accepted,
                r=r, dr=dr, psi=psi, launch=li)


def fp64_at_radius(params, psi0, r_end, r0=None):
    """The FP64 stepper from the same state, stopped exactly on r_end."""
    from xflat_amd import capi
    with capi.Context(params, precision=capi.XF_FP64) as c:
        c.upload(psi0, params.R0 if r0 is None else r0, DR0)
        n = c.run_to(r_end)
        psi, r, _ = c.download()
    assert r == r_end
    return psi, n


def equal_radius_errors(params, psi0, res):
    """Errors of a timed run's final state against FP64 stopped on the same radius."""
    ref, n = fp64_at_radius(params, psi0, res["r"])
    return dict(_errors(res["psi"], ref), radius_km=res["r"], fp64_attempts=n, attempts=res["steps"],
                note="against the FP64 stepper from the same state stopped exactly on the same radius")


def accuracy_window(prec_name, w=ACC_WINDOW):
    """Untimed accuracy leg (the 20 output radii of the mixed-precision
    workload): the stepper and the FP64 stepper advance from the same seeded
    state at R0 = 300 km and stop exactly on each of 20 equally spaced radii
    up to 340 km (their end radius Rn moved to each in turn); P_ee, Pbar_ee
    and the unitarity drift are compared there.  The window is where the
    inverted-hierarchy antineutrinos convert, so the states move: the FP64
    run's change of P_ee and Pbar_ee since R0 is reported with each point."""
    from xflat_amd import capi
    from xflat_amd.seed import seeded_psi
    p = _cfg(**dict(WORKLOADS["ma1000x400"]["params"], R0=w["R0"]))
    psi0 = seeded_psi(p.abins, p.ebins, seed=2024, amp=1e-6)
    pe0, pb0 = _pee(psi0)
    a = capi.Context(p, precision=_prec(prec_name))
    b = capi.Context(p, precision=capi.XF_FP64)
    pts = []
    try:
        for c in (a, b):
            c.upload(psi0, w["R0"], DR0)
        for k in range(1, w["points"] + 1):
            rk = w["R0"] + (w["R1"] - w["R0"]) * k / w["points"]
            na, nb = a.run_to(rk), b.run_to(rk)
            pa, ra, _ = a.download()
            pb, rb, _ = b.download()
            assert ra == rk and rb == rk
            e, eb = _pee(pb)
            pts.append(dict(_errors(pa, pb), radius_km=rk, attempts=na, fp64_attempts=nb,
                            fp64_change_Pee=float(np.abs(e - pe0).max()),
                            fp64_change_Pbar_ee=float(np.abs(eb - pb0).max())))
    finally:
        a.close()
        b.close()
    return {"window": f"MA 1000 x 400, inverted hierarchy, eps0 = 1e-9, from R0 = {w['R0']:g} km (seeded 1e-6 "
                      f"perturbation) to {w['R1']:g} km; both steppers stop exactly on each of {w['points']} "
                      "equally spaced radii",
            "points": w["points"],
            "max_abs_dPee": max(x["max_abs_dPee"] for x in pts),
            "max_abs_dPbar_ee": max(x["max_abs_dPbar_ee"] for x in pts),
            "unitarity_drift": max(x["unitarity_drift"] for x in pts),
            "fp64_change_Pee": max(x["fp64_change_Pee"] for x in pts),
            "fp64_change_Pbar_ee": max(x["fp64_change_Pbar_ee"] for x in pts),
            "attempts": sum(x["attempts"] for x in pts), "fp64_attempts": sum(x["fp64_attempts"] for x in pts),
            "per_point": pts}


def time_short_launches(params, prec_name, psi0, launch, launches=50, per=20):
    """A radial loop cut into short launches (a host that stops every few
    steps to test its stop rules or write output): `launches` eager
    xf_run_steps calls of `per` attempts, one xf_run_graph replay of the same
    launches, and one launch of all launches x per attempts, each from the
    same state, after one untimed pass of each.  A launch resumes where the
    previous one stopped (the same attempts, bit for bit), so the three cover
    the same radii and the difference is the per-call cost."""
    from xflat_amd import capi
    ctx = capi.Context(params, precision=_prec(prec_name))
    ctx.upload(psi0, params.R0, DR0)
    if launch:
        _set_shape(ctx, params, launch)
    cells = params.abins * params.ebins
    out = {"launches": launches, "steps_per_launch": per}
    wall = {}
    for name in ("eager", "graph", "one_launch"):
        for timed in (False, True):
            ctx.upload(psi0, params.R0, DR0)
            t0 = time.perf_counter()
            if name == "eager":
                steps = sum(ctx.run(per)[0].steps for _ in range(launches))
            elif name == "graph":
                steps = ctx.run_graph(launches, per).steps
            else:
                steps = ctx.run(launches * per)[0].steps
            t1 = time.perf_counter()
        wall[name] = t1 - t0
        out[f"{name}_value"] = cells * steps / (t1 - t0)
        out[f"{name}_ms_per_step"] = 1e3 * (t1 - t0) / max(1, steps)
        out[f"{name}_radius_km"] = ctx.download()[1]
    ctx.close()
    out["eager_fraction_of_one_launch"] = wall["one_launch"] / wall["eager"]
    out["fixed_us_per_call"] = 1e6 * (wall["eager"] - wall["one_launch"]) / launches
    return out


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


# ---------------------------------------------------------------- CPU side
def _host():
    model, sockets = "?", set()
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name") and model == "?":
                    model = line.split(":", 1)[1].strip()
                elif line.startswith("physical id"):
                    sockets.add(line.split(":", 1)[1].strip())
    except OSError:
        pass
    aff = len(os.sched_getaffinity(0)) if hasattr(os, "sched_getaffinity") else (os.cpu_count() or 1)
    return {"model": model, "nproc": os.cpu_count(), "affinity_cpus": aff, "sockets": max(1, len(sockets)),
            "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS")}


def _ncores():
    """Threads for the CPU baseline: every CPU in this process's affinity
    mask, or the host's CPU share when the environment states one
    (OMP_NUM_THREADS), whichever is smaller."""
    h = _host()
    n = h["affinity_cpus"]
    try:
        share = int(h["omp_num_threads_env"] or 0)
    except ValueError:
        share = 0
    return max(1, min(n, share) if share > 0 else n), h


def _ref_exe(maa):
    isa = "v3"
    try:
        with open("/proc/cpuinfo") as f:
            if " avx512f" in f.read():
                isa = "v4"
    except OSError:
        pass
    return os.path.join(ROOT, "oracle", "_ref", f"xflat_ref_{'maa' if maa else 'ma'}_{isa}")


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


def cpu_baseline(workload, psi0=None, ens_cfgs=None, budget_cells=3.2e8):
    """Reference evolution path on this host's CPU share, bounded sample
    (budget_cells beam-steps at 16 threads, about 7 s; scaled with the thread
    count); the project's CPU restatement when the reference build is absent."""
    ncores, host = _ncores()
    w = WORKLOADS[workload]
    base = dict(unit="beam-steps/s", cores=ncores, host=host)
    if workload == "ensemble96":
        # the first three instances, each for a bounded number of attempts
        from xflat_amd.seed import seeded_psi
        cells = secs = 0.0
        for d in ens_cfgs[:3]:
            over = dict(d, R0=R0, Rn=1000)
            n_cells = d["Abins"] * d["Ebins"]
            n = max(8, int(budget_cells / 3 * ncores / 16 / n_cells))
            res = ref_run(over, seeded_psi(d["Abins"], d["Ebins"], seed=7, amp=1e-6), n, ncores)
            if res is None:
                return None
            cells += n_cells * res[0]
            secs += res[1]
        return dict(base, value=cells / secs, kind="reference",
                    sample=f"the first 3 ensemble instances run one after another on the reference's own evolution "
                           f"path (oracle/_ref, OpenMP {ncores} threads), {secs:.2f} s loop time")
    params = dict(w["params"], Lve=w["lve"][1]) if "lve" in w else w["params"]
    n_cells = np.prod([params[k] for k in ("Abins", "Ebins")]) * int(params.get("Pbins", 1))
    steps = max(8, int(budget_cells * ncores / 16 / n_cells))
    maa = workload.startswith("maa")
    res = ref_run(params, None if maa else psi0, steps, ncores)
    if res is not None:
        n, secs = res
        return dict(base, value=n_cells * n / secs, kind="reference",
                    sample=f"{w['desc']}, FP64, {n} attempted steps on the reference's own evolution path "
                           f"(oracle/_ref/{os.path.basename(_ref_exe(maa))}, OpenMP {ncores} threads, "
                           f"OMP_PROC_BIND=close, {secs:.2f} s loop time"
                           f"{', pure flavour initial states and Lve = 1.25e51: the reference cannot load an MAA state' if maa else ''})")
    from oracle.oracle import Oracle, set_threads
    set_threads(ncores)
    o = Oracle(_cfg(**params))
    if maa:
        o.init_state_pure()
    else:
        o.set_state(psi0, R0, DR0)
    n, _, secs = o.run(steps, log=False)
    return dict(base, value=n_cells * n / secs, kind="port",
                sample=f"{w['desc']}, FP64, {n} attempted steps on the project's CPU restatement "
                       f"(OpenMP {ncores} threads, {secs:.2f} s)")


# ---------------------------------------------------------------- roofline
def _profile_value(path, pattern):
    try:
        with open(path) as f:
            m = re.search(pattern, f.read(), re.M)
        return float(m.group(1).replace(",", "")) if m else None
    except OSError:
        return None


def _profile(workload, prec_name, launch):
    """(HBM file, SQ file) of the committed profiles when they were taken at
    this launch shape, else None."""
    for (mode, cpt, bpt, block), hbm, valu in PROFILES.get((workload, prec_name), []):
        if (launch["mode"], launch["cpt"]) != (mode, cpt):
            continue
        if (bpt is not None and launch.get("beams_per_tile") != bpt) or (block is not None and launch.get("block") != block):
            continue
        return os.path.join(PROF, hbm), os.path.join(PROF, valu)
    return None


def _roofline(workload, prec_name, cells, steps, kernel_ms, launch):
    """HBM roofline by algorithmic bytes (the metric's figure), with the
    measured HBM traffic of the committed FETCH_SIZE / WRITE_SIZE profile when
    it covers this launch shape."""
    b = _state_bytes(prec_name)
    secs = kernel_ms / 1e3
    achieved = b * cells * steps / secs / 1e9
    prof = _profile(workload, prec_name, launch)
    per_step = _profile_value(prof[0], r"^traffic_bytes_per_step:\s*([0-9.eE+]+)") if prof else None
    out = {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": achieved / HBM_PEAK_GBS,
           "traffic": per_step * steps if per_step else None,
           "traffic_source": os.path.relpath(prof[0], ROOT) if per_step else None,
           "note": f"achieved = algorithmic {b} B per beam-step (nu + anti-nu 2-spinors read and written once) over "
                   f"the stepper kernel's HIP-event time (mode {launch['mode']})"}
    if per_step:
        meas = per_step * steps / secs / 1e9
        out["measured_hbm_gbs"] = meas
        out["measured_frac"] = meas / HBM_PEAK_GBS
        out["traffic_vs_algorithmic"] = per_step / (b * cells)
        if launch["mode"] in (0, 1):
            out["note"] += ("; the resident mode keeps the state in registers across steps, so the measured HBM "
                            "traffic (the exchange rows) is a small fraction of the algorithmic bytes and HBM does "
                            "not bound the kernel: VALU issue and the exchanges' latency do (roofline_valu)")
        else:
            out["note"] += ("; the streaming mode reads the state in each of the step's three passes and writes it "
                            "once (traffic_vs_algorithmic)")
    return out


def _roofline_valu(workload, prec_name, steps, kernel_ms, launch):
    """VALU issue rate and FP32 / FP64 FLOP rates of the timed kernel from the
    committed SQ profile of the same launch shape (instructions per wave per
    step x waves x steps over the HIP-event time)."""
    prof = _profile(workload, prec_name, launch)
    if not prof:
        return None
    row = r"^\| {} \|(?: [0-9,]+ \|)? ([0-9,.]+) \|"
    per = {k: _profile_value(prof[1], row.format(k)) for k in (
        "SQ_INSTS_VALU", "SQ_INSTS_VALU_ADD_F32", "SQ_INSTS_VALU_MUL_F32", "SQ_INSTS_VALU_FMA_F32",
        "SQ_INSTS_VALU_ADD_F64", "SQ_INSTS_VALU_MUL_F64", "SQ_INSTS_VALU_FMA_F64")}
    if per["SQ_INSTS_VALU"] is None:
        return None
    per = {k: (v or 0.) for k, v in per.items()}        # a class the profile does not list issued none
    waves = launch["grid"] * (launch["block"] // 64)
    secs = kernel_ms / 1e3
    ach = per["SQ_INSTS_VALU"] * waves * steps / secs
    f32 = (per["SQ_INSTS_VALU_ADD_F32"] + per["SQ_INSTS_VALU_MUL_F32"] + 2 * per["SQ_INSTS_VALU_FMA_F32"]) * 64
    f64 = (per["SQ_INSTS_VALU_ADD_F64"] + per["SQ_INSTS_VALU_MUL_F64"] + 2 * per["SQ_INSTS_VALU_FMA_F64"]) * 64
    f32t = f32 * waves * steps / secs / 1e12
    f64t = f64 * waves * steps / secs / 1e12
    return {"bound": "valu", "achieved": ach, "peak": VALU_PEAK_WAVE_INSTR, "unit": "wave-instructions/s",
            "frac": ach / VALU_PEAK_WAVE_INSTR, "valu_per_wave_step": per["SQ_INSTS_VALU"], "waves": waves,
            "fp32_tflops": f32t, "fp32_frac": f32t / FP32_PEAK_TFLOPS,
            "fp64_tflops": f64t, "fp64_frac": f64t / FP64_PEAK_TFLOPS,
            "source": os.path.relpath(prof[1], ROOT),
            "note": "peak: 1,024 SIMDs x one wave64 VALU instruction per 4 cycles at 2.4 GHz; FLOP rates count an "
                    "FMA as 2 against 157.3 TF FP32 and 78.6 TF FP64"}


def _line(prec_name, value, steps, warmup, ms_per_step, extra):
    return dict({
        "metric": METRIC, "value": value, "unit": "beam-steps/s", "n_gpus": 1, "steps": steps,
        "warmup": warmup, "ms_per_step": ms_per_step, "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None,
        "dtype": {"mixed": "fp32 (propagator) + fp64 (state, moments, error, renorm)", "fp64": "fp64",
                  "fp32": "fp32 (state, propagator) + fp64 (moments, error, renorm)"}[prec_name],
    }, **extra)


# ---------------------------------------------------------------- workloads
def bench_ma(a, prec_name):
    from xflat_amd.seed import seeded_psi
    w = WORKLOADS["ma1000x400"]
    p = _cfg(**w["params"])
    cells = p.abins * p.ebins
    psi0 = seeded_psi(p.abins, p.ebins, seed=2024, amp=1e-6)
    launch = tuple(int(x) for x in a.launch.split(",")) if a.launch else None
    log(f"{prec_name} MA 1000x400: autotune, warmup {a.warmup}, timing {a.steps} steps")
    res = time_ctx(p, prec_name, psi0, a.steps, a.warmup, a.tune_steps, launch)
    value = cells * res["steps"] / res["wall_s"]
    log(f"{prec_name}: {value:.3e} beam-steps/s, launch {res['launch']}")
    out = _line(prec_name, value, res["steps"], a.warmup, 1e3 * res["wall_s"] / max(1, res["steps"]), {
        "data": "synthetic: pure flavour states with a seeded 1e-6 off-diagonal perturbation",
        "config": {"workload": w["desc"], "beams": cells, "precision": prec_name, "launch": res["launch"]},
        "accepted_steps": res["accepted"], "radius_km": res["r"],
        "kernel_ms_per_step": res["kernel_ms"] / max(1, res["steps"]),
        "roofline": _roofline("ma1000x400", prec_name, cells, res["steps"], res["kernel_ms"], res["launch"]),
    })
    rv = _roofline_valu("ma1000x400", prec_name, res["steps"], res["kernel_ms"], res["launch"])
    if rv:
        out["roofline_valu"] = rv
    if a.no_fp64 or prec_name == "fp64":
        return out
    ref = time_ctx(p, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
    out["fp64"] = {"value": cells * ref["steps"] / ref["wall_s"], "ms_per_step": 1e3 * ref["wall_s"] / max(1, ref["steps"]),
                   "accepted_steps": ref["accepted"], "radius_km": ref["r"], "launch": ref["launch"],
                   "roofline": _roofline("ma1000x400", "fp64", cells, ref["steps"], ref["kernel_ms"], ref["launch"]),
                   "roofline_valu": _roofline_valu("ma1000x400", "fp64", ref["steps"], ref["kernel_ms"], ref["launch"])}
    log(f"fp64: {out['fp64']['value']:.3e} beam-steps/s")
    out["radius_per_attempt_vs_fp64"] = (res["r"] - R0) / max(ref["r"] - R0, 1e-300)
    out["pee_error_vs_fp64"] = equal_radius_errors(p, psi0, res)
    out["pee_error_vs_fp64_at_20_radii"] = accuracy_window(prec_name)
    acc = out["pee_error_vs_fp64_at_20_radii"]
    log(f"20 radii {ACC_WINDOW['R0']:g}-{ACC_WINDOW['R1']:g} km: max |dP_ee| {acc['max_abs_dPee']:.2e}, "
        f"|dPbar_ee| {acc['max_abs_dPbar_ee']:.2e}; FP64 change of Pbar_ee {acc['fp64_change_Pbar_ee']:.2f}")
    if a.no_extra:
        return out
    # the same workload in the normal hierarchy (the production configuration runs both)
    pn = _cfg(**dict(w["params"], dm2="2.4E-3"))
    rn = time_ctx(pn, prec_name, psi0, a.steps, a.warmup, a.tune_steps, launch)
    fn = time_ctx(pn, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
    out["normal_hierarchy"] = dict(
        {"value": cells * rn["steps"] / rn["wall_s"], "ms_per_step": 1e3 * rn["wall_s"] / max(1, rn["steps"]),
         "accepted_steps": rn["accepted"], "radius_km": rn["r"],
         "fp64_value": cells * fn["steps"] / fn["wall_s"], "fp64_accepted_steps": fn["accepted"],
         "fp64_radius_km": fn["r"]}, **equal_radius_errors(pn, psi0, rn))
    log(f"normal hierarchy: {prec_name} {out['normal_hierarchy']['value']:.3e}, fp64 "
        f"{out['normal_hierarchy']['fp64_value']:.3e} beam-steps/s")
    li = res["launch"]                        # the headline's (tuned) shape
    out["short_launches"] = time_short_launches(p, prec_name, psi0, (li["mode"], li["cpt"], li["beams_per_tile"]))
    s = out["short_launches"]
    log(f"{s['launches']} launches x {s['steps_per_launch']} steps: eager {s['eager_value']:.3e}, "
        f"graph {s['graph_value']:.3e}, one launch {s['one_launch_value']:.3e} beam-steps/s "
        f"({s['fixed_us_per_call']:.1f} us per call)")
    if prec_name == "fp32":
        mx = time_ctx(p, "mixed", psi0, a.steps, a.warmup, a.tune_steps)
        out["fp64_state_mixed"] = dict(
            {"value": cells * mx["steps"] / mx["wall_s"], "ms_per_step": 1e3 * mx["wall_s"] / max(1, mx["steps"]),
             "accepted_steps": mx["accepted"], "radius_km": mx["r"], "launch": mx["launch"]},
            **equal_radius_errors(p, psi0, mx))
        log(f"mixed (fp64 state): {out['fp64_state_mixed']['value']:.3e} beam-steps/s")
    return out


def bench_maa(a, prec_name):
    """MAA 256 x 64 x 200 for each luminosity of the sweep; the timed figure
    is the middle one's, the others are listed."""
    from xflat_amd.seed import azimuthal_psi
    w = WORKLOADS["maa256x64x200"]
    sweep = []
    for lve in w["lve"]:
        p = _cfg(**dict(w["params"], Lve=lve))
        cells = p.abins * p.pbins * p.ebins
        psi0 = azimuthal_psi(p.abins, p.pbins, p.ebins, seed=31, amp=1e-7)
        launch = tuple(int(x) for x in a.launch.split(",")) if a.launch else None
        log(f"{prec_name} MAA 256x64x200 Lve = {lve}: {'launch ' + a.launch if launch else 'autotune'}, "
            f"warmup {a.warmup}, timing {a.steps} steps")
        res = time_ctx(p, prec_name, psi0, a.steps, a.warmup, a.tune_steps, launch)
        pt = {"Lve_erg_s": float(lve), "value": cells * res["steps"] / res["wall_s"],
              "ms_per_step": 1e3 * res["wall_s"] / max(1, res["steps"]),
              "kernel_ms_per_step": res["kernel_ms"] / max(1, res["steps"]),
              "accepted_steps": res["accepted"], "radius_km": res["r"], "launch": res["launch"],
              "roofline": _roofline("maa256x64x200", prec_name, cells, res["steps"], res["kernel_ms"], res["launch"]),
              "roofline_valu": _roofline_valu("maa256x64x200", prec_name, res["steps"], res["kernel_ms"], res["launch"])}
        if not a.no_fp64 and prec_name != "fp64":
            ref = time_ctx(p, "fp64", psi0, a.steps, a.warmup, a.tune_steps)
            pt["fp64"] = {"value": cells * ref["steps"] / ref["wall_s"],
                          "ms_per_step": 1e3 * ref["wall_s"] / max(1, ref["steps"]),
                          "accepted_steps": ref["accepted"], "radius_km": ref["r"], "launch": ref["launch"]}
            pt["radius_per_attempt_vs_fp64"] = (res["r"] - R0) / max(ref["r"] - R0, 1e-300)
            pt["pee_error_vs_fp64"] = equal_radius_errors(p, psi0, res)
        log(f"{prec_name}: {pt['value']:.3e} beam-steps/s, launch {res['launch']}")
        sweep.append(pt)
    mid = sweep[len(sweep) // 2]
    out = _line(prec_name, mid["value"], a.steps, a.warmup, mid["ms_per_step"], {
        "data": "synthetic: pure flavour states with a seeded azimuthal perturbation "
                "1e-7 sum_{m=1..4} cos(m phi + phi_m), random phases",
        "config": {"workload": w["desc"], "beams": 256 * 64 * 200, "precision": prec_name, "launch": mid["launch"]},
        "kernel_ms_per_step": mid["kernel_ms_per_step"], "roofline": mid["roofline"], "lve_sweep": sweep})
    if "pee_error_vs_fp64" in mid:
        out["pee_error_vs_fp64"] = {k: max(x["pee_error_vs_fp64"][k] for x in sweep)
                                    for k in ("max_abs_dPee", "max_abs_dPbar_ee", "unitarity_drift")}
    if not a.no_cpu:
        log("cpu baseline (MAA) ...")
        out["cpu_baseline"] = cpu_baseline("maa256x64x200", budget_cells=a.cpu_cells * (0.25 if a.workload is None else 1.0))
    return out


def bench_ensemble(a, prec_name):
    from xflat_amd.seed import ensemble_configs, seeded_psi
    w = WORKLOADS["ensemble96"]
    cfgs = ensemble_configs(96)
    params = [_cfg(**dict(d, R0=R0, Rn=1000)) for d in cfgs]
    inits = lambda i: seeded_psi(params[i].abins, params[i].ebins, seed=5000 + i, amp=1e-6)  # noqa: E731
    log(f"{prec_name} ensemble of 96: autotune, warmup {a.warmup}, timing {a.steps} steps")
    ens, res = time_ensemble(params, prec_name, inits, a.steps, a.warmup, a.tune_steps)
    cells = res["cells"]
    done = sum(p.abins * p.ebins * n for p, n in zip(params, res["steps"]))
    value = done / res["wall_s"]
    out = _line(prec_name, value, int(max(res["steps"])), a.warmup, 1e3 * res["wall_s"] / max(1, max(res["steps"])), {
        "data": "synthetic: seeded instance parameters; pure flavour states with a seeded 1e-6 off-diagonal "
                "perturbation",
        "config": {"workload": w["desc"], "instances": 96, "beams": cells, "precision": prec_name,
                   "launch": res["launch"]},
        "kernel_ms_per_step": res["kernel_ms"] / max(1, max(res["steps"])),
        "accepted_steps_total": int(sum(res["accepted"])),
        "roofline": _roofline("ensemble96", prec_name, cells, max(res["steps"]), res["kernel_ms"], res["launch"]),
    })
    rv = _roofline_valu("ensemble96", prec_name, max(res["steps"]), res["kernel_ms"], res["launch"])
    if rv:
        out["roofline_valu"] = rv
    log(f"{prec_name}: {value:.3e} beam-steps/s, launch {res['launch']}")
    if not a.no_fp64 and prec_name != "fp64":
        ens64, ref = time_ensemble(params, "fp64", inits, a.steps, a.warmup, a.tune_steps)
        out["fp64"] = {"value": sum(p.abins * p.ebins * n for p, n in zip(params, ref["steps"])) / ref["wall_s"],
                       "ms_per_step": 1e3 * ref["wall_s"] / max(1, max(ref["steps"])),
                       "accepted_steps_total": int(sum(ref["accepted"])), "launch": ref["launch"]}
        log(f"fp64: {out['fp64']['value']:.3e} beam-steps/s")
        # every FP64 instance from its initial state, stopped exactly on the
        # radius its FP32 instance reached
        for i, p in enumerate(params):
            ens64.upload(i, inits(i), p.R0, DR0)
        n64 = ens64.run_to(res["r"])
        worst = []
        for i in range(96):
            psi, _, _ = ens.download(i)
            psi64, r64, _ = ens64.download(i)
            assert r64 == res["r"][i]
            worst.append(_errors(psi, psi64))
        ens64.close()
        out["pee_error_vs_fp64"] = {
            "note": "per instance, against the FP64 stepper from the same state stopped exactly on the same radius",
            "per_instance_max_abs_dPee": [x["max_abs_dPee"] for x in worst],
            "max_abs_dPee": max(x["max_abs_dPee"] for x in worst),
            "max_abs_dPbar_ee": max(x["max_abs_dPbar_ee"] for x in worst),
            "unitarity_drift": max(x["unitarity_drift"] for x in worst),
            "fp64_attempts_total": int(sum(n64)), "attempts_total": int(sum(res["steps"]))}
        log(f"per-instance max |dP_ee| vs fp64 at equal radii: {out['pee_error_vs_fp64']['max_abs_dPee']:.2e}")
    ens.close()
    if not a.no_cpu:
        log("cpu baseline (ensemble) ...")
        out["cpu_baseline"] = cpu_baseline("ensemble96", ens_cfgs=cfgs,
                                           budget_cells=a.cpu_cells * (0.25 if a.workload is None else 1.0))
    return out


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


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--workload", default=None, choices=sorted(WORKLOADS),
                    help="one workload as its own line (default: the headline with the others as extra keys)")
    ap.add_argument("--steps", type=int, default=None)
    ap.add_argument("--warmup", type=int, default=None)
    ap.add_argument("--precision", default="fp32", choices=["fp32", "mixed", "fp64"])
    ap.add_argument("--tune-steps", type=int, default=30)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--no-fp64", action="store_true")
    ap.add_argument("--no-extra", action="store_true",
                    help="headline only: no normal hierarchy, short launches, mixed leg or other workloads")
    ap.add_argument("--cpu-cells", type=float, default=3.2e8, help="beam-steps of each CPU-baseline sample at 16 threads")
    ap.add_argument("--calibrate-copy", action="store_true")
    ap.add_argument("--launch", default=None,
                    help="mode,cpt,bpt: fixed launch shape instead of autotuning (MA and MAA workloads)")
    a = ap.parse_args()
    if a.gpus != 1:
        print(f"bench.py measures one GPU: --gpus {a.gpus} is not supported (several GPUs would shard the "
              "angle bins, one all-reduce of the moments per exchange; that path is not built).", file=sys.stderr)
        sys.exit(2)
    wl = a.workload or "ma1000x400"
    w = WORKLOADS[wl]
    a.steps = a.steps if a.steps is not None else w["steps"]
    a.warmup = a.warmup if a.warmup is not None else w["warmup"]
    fn = {"ma1000x400": bench_ma, "maa256x64x200": bench_maa, "ensemble96": bench_ensemble}[wl]
    out = fn(a, a.precision)
    if wl == "ma1000x400" and not a.no_cpu:
        log("cpu baseline ...")
        from xflat_amd.seed import seeded_psi
        out["cpu_baseline"] = cpu_baseline("ma1000x400", psi0=seeded_psi(1000, 400, seed=2024, amp=1e-6),
                                           budget_cells=a.cpu_cells)
    if a.workload is None and not a.no_extra:
        out["other_workloads"] = {"maa256x64x200": bench_maa(a, a.precision),
                                  "ensemble96": bench_ensemble(a, a.precision)}
    if a.calibrate_copy:
        out["copy_calibration_bytes"] = calibrate_copy()
    print(json.dumps(out))


if __name__ == "__main__":
    main()

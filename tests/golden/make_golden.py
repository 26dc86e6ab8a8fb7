"""Generates the golden fixtures tests/golden/*.npz from the reference itself.

Runs the reference's evolution path, compiled from its unmodified sources by
oracle/Makefile into oracle/_ref/ (with the project's raw-file IO plugin
oracle/ref_io_recorder.cpp), on small configurations:

* MA and SA cases start from a seeded perturbed state injected through the
  reference's resume path (fio::fillInitData, src/IO_f.cpp:197-250, which sets
  R0, dr and psi; SA calls it from Phy_SA.cpp:231);
* MAA cases start from pure flavour states (the reference's MAA resume and
  perturbation are switched off, Phy_MAA.cpp:281,412), or, for maa_pert, from
  the reference's own -DPERTERB perturbation (NBeam.cpp:260-275): each NBeam
  draws two glibc rand() values per energy bin, nubm_0 first (Nmr.cpp:61),
  in initBeam's loop order (Phy_MAA.cpp:358-361) on one OpenMP thread; the
  generator reproduces that state by calling the same libc rand() stream
  (default seed) in the same order.  The states are not SU(2) partners, so
  they exercise the 4-species path.

Every accepted step is recorded by the IO hook of Nmr.cpp:238 as
(it = attempt index, r after the step, dr used); the wave functions are kept
for a few records.  Run in the build container (the reference is not on the
GPU host):  python tests/golden/make_golden.py
"""
from __future__ import annotations

import json
import os
import re
import subprocess
import sys
import tempfile

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(os.path.dirname(HERE))
sys.path.insert(0, ROOT)

from xflat_amd.config import config_text  # noqa: E402
from xflat_amd.seed import seeded_psi, pure_psi  # noqa: E402

REF_DIR = os.path.join(ROOT, "oracle", "_ref")

# name -> (config overrides, seed or None for pure states, amplitude, records to keep)
CASES = {
    # the bulb MA benchmark physics (inverted hierarchy) at R_nu, small grid:
    # the first attempts with dr = 1e-3 km are rejected until dr ~ 4e-8 km.
    "ma_r10_ih": (dict(Abins=32, Ebins=16, R0=10, Ts=400, dm2="-2.4E-3"), 11, 1e-6, 6),
    "ma_r10_nh": (dict(Abins=32, Ebins=16, R0=10, Ts=300, dm2="2.4E-3"), 12, 1e-6, 6),
    # further out: larger steps, tighter tolerance (the 1000 x 400 benchmark uses eps0 = 1e-9)
    "ma_r50_ih": (dict(Abins=48, Ebins=24, R0=50, Ts=300, dm2="-2.4E-3", eps0="1.E-9"), 13, 1e-6, 6),
    # no matter term (hasMatter= 0), vacuum + nu-nu only
    "ma_vac": (dict(Abins=16, Ebins=16, R0=30, Ts=200, hasMatter=0), 14, 1e-6, 6),
    # MAA: 4-moment kernel, pure flavour initial states
    "maa_r50": (dict(Abins=8, Pbins=6, Ebins=12, R0=50, Ts=200), None, 0.0, 6),
    "maa_r10": (dict(Abins=8, Pbins=4, Ebins=12, R0=10, Ts=200, dm2="2.4E-3"), None, 0.0, 6),
    # MAA from the reference's -DPERTERB initial states (glibc rand())
    "maa_pert": (dict(_module="maa_pert", Abins=8, Pbins=6, Ebins=12, R0=50, Ts=200), None, 0.0, 6),
    # SA: one trajectory (reference built with -DSA), H_vv scaled by D(r)
    "sa_r10_ih": (dict(_module="sa", Abins=1, Ebins=64, R0=10, Ts=400, dm2="-2.4E-3"), 21, 1e-6, 8),
    "sa_r50_nh": (dict(_module="sa", Abins=1, Ebins=100, R0=50, Ts=300, dm2="2.4E-3", eps0="1.E-9"), 22, 1e-6, 8),
    "sa_vac": (dict(_module="sa", Abins=1, Ebins=48, R0=30, Ts=200, hasMatter=0), 23, 1e-6, 6),
}


def _perterb_psi(nbeam, ne):
    """The -DPERTERB initial state of NBeam::NBeam(prtc) (NBeam.cpp:246-275)
    for the first beam set, from libc's rand() with its default seed:
    eps = (rand() 1e-8 - 0.5e-8) / RAND_MAX, eta likewise, per energy bin;
    (one_sq, 0, eps, eta) for nu_e / anti-nu_e, (eps, 0, one_sq, eta) else."""
    import ctypes
    libc = ctypes.CDLL("libc.so.6")
    libc.srand(1)                         # glibc's state before any srand()
    rand_max = 2147483647.0
    psi = np.zeros((nbeam, 4, 4, ne))
    for b in range(nbeam):
        for n in range(4):
            for e in range(ne):
                e1 = libc.rand() * 1.e-8 - (.5 * 1.e-8)
                e2 = libc.rand() * 1.e-8 - (.5 * 1.e-8)
                eps, eta = e1 / rand_max, e2 / rand_max
                one_sq = np.sqrt(1 - eps * eps - eta * eta)
                psi[b, n, :, e] = (one_sq, 0., eps, eta) if n < 2 else (eps, 0., one_sq, eta)
    return psi


def _run_case(name, over, seed, amp, keep, isa="v3"):
    over = dict(over)
    module = over.pop("_module", None)
    maa = int(over.get("Pbins", 1)) > 1
    sa = module == "sa"
    pert = module == "maa_pert"
    kind = "sa" if sa else "maa_pert" if pert else "maa" if maa else "ma"
    exe = os.path.join(REF_DIR, f"xflat_ref_{kind}_{'v3' if pert else isa}")
    if not os.path.exists(exe):
        raise SystemExit(f"{exe} missing: run `make -C oracle ref` first")
    txt = config_text(**over)
    nth = int(over["Abins"]); nph = int(over.get("Pbins", 1)); ne = int(over["Ebins"])
    nbeam = 1 if sa else nth * nph
    r0 = float(over["R0"]); dr0 = 0.001
    if seed is not None:
        psi0 = seeded_psi(nbeam, ne, seed, amp)
    elif pert:
        psi0 = _perterb_psi(nbeam, ne)
    else:
        psi0 = pure_psi(nbeam, ne)
    with tempfile.TemporaryDirectory(dir=os.path.join(ROOT, "_work") if os.path.isdir(os.path.join(ROOT, "_work")) else None) as td:
        cfg = os.path.join(td, "config.txt")
        with open(cfg, "w") as f:
            f.write(txt)
        out = os.path.join(td, "rec.bin")
        env = dict(os.environ, XFLAT_REF_OUT=out, XFLAT_REF_PSI_EVERY="1", OMP_NUM_THREADS="1" if pert else "4")
        args = [exe, cfg]
        if seed is not None:
            init = os.path.join(td, "init.bin")
            with open(init, "wb") as f:
                np.array([nbeam * 4, 4, ne], dtype=np.int64).tofile(f)
                np.array([r0, dr0], dtype=np.float64).tofile(f)
                psi0.astype(np.float64).tofile(f)
            env["XFLAT_REF_INIT"] = init
            args.append(init)
        res = subprocess.run(args, env=env, capture_output=True, text=True, cwd=td, timeout=600)
        if res.returncode != 0:
            raise SystemExit(f"{name}: reference failed\n{res.stdout}\n{res.stderr}")
        m = re.search(r"Total iterations:\s*(\d+)\s+Correct iterations:\s*(\d+)\s+at radius\(Km\):\s*([0-9.eE+-]+)", res.stdout)
        tot, acc = int(m.group(1)), int(m.group(2))
        raw = np.fromfile(out, dtype=np.uint8)
    hdr = raw[:32].view(np.int64)
    nnb, ncmp, nebin = int(hdr[0]), int(hdr[1]), int(hdr[2])
    per = nnb * ncmp * nebin
    pos = 32
    its, rs, drs, psis = [], [], [], []
    while pos < raw.size:
        it = raw[pos:pos + 8].view(np.int64)[0]
        r, dr = raw[pos + 8:pos + 24].view(np.float64)
        has = raw[pos + 24:pos + 32].view(np.int64)[0]
        pos += 32
        its.append(it); rs.append(r); drs.append(dr)
        if has:
            psis.append(raw[pos:pos + 8 * per].view(np.float64).reshape(nbeam, 4, 4, ne).copy())
            pos += 8 * per
        else:
            psis.append(None)
    nrec = len(its)
    sel = sorted(set(np.linspace(0, nrec - 1, keep).round().astype(int).tolist())) if nrec else []
    return dict(
        config=np.array(txt), module=np.int64(2 if sa else 1 if maa else 0),
        psi0=psi0, r0=np.float64(r0), dr0=np.float64(dr0),
        it=np.array(its, dtype=np.int64), r=np.array(rs), dr=np.array(drs),
        psi_idx=np.array(sel, dtype=np.int64),
        psi=np.stack([psis[i] for i in sel]) if sel else np.zeros((0, nbeam, 4, 4, ne)),
        total=np.int64(tot), accepted=np.int64(acc),
        meta=np.array(json.dumps(dict(case=name, seed=seed, amp=amp, isa=isa,
                                      source="oracle/_ref reference build"))),
    )


def main(names=None):
    for name, (over, seed, amp, keep) in CASES.items():
        if names and name not in names:
            continue
        d = _run_case(name, over, seed, amp, keep)
        path = os.path.join(HERE, f"{name}.npz")
        np.savez_compressed(path, **d)
        print(f"{name}: {int(d['total'])} attempts, {int(d['accepted'])} accepted, "
              f"r -> {d['r'][-1] if d['r'].size else d['r0']:.9f}, {os.path.getsize(path)//1024} KiB")


if __name__ == "__main__":
    main(sys.argv[1:])

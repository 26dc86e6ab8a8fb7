"""The batched ensemble at its full size: the 96 seeded instances of the
ensemble workload (angle bins 64..1500, energy bins 32..400, 14.2 M beams in
all) in one cooperative launch with the autotuner's shape.

- FP64: every instance against its own oracle run from the same state (the
  per-instance reference is one run of nmr::evolutionLoop, src/Nmr.cpp:89-262):
  the same attempted and accepted step counts, the same radius to 1e-12, and
  P_ee / Pbar_ee within 1e-9.
- FP32 state: every instance within 1e-5 of the FP64 stepper stopped exactly
  on the radius the FP32 instance reached, unitarity drift below 1e-6.
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu

PEE_TOL = 1e-9
STEPS = 12


def _pee(psi):
    return psi[:, 0, 0] ** 2 + psi[:, 0, 1] ** 2, psi[:, 1, 0] ** 2 + psi[:, 1, 1] ** 2


def test_ensemble96_full_size_fp64_matches_oracle_fp32_close():
    from xflat_amd import capi
    from xflat_amd.config import bulb_params
    from xflat_amd.seed import ensemble_configs, seeded_psi
    from oracle.oracle import Oracle
    if capi.device_count() < 1:
        pytest.fail("no HIP device visible")
    cfgs = ensemble_configs(96)
    ps = [bulb_params(**dict(d, R0=10, Rn=1000)) for d in cfgs]
    inits = [seeded_psi(p.abins, p.ebins, seed=5000 + i, amp=1e-6) for i, p in enumerate(ps)]
    cells = sum(p.abins * p.ebins for p in ps)
    print(f"[ensemble96] {len(ps)} instances, {cells} beams", flush=True)
    out = {}
    for name, pr in (("fp64", capi.XF_FP64), ("fp32", capi.XF_FP32)):
        ens = capi.Ensemble(ps, precision=pr)
        for i, p in enumerate(ps):
            ens.upload(i, inits[i], p.R0, 1e-3)
        _, ms = ens.autotune(4)
        for i, p in enumerate(ps):
            ens.upload(i, inits[i], p.R0, 1e-3)
        infos = ens.run(STEPS)
        print(f"[ensemble96] {name}: launch {ens.launch()}, tuned {ms:.3f} ms per step", flush=True)
        out[name] = (ens, infos)
    ens64, infos64 = out["fp64"]
    ens32, infos32 = out["fp32"]
    d64 = []
    for i, p in enumerate(ps):
        o = Oracle(p)
        o.set_state(inits[i], p.R0, 1e-3)
        o.run(STEPS, log=False)
        opsi, orr, _, (tot, acc) = o.state()
        o.close()
        psi, r, _ = ens64.download(i)
        info = infos64[i]
        assert (info.total_steps, info.total_accepted) == (tot, acc), i
        assert abs(r - orr) <= 1e-12 * orr, i
        (pe, pb), (oe, ob) = _pee(psi), _pee(opsi)
        d = max(np.abs(pe - oe).max(), np.abs(pb - ob).max())
        assert d <= PEE_TOL, (i, d)
        d64.append(d)
        if i % 16 == 15:
            print(f"[ensemble96] fp64 vs oracle: instances 0..{i} agree, max |dP| {max(d64):.2e}", flush=True)
    # FP64 from the same states, stopped on each FP32 instance's radius
    r32 = [x.r for x in infos32]
    for i, p in enumerate(ps):
        ens64.upload(i, inits[i], p.R0, 1e-3)
    ens64.run_to(r32)
    d32, drift = [], []
    for i in range(len(ps)):
        psi, r, _ = ens32.download(i)
        ref, rr, _ = ens64.download(i)
        assert r == rr == r32[i]
        (fe, fb), (e, b) = _pee(psi), _pee(ref)
        d32.append(max(np.abs(fe - e).max(), np.abs(fb - b).max()))
        drift.append(float(np.abs(np.sum(psi[:, 0] ** 2, axis=1) - 1).max()))
    print(f"[ensemble96] fp32 vs fp64 at equal radii: max |dP| {max(d32):.2e}, unitarity drift {max(drift):.2e}",
          flush=True)
    ens64.close()
    ens32.close()
    assert max(d32) <= 1e-5
    assert max(drift) <= 1e-6

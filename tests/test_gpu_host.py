"""GPU checks of the command-line host (xflat_amd/bin/xflat_run): the
reference's driver and evolution loop (src/main.cpp, src/Nmr.cpp:89-262)
with its NetCDF output (src/IO_f.cpp) and resume (IO_f.cpp:197-250), over the
C-ABI.  The Snapshot records and a resumed run are compared with the CPU
oracle on the same inputs (FP64 bar: max |dP_ee| <= 1e-9)."""
import os
import subprocess

import numpy as np
import pytest
from scipy.io import netcdf_file

from conftest import ROOT

pytestmark = pytest.mark.gpu

EXE = os.path.join(ROOT, "xflat_amd", "bin", "xflat_run")


def _pee(psi):
    return psi[..., 0, 0, :] ** 2 + psi[..., 0, 1, :] ** 2


def _run(args, cwd):
    if not os.path.exists(EXE):
        pytest.fail("xflat_amd/bin/xflat_run is not built")
    print("[host]", " ".join(args), flush=True)
    r = subprocess.run([EXE] + args, cwd=cwd, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-2000:]
    return r.stdout


def test_host_snapshots_probability_and_resume_match_oracle(tmp_path):
    from xflat_amd.config import config_text, parse_config_text
    from xflat_amd.seed import pure_psi
    from oracle.oracle import Oracle

    over = dict(Abins=24, Ebins=16, R0=30, Rn=250, dr="1e-3", eps0="1.E-8", dumpMode=3,
                filePrefix=str(tmp_path) + "/", itr_step1=1, r_step1="1e9", t_step1=100000,
                itr_step2=4, r_step2="1e9", t_step2=100000, newFile_step=1000000, sync_step=7)
    cfg = tmp_path / "config.txt"
    cfg.write_text(config_text(Ts=150, **over))
    out = _run([str(cfg), "--chunk", "37"], str(tmp_path))
    assert "Total iterations: 150" in out
    p = parse_config_text(config_text(Ts=150, **over))

    # oracle from the same pure initial state
    o = Oracle(p)
    psi0 = pure_psi(24, 16)
    o.set_state(psi0, 30.0, 1e-3)
    n, rec, _ = o.run(150)
    acc = np.nonzero(rec["accepted"])[0]
    with netcdf_file(str(tmp_path / "Snapshot0_0_.nc"), "r", mmap=False) as nc:
        r = nc.variables["r"][:].copy()
        dr = nc.variables["dr"][:].copy()
        snaps = nc.variables["psi"][:].copy()
    # one record per accepted step (itr_step1 = 1), r after the step, the dr used
    assert len(r) == len(acc)
    np.testing.assert_allclose(r, rec["r"][acc] + rec["dr"][acc], rtol=1e-12)
    np.testing.assert_allclose(dr, rec["dr"][acc], rtol=2e-6)
    opsi, orr, odr, _ = o.state()
    assert np.abs(_pee(snaps[-1]) - _pee(opsi)).max() <= 1e-9
    assert abs(r[-1] - orr) <= 1e-12 * orr

    # Probability records at accepted attempts with it % 4 == 0 (IO_f.cpp:408)
    with netcdf_file(str(tmp_path / "Probability0_0_.nc"), "r", mmap=False) as nc:
        pr = nc.variables["r"][:].copy()
        p0 = nc.variables["prtcl0"][:].copy()
    want = [k for k, it in enumerate(acc) if it % 4 == 0]
    assert len(pr) == len(want)
    np.testing.assert_array_equal(pr, r[want])
    # prtcl_q = sum_e |a_q|^2 fv of species q % 2 of the same state
    fv = o.tables()["fv"]
    x = snaps[want[-1]]
    np.testing.assert_allclose(p0[-1, :, 0], ((x[:, 0, 0] ** 2 + x[:, 0, 1] ** 2) * fv[0]).sum(axis=1), rtol=1e-13)

    # resume from the Snapshot: R0, dr and psi of its last record
    cfg2 = tmp_path / "config2.txt"
    over2 = dict(over, filePrefix=str(tmp_path) + "/resumed_", dumpMode=1, itr_step1=10)
    cfg2.write_text(config_text(Ts=80, **over2))
    out2 = _run([str(cfg2), str(tmp_path / "Snapshot0_0_.nc"), "--chunk", "25"], str(tmp_path))
    assert "Loading the provided data file" in out2 and "Total iterations: 80" in out2
    o2 = Oracle(p)
    o2.set_state(snaps[-1].reshape(24, 4, 4, 16), float(r[-1]), float(dr[-1]))
    n2, rec2, _ = o2.run(80)
    opsi2, orr2, _, _ = o2.state()
    with netcdf_file(str(tmp_path / "resumed_Snapshot0_0_.nc"), "r", mmap=False) as nc:
        r2 = nc.variables["r"][:].copy()
        last2 = nc.variables["psi"][-1].copy()
    acc2 = np.nonzero(rec2["accepted"])[0]
    want2 = [k for k, it in enumerate(acc2) if it % 10 == 0]
    np.testing.assert_allclose(r2, rec2["r"][acc2][want2] + rec2["dr"][acc2][want2], rtol=1e-12)
    # the last record is the last accepted attempt with it % 10 == 0
    k_last = acc2[want2[-1]]
    o3 = Oracle(p)
    o3.set_state(snaps[-1].reshape(24, 4, 4, 16), float(r[-1]), float(dr[-1]))
    o3.run(int(k_last) + 1)
    opsi3, _, _, _ = o3.state()
    assert np.abs(_pee(last2) - _pee(opsi3)).max() <= 1e-9


def test_host_single_angle_module_matches_oracle(tmp_path):
    """--module sa (the reference's -DSA build, Phy_SA.cpp): one trajectory,
    a Snapshot without the theta dimension, every accepted step recorded."""
    from xflat_amd.config import config_text, parse_config_text, XF_SA
    from oracle.oracle import Oracle
    from xflat_amd.seed import pure_psi

    over = dict(Abins=1, Ebins=40, R0=30, Rn=250, dr="1e-3", eps0="1.E-8", dumpMode=1,
                filePrefix=str(tmp_path) + "/", itr_step1=1, r_step1="1e9", t_step1=100000,
                newFile_step=1000000, sync_step=7)
    cfg = tmp_path / "config.txt"
    cfg.write_text(config_text(Ts=120, **over))
    out = _run([str(cfg), "--module", "sa", "--chunk", "29"], str(tmp_path))
    assert "Total iterations: 120" in out
    p = parse_config_text(config_text(Ts=120, **over), module=XF_SA)
    o = Oracle(p)
    o.set_state(pure_psi(1, 40), 30.0, 1e-3)
    n, rec, _ = o.run(120)
    acc = np.nonzero(rec["accepted"])[0]
    with netcdf_file(str(tmp_path / "Snapshot0_0_.nc"), "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "prtcl", "comp", "ebin"]
        r = nc.variables["r"][:].copy()
        snaps = nc.variables["psi"][:].copy()
    assert len(r) == len(acc)
    np.testing.assert_allclose(r, rec["r"][acc] + rec["dr"][acc], rtol=1e-12)
    opsi, orr, _, _ = o.state()
    assert np.abs(_pee(snaps[-1][None]) - _pee(opsi)).max() <= 1e-9

"""CPU checks of the NetCDF output and resume layer (include/xflat_io.h).

The files must follow the reference's schema (src/IO_f.cpp:90-160, dimension
names of Phy_MA.cpp:92-99 / Phy_MAA.cpp:94-102) and be readable by an
independent NetCDF classic reader: scipy.io.netcdf_file checks the codec.
Resume (IO_f.cpp:197-250) reads the last record back, also from files written
by scipy in both classic versions.
"""
import numpy as np
import pytest
from scipy.io import netcdf_file

from xflat_amd import capi
from xflat_amd.config import config_text, parse_config_text, n_beams, XF_SA


def _params(module=None, **over):
    return parse_config_text(config_text(**over), module=module)


def _rand_psi(p, rng):
    return rng.standard_normal((n_beams(p), 4, 4, p.ebins))


def test_snapshot_schema_and_records_ma(tmp_path):
    p = _params(Abins=5, Ebins=3, Tve=2.76, Tv_e=4.01, Tvt=6.26, Tv_t=6.26, eta_ve=0.5)
    rng = np.random.default_rng(1)
    path = str(tmp_path / "snap.nc")
    recs = [(10.0 + k, 1e-3 * (k + 1), _rand_psi(p, rng)) for k in range(3)]
    with capi.NcFile(p, path) as f:
        for r, dr, psi in recs:
            f.append_snapshot(r, dr, psi)
        assert f.records() == 3
    with netcdf_file(path, "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "theta", "prtcl", "comp", "ebin"]
        assert nc.dimensions["r"] is None                      # unlimited
        assert (nc.dimensions["theta"], nc.dimensions["prtcl"], nc.dimensions["comp"],
                nc.dimensions["ebin"]) == (5, 4, 4, 3)
        assert list(nc.variables) == ["r", "dr", "psi"]
        psi = nc.variables["psi"]
        assert psi.dimensions == ("r", "theta", "prtcl", "comp", "ebin")
        np.testing.assert_array_equal(nc.variables["r"][:], [r for r, _, _ in recs])
        np.testing.assert_array_equal(nc.variables["dr"][:], [dr for _, dr, _ in recs])
        for k, (_, _, x) in enumerate(recs):
            np.testing.assert_array_equal(psi[k], x)
        # attributes of IO_f.cpp:119-128, with the reference's names: for
        # flavour i the literal is offset by i characters
        a = psi._attributes
        assert a["E0"] == pytest.approx(p.E0) and a["E1"] == pytest.approx(p.E1)
        assert a["Teprature for particle"] == pytest.approx(p.T[0])
        assert a["Teprature for anti-particle"] == pytest.approx(p.T[1])
        assert a["eprature for particle"] == pytest.approx(p.T[1])
        assert a["eprature for anti-particle"] == pytest.approx(p.T[2])
        assert a["eta for particle"] == pytest.approx(p.eta[0])
        assert a["ta for anti-particle"] == pytest.approx(p.eta[2])


def test_snapshot_schema_maa(tmp_path):
    p = _params(Abins=3, Pbins=4, Ebins=2)
    assert p.module == 1
    rng = np.random.default_rng(2)
    x = _rand_psi(p, rng)
    path = str(tmp_path / "snap.nc")
    with capi.NcFile(p, path) as f:
        f.append_snapshot(50.0, 2e-3, x)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "theta", "phi", "prtcl", "comp", "ebin"]
        v = nc.variables["psi"][:]
        assert v.shape == (1, 3, 4, 4, 4, 2)
        np.testing.assert_array_equal(v[0].reshape(12, 4, 4, 2), x)


def test_snapshot_probability_and_resume_sa(tmp_path):
    """SA (Phy_SA.cpp): one trajectory; the Snapshot has no theta dimension
    (getDimInfo, Phy_SA.cpp:74-80) and the Probability grid is 1 x 1
    (theta = beamLen / (Pbins * 4), IO_f.cpp:93-94)."""
    p = _params(module=XF_SA, Abins=7, Ebins=5)
    assert p.module == XF_SA and n_beams(p) == 1
    rng = np.random.default_rng(9)
    xs = [_rand_psi(p, rng) for _ in range(2)]
    fv = rng.random((4, p.ebins))
    snap, prob = str(tmp_path / "snap.nc"), str(tmp_path / "prob.nc")
    with capi.NcFile(p, snap) as f, capi.NcFile(p, prob, probability=True) as g:
        for k, x in enumerate(xs):
            f.append_snapshot(10.0 + k, 1e-6 * (k + 1), x)
            g.append_probability(10.0 + k, x, fv)
    with netcdf_file(snap, "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "prtcl", "comp", "ebin"]
        v = nc.variables["psi"]
        assert v.dimensions == ("r", "prtcl", "comp", "ebin")
        np.testing.assert_array_equal(v[1], xs[1][0])
    with netcdf_file(prob, "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "theta", "phi"]
        assert (nc.dimensions["theta"], nc.dimensions["phi"]) == (1, 1)
        want = ((xs[0][0, 0, 0] ** 2 + xs[0][0, 0, 1] ** 2) * fv[0]).sum()
        np.testing.assert_allclose(nc.variables["prtcl0"][0, 0, 0], want, rtol=1e-14)
    psi, r, dr = capi.snapshot_read_last(snap, p)
    assert (r, dr) == (11.0, 2e-6)
    np.testing.assert_array_equal(psi, xs[1])


def test_probability_values(tmp_path):
    p = _params(Abins=4, Pbins=1, Ebins=6)
    rng = np.random.default_rng(3)
    x = _rand_psi(p, rng)
    fv = rng.random((4, p.ebins))
    path = str(tmp_path / "prob.nc")
    with capi.NcFile(p, path, probability=True) as f:
        f.append_probability(12.5, x, fv)
        f.append_probability(13.5, 2 * x, fv)
    with netcdf_file(path, "r", mmap=False) as nc:
        assert list(nc.dimensions) == ["r", "theta", "phi"]
        assert list(nc.variables) == ["r", "prtcl0", "prtcl1", "prtcl2", "prtcl3"]
        np.testing.assert_array_equal(nc.variables["r"][:], [12.5, 13.5])
        for q in range(4):
            # sum_e |a|^2 fv of species q % 2 (IO_f.cpp:346-351)
            want = ((x[:, q, 0] ** 2 + x[:, q, 1] ** 2) * fv[q % 2]).sum(axis=1)
            got = nc.variables[f"prtcl{q}"][:]
            assert got.shape == (2, 4, 1)
            np.testing.assert_allclose(got[0, :, 0], want, rtol=1e-14)
            np.testing.assert_allclose(got[1, :, 0], 4 * want, rtol=1e-14)


def test_resume_reads_last_record(tmp_path):
    p = _params(Abins=6, Ebins=4)
    rng = np.random.default_rng(4)
    path = str(tmp_path / "snap.nc")
    last = None
    with capi.NcFile(p, path) as f:
        for k in range(4):
            last = (20.0 + k * 0.1, 1e-4 * (k + 2), _rand_psi(p, rng))
            f.append_snapshot(*last)
            f.sync()
            psi, r, dr = capi.snapshot_read_last(path, p)     # readable while open
            assert (r, dr) == last[:2]
            np.testing.assert_array_equal(psi, last[2])


@pytest.mark.parametrize("version", [1, 2])
def test_resume_from_a_file_written_by_another_library(tmp_path, version):
    """A Snapshot written by scipy (classic or 64-bit offset), with an extra
    variable and global attribute, is read back by the resume path."""
    p = _params(Abins=3, Ebins=5)
    rng = np.random.default_rng(5 + version)
    path = str(tmp_path / f"scipy_v{version}.nc")
    xs = [_rand_psi(p, rng) for _ in range(3)]
    with netcdf_file(path, "w", version=version) as nc:
        nc.history = "written by scipy"
        nc.createDimension("r", None)
        nc.createDimension("theta", 3)
        nc.createDimension("prtcl", 4)
        nc.createDimension("comp", 4)
        nc.createDimension("ebin", 5)
        extra = nc.createVariable("theta_bm", "i4", ("theta",))
        extra[:] = np.arange(3)
        psi = nc.createVariable("psi", "f8", ("r", "theta", "prtcl", "comp", "ebin"))
        r = nc.createVariable("r", "f8", ("r",))
        dr = nc.createVariable("dr", "f8", ("r",))
        for k, x in enumerate(xs):
            psi[k] = x
            r[k] = 30.0 + k
            dr[k] = 1e-5 * (k + 1)
    got, rr, dd = capi.snapshot_read_last(path, p)
    assert (rr, dd) == (32.0, 1e-5 * 3)
    np.testing.assert_array_equal(got, xs[-1])


def test_resume_rejects_a_mismatched_grid(tmp_path):
    p = _params(Abins=6, Ebins=4)
    path = str(tmp_path / "snap.nc")
    with capi.NcFile(p, path) as f:
        f.append_snapshot(10.0, 1e-3, np.zeros((6, 4, 4, 4)))
    with pytest.raises(capi.XfError) as e:
        capi.snapshot_read_last(path, _params(Abins=7, Ebins=4))
    assert e.value.code == -5
    with pytest.raises(capi.XfError):
        capi.snapshot_read_last(str(tmp_path / "missing.nc"), p)


def test_output_names():
    p = _params(Abins=2, Ebins=2, filePrefix="out/run1_")
    assert capi.output_name(p, False, 0, 0) == "out/run1_Snapshot0_0_.nc"        # IO_f.cpp:99
    assert capi.output_name(p, True, 3, 0) == "out/run1_Probability3_0_.nc"      # IO_f.cpp:140

"""CPU checks of the C-ABI library: it loads without a GPU, exports every
entry point include/*.h declares, and its config reader (Parser.cpp
keywords and unit scaling) agrees with the Python one."""
import os
import re

import numpy as np
import pytest

from conftest import ROOT
from xflat_amd import capi
from xflat_amd.config import config_text, parse_config_text, MEV, EV2, ERG_S


def _declared():
    names = set()
    for h in sorted(os.listdir(os.path.join(ROOT, "include"))):
        if h.endswith(".h"):
            txt = open(os.path.join(ROOT, "include", h)).read()
            names |= set(re.findall(r"^\s*(?:int|const char\*|double)\s+(xf_[a-z_0-9]+)\(", txt, re.M))
    return sorted(names)


def test_library_exports_every_declared_symbol():
    L = capi.load()
    names = _declared()
    assert len(names) >= 40
    missing = [n for n in names if not hasattr(L, n)]
    assert not missing, missing
    # the ctypes signatures cover the header too
    assert set(names) <= set(capi.SIGNATURES)


def test_no_device_is_reported_cleanly():
    if capi.device_count() > 0:
        pytest.skip("a GPU is visible")
    p = parse_config_text(config_text(Abins=8, Ebins=8))
    with pytest.raises(capi.XfError) as e:
        capi.Context(p)
    assert e.value.code == -7


@pytest.mark.parametrize("over", [dict(), dict(Abins=256, Pbins=64, Ebins=200, Rn=400),
                                  dict(dm2="+3E-3", nb0="1.63E+36", hNS=0.18, theta=0.6, eta_ve=3, Lve="1.5E+51")])
def test_cpp_and_python_config_readers_agree(over):
    txt = config_text(**over) + "# trailing comment\nunknownKey= 5\nt_step1= 300000.\n"
    a = capi.parse_config_text(txt).as_dict()
    b = parse_config_text(txt).as_dict()
    assert a == b


def test_unit_scaling_of_parser():
    p = capi.parse_config_text("E0= 1\nE1= 60\ndm2= -2.4E-3\nLve= 1E+51\nTve= 2.76\nnb0= 2\ntheta= 0.1\nPbins= 1\n")
    assert p.E0 == 1 * MEV and p.E1 == 60 * MEV
    assert p.dm2 == -2.4e-3 * EV2
    assert p.L[0] == 1e51 * ERG_S and p.T[0] == 2.76 * MEV
    assert p.nb0 == 2e15 and p.theta2 == 0.2
    assert p.module == 0


def test_configs_parse(tmp_path):
    for f in sorted(os.listdir(os.path.join(ROOT, "configs"))):
        p = capi.parse_config(os.path.join(ROOT, "configs", f))
        assert p.abins > 0 and p.ebins > 0 and p.flvs == 2, f

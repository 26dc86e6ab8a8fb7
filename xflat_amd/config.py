"""XFLAT config*.txt format: ``key= value`` lines, ``#`` comments.

Keywords and unit scaling follow Parser::readLines (src/Parser.cpp:25-127):
E0/E1 and the temperatures are scaled by MEV, nb0 by 1e15, dm2 by EV2, the
luminosities by ERG_S, and theta2 = 2*theta.  Unit constants are those of
include/global.h:43-46.  Unknown keys are ignored; unset keys stay zero.
"""
from __future__ import annotations

import ctypes

ERG_S = 1.053613095584080e+16   # erg/s -> 1/km^2
MEV = 5.067731161976211e+15     # MeV -> 1/km
MEV_2 = 3.893793036626876e-32   # MeV^-2 -> km^2
EV2 = 2.568189913006477e+19     # eV^2 -> 1/km^2

XF_MA, XF_MAA, XF_SA = 0, 1, 2
XF_FP64, XF_MIXED, XF_FP32 = 0, 1, 2


class XfParams(ctypes.Structure):
    """Mirror of ``xf_params`` in include/xflat.h (field order matters)."""
    _fields_ = [
        ("module", ctypes.c_int),
        ("abins", ctypes.c_int), ("pbins", ctypes.c_int), ("ebins", ctypes.c_int),
        ("flvs", ctypes.c_int), ("spoints", ctypes.c_int),
        ("E0", ctypes.c_double), ("E1", ctypes.c_double),
        ("R0", ctypes.c_double), ("Rn", ctypes.c_double),
        ("dr", ctypes.c_double), ("max_dr", ctypes.c_double),
        ("eps0", ctypes.c_double), ("ksi", ctypes.c_double), ("mu", ctypes.c_double),
        ("Ts", ctypes.c_int), ("Tn", ctypes.c_int),
        ("theta", ctypes.c_double), ("theta2", ctypes.c_double), ("dm2", ctypes.c_double),
        ("L", ctypes.c_double * 4), ("T", ctypes.c_double * 4), ("eta", ctypes.c_double * 4),
        ("has_matter", ctypes.c_int),
        ("Ye", ctypes.c_double), ("Rv", ctypes.c_double), ("Mns", ctypes.c_double),
        ("gs", ctypes.c_double), ("S", ctypes.c_double), ("hNS", ctypes.c_double),
        ("nb0", ctypes.c_double),
        ("dump_mode", ctypes.c_int), ("new_file_step", ctypes.c_int), ("sync_step", ctypes.c_int),
        ("r_step1", ctypes.c_double), ("t_step1", ctypes.c_int), ("itr_step1", ctypes.c_int),
        ("r_step2", ctypes.c_double), ("t_step2", ctypes.c_int), ("itr_step2", ctypes.c_int),
        ("start_beam", ctypes.c_int), ("end_beam", ctypes.c_int),
        ("multi_node_bench", ctypes.c_int), ("min_nodes", ctypes.c_int), ("max_nodes", ctypes.c_int),
        ("file_prefix", ctypes.c_char * 256),
    ]

    def as_dict(self):
        out = {}
        for name, _ in self._fields_:
            v = getattr(self, name)
            if isinstance(v, ctypes.Array):
                v = list(v)
            out[name] = v
        return out


# keyword -> (field, index or None, scale, kind); kind 'i' int, 'd' double, 's' str
_KEYS = {
    "filePrefix=": ("file_prefix", None, None, "s"),
    "dumpMode=": ("dump_mode", None, None, "i"),
    "newFile_step=": ("new_file_step", None, None, "i"),
    "sync_step=": ("sync_step", None, None, "i"),
    "r_step1=": ("r_step1", None, None, "d"),
    "t_step1=": ("t_step1", None, None, "i"),
    "itr_step1=": ("itr_step1", None, None, "i"),
    "r_step2=": ("r_step2", None, None, "d"),
    "t_step2=": ("t_step2", None, None, "i"),
    "itr_step2=": ("itr_step2", None, None, "i"),
    "start_beam=": ("start_beam", None, None, "i"),
    "end_beam=": ("end_beam", None, None, "i"),
    "multiNodeBench=": ("multi_node_bench", None, None, "i"),
    "minNodes=": ("min_nodes", None, None, "i"),
    "maxNodes=": ("max_nodes", None, None, "i"),
    "hasMatter=": ("has_matter", None, None, "i"),
    "eps0=": ("eps0", None, None, "d"),
    "ksi=": ("ksi", None, None, "d"),
    "mu=": ("mu", None, None, "d"),
    "Tn=": ("Tn", None, None, "i"),
    "Ts=": ("Ts", None, None, "i"),
    "R0=": ("R0", None, None, "d"),
    "Rn=": ("Rn", None, None, "d"),
    "dr=": ("dr", None, None, "d"),
    "max_dr=": ("max_dr", None, None, "d"),
    "E0=": ("E0", None, MEV, "d"),
    "E1=": ("E1", None, MEV, "d"),
    "SPoints=": ("spoints", None, None, "i"),
    "Pbins=": ("pbins", None, None, "i"),
    "Abins=": ("abins", None, None, "i"),
    "Ebins=": ("ebins", None, None, "i"),
    "Flvs=": ("flvs", None, None, "i"),
    "Ye=": ("Ye", None, None, "d"),
    "Rv=": ("Rv", None, None, "d"),
    "Mns=": ("Mns", None, None, "d"),
    "gs=": ("gs", None, None, "d"),
    "S=": ("S", None, None, "d"),
    "hNS=": ("hNS", None, None, "d"),
    "nb0=": ("nb0", None, 1.E+15, "d"),
    "theta=": ("theta", None, None, "d"),
    "dm2=": ("dm2", None, EV2, "d"),
    "Lve=": ("L", 0, ERG_S, "d"), "Lv_e=": ("L", 1, ERG_S, "d"),
    "Lvt=": ("L", 2, ERG_S, "d"), "Lv_t=": ("L", 3, ERG_S, "d"),
    "Tve=": ("T", 0, MEV, "d"), "Tv_e=": ("T", 1, MEV, "d"),
    "Tvt=": ("T", 2, MEV, "d"), "Tv_t=": ("T", 3, MEV, "d"),
    "eta_ve=": ("eta", 0, None, "d"), "eta_v_e=": ("eta", 1, None, "d"),
    "eta_vt=": ("eta", 2, None, "d"), "eta_v_t=": ("eta", 3, None, "d"),
}


def _c_int(tok: str) -> int:
    """istream >> int: leading integer part ('1E+51' -> 1, '.5' -> fails)."""
    s = tok.strip()
    i = 0
    if i < len(s) and s[i] in "+-":
        i += 1
    j = i
    while j < len(s) and s[j].isdigit():
        j += 1
    if j == i:
        raise ValueError(tok)
    return int(s[:j])


def _c_double(tok: str) -> float:
    """istream >> double: longest valid floating prefix ('.001' ok, '10.' ok)."""
    s = tok.strip()
    for end in range(len(s), 0, -1):
        try:
            return float(s[:end])
        except ValueError:
            continue
    raise ValueError(tok)


def n_beams(p) -> int:
    """Beams of the state: Abins x Pbins for MAA, Abins for MA and one
    trajectory for SA (beam_len = nbm_prtcls, src/Phy_SA.cpp:177)."""
    if p.module == XF_SA:
        return 1
    return p.abins * (p.pbins if p.module == XF_MAA else 1)


def parse_config_text(text: str, module: int | None = None) -> XfParams:
    """Parse config text like Parser::readLines (src/Parser.cpp:25-127).

    ``module`` selects MA/MAA/SA (a compile-time choice in the reference,
    global.h:5-19); by default it is MAA when Pbins > 1, else MA.
    """
    p = XfParams()
    for line in text.splitlines():
        toks = line.split()
        if not toks or toks[0].startswith("#"):
            continue
        key = toks[0]
        if key not in _KEYS or len(toks) < 2:
            continue
        field, idx, scale, kind = _KEYS[key]
        if kind == "s":
            setattr(p, field, toks[1].encode()[:255])
            continue
        try:
            v = _c_int(toks[1]) if kind == "i" else _c_double(toks[1])
        except ValueError:
            continue
        if scale is not None:
            v = v * scale
        if idx is None:
            setattr(p, field, v)
        else:
            getattr(p, field)[idx] = v
        if field == "theta":
            p.theta2 = p.theta * 2
    if module is None:
        module = XF_MAA if p.pbins > 1 else XF_MA
    p.module = module
    return p


# Physics of the bulb benchmark (every benchmark configuration): flavours
# with 1e51 erg/s each, Fermi-Dirac spectra with zero chemical potential at
# T = 2.76, 4.01, 6.26 MeV, delta m^2 = 2.4e-3 eV^2, theta_v = 0.1, and an
# electron density n_e = 1e32 cm^-3 (10 km / r)^3: nb0 = 0, Ye = 0.5,
# gs = 5.5, Mns = 1.4 and S = 58.2969 make the far term of Matt.cpp:31-36
# equal 2e32 (10/r)^3 baryons/cm^3.
BULB = dict(
    dumpMode=0, filePrefix="xflat", newFile_step=1000000, sync_step=1000000,
    r_step1=10, t_step1=300000, itr_step1=3000000,
    r_step2=0.5, t_step2=30000, itr_step2=3000000,
    start_beam=-1, end_beam=-1, multiNodeBench=0, minNodes=1, maxNodes=1,
    hasMatter=1, eps0="1.E-8", ksi=".94", mu=20,
    dm2="-2.4E-3", theta="0.1",
    Tn=10000000, Ts=30000000, R0=10, Rn=250, dr=".001", max_dr=".1", E0=1, E1=60,
    SPoints=1, Abins=200, Pbins=1, Ebins=100, Flvs=2,
    Ye=0.5, nb0=0, Rv="10.", Mns=1.4, gs=5.5, S=58.2969, hNS=0.18,
    Lve="1E+51", Lv_e="1E+51", Lvt="1E+51", Lv_t="1E+51",
    Tve=2.76, Tv_e=4.01, Tvt=6.26, Tv_t=6.26,
    eta_ve=0, eta_v_e=0, eta_vt=0, eta_v_t=0,
)


def config_text(**overrides) -> str:
    """Config text in the reference's format: BULB with overrides."""
    d = dict(BULB)
    d.update(overrides)
    lines = ["# XFLAT parameter file (key= value)"]
    for k, v in d.items():
        lines.append(f"{k}= {v}")
    return "\n".join(lines) + "\n"


def bulb_params(**overrides) -> XfParams:
    """Parsed XfParams for the bulb physics with keyword overrides.

    MAA is selected when Pbins > 1 (or pass module=XF_MA/XF_MAA/XF_SA)."""
    module = overrides.pop("module", None)
    return parse_config_text(config_text(**overrides), module=module)

"""MI355X-native multi-angle evolution step of XFLAT (HIP kernels behind a C-ABI).

The host side mirrors the reference's Phy module interface; see
include/xflat.h for the C entry points and xflat_amd.capi for the binding.
"""
from .config import (XfParams, parse_config_text as parse_config_text_py, config_text, bulb_params,  # noqa: F401
                     XF_MA, XF_MAA, XF_FP64, XF_MIXED, XF_FP32, MEV, ERG_S, EV2, MEV_2)
from .seed import seeded_psi, pure_psi  # noqa: F401

__version__ = "1.0.0"

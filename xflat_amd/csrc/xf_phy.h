// xf_phy.h -- the fine-grained Phy.h mirror (four FP64 beam sets on the device).
#pragma once
#include "xf_internal.h"

struct XfPhy;
void xf_phy_destroy(XfPhy* p);

// accessors of the context (xf_capi.cpp)
XfPhy*& xf_ctx_phy(xf_ctx* ctx);
const xf_params& xf_ctx_params(const xf_ctx* ctx);
const XfTables& xf_ctx_tables(const xf_ctx* ctx);

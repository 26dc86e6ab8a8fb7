/*
 * xf_ref_params.h -- shared by the reference-side bindings (Nmr_gpu.cpp,
 * Phy_gpu.cpp): the C-ABI's xf_params filled from the reference's own
 * configuration store, as Parser::readLines (src/Parser.cpp:25-127) left it in
 * Util and nbm, already in internal units; R0 and dr as fio::fillInitData set
 * them on resume (IO_f.cpp:197-250).  Compiled against the reference's
 * include/ (Util.h, NBeam.h).
 */
#ifndef XF_REF_PARAMS_H_
#define XF_REF_PARAMS_H_

#include <cstdio>
#include <cstdlib>
#include <cstring>

#include "Util.h"
#include "NBeam.h"
#include "xflat.h"

namespace xf_ref {

inline int module_of_build()
{
#if defined(MAA)
    return XF_MAA;
#elif defined(SA)
    return XF_SA;
#elif defined(MA)
    return XF_MA;
#else
    return XF_MAA;              // global.h's default module
#endif
}

inline void params_from_reference(xf_params* p)
{
    std::memset(p, 0, sizeof *p);
    p->module = module_of_build();
    p->abins = Util::Abins(); p->pbins = Util::Pbins(); p->ebins = Util::Ebins();
    p->flvs = Util::Flvs(); p->spoints = Util::SPoints();
    p->E0 = Util::E0(); p->E1 = Util::E1();
    p->R0 = Util::R0(); p->Rn = Util::Rn(); p->dr = Util::dr(); p->max_dr = Util::maxDr();
    p->eps0 = Util::eps0(); p->ksi = Util::ksi(); p->mu = Util::mu();
    p->Ts = Util::Ts(); p->Tn = Util::Tn();
    p->theta = Util::theta(); p->theta2 = Util::theta2(); p->dm2 = Util::dm2();
    for (int i = 0; i < 4; ++i) {
        p->L[i] = nbm::Lv(i);
        p->T[i] = nbm::Tv(i);
        p->eta[i] = nbm::eta(i);
    }
    p->has_matter = Util::hasMatter();
    p->Ye = Util::Ye(); p->Rv = Util::Rv(); p->Mns = Util::Mns(); p->gs = Util::gs(); p->S = Util::S();
    p->hNS = Util::hNS(); p->nb0 = Util::nb0();
    p->dump_mode = Util::dumpMode(); p->new_file_step = Util::newFilestep(); p->sync_step = Util::syncstep();
    p->r_step1 = Util::rstep1(); p->t_step1 = Util::tstep1(); p->itr_step1 = Util::itrstep1();
    p->r_step2 = Util::rstep2(); p->t_step2 = Util::tstep2(); p->itr_step2 = Util::itrstep2();
    p->start_beam = Util::startBeam(); p->end_beam = Util::endBeam();
    p->multi_node_bench = Util::multiNodeBench(); p->min_nodes = Util::minNodes(); p->max_nodes = Util::maxNodes();
    std::strncpy(p->file_prefix, Util::filePrefix().c_str(), sizeof p->file_prefix - 1);
}

// FP64 (the parity mode) unless XFLAT_GPU_PRECISION names another
inline int precision_from_env()
{
    const char* s = getenv("XFLAT_GPU_PRECISION");
    if (!s || !std::strcmp(s, "fp64")) return XF_FP64;
    if (!std::strcmp(s, "fp32")) return XF_FP32;
    if (!std::strcmp(s, "mixed")) return XF_MIXED;
    printf("xflat: XFLAT_GPU_PRECISION=%s: fp64, mixed or fp32\n", s);
    exit(-10);
}

// a failed C-ABI call ends the run with the reference's IO error code (IO_f.cpp:88)
inline void check(int rc, const char* what)
{
    if (rc != XF_OK) {
        printf("xflat: %s failed (%d): %s\n", what, rc, xf_last_error());
        exit(-10);
    }
}

}  // namespace xf_ref

#endif

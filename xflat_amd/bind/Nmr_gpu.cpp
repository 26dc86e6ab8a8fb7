/*
 * Nmr_gpu.cpp -- the reference's numerics module (include/Nmr.h: nmr::init,
 * nmr::freemem, nmr::evolutionLoop) over the xflat_amd C-ABI (include/xflat.h).
 *
 * A source file for the reference's tree: it is compiled instead of
 * src/Nmr.cpp, against the reference's own include/ and with its other
 * translation units unchanged (main, Parser, Util, NBGroup, NBeam, Fenergy,
 * Matt, the Phy module chosen by -DMA / -DMAA / -DSA, the IO module), and
 * linked with libxflat.so.  The Phy module still sets up the host NBeam array
 * (phy::initBeam, with fio::fillInitData on resume) and serves the IO
 * module's dimension queries; the radial loop of Nmr.cpp:89-262 runs on the
 * GPU in xf_run_steps, many attempted steps per call, accept/reject and dr
 * control included.
 *
 * Dumps (Nmr.cpp:238 calls fio::dumpToFile after every accepted step; the IO
 * module decides whether to write, IO_f.cpp:387-424).  The device stops after
 * every accepted step at which the IO module's radius or iteration rule for an
 * active stream (dumpMode bits 1 and 2: r_step1/itr_step1, r_step2/itr_step2),
 * or r == Rn, can fire; the state is copied back into the host array and
 * fio::dumpToFile(state, it, r, dr) is called with the attempt index and the
 * dr of that step, so the reference's IO module takes its own decision
 * unchanged.  Wall-clock rules (t_step1, t_step2) and the Tn budget are
 * evaluated between launches of at most XFLAT_GPU_CHUNK attempts (default
 * 2000); with dumpMode 0 the module is called at the end of every launch that
 * accepted a step (it only prints the radius).
 *
 * Precision: FP64 (the parity mode) unless XFLAT_GPU_PRECISION is "fp32" or
 * "mixed".  Errors print xf_last_error() and exit(-10), the reference's IO
 * error convention (IO_f.cpp:88).
 */
#include <algorithm>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <sys/time.h>
#include <vector>

#include "global.h"
#include "Util.h"
#include "Nmr.h"
#include "NBeam.h"
#include "NBGroup.h"
#include "Fio.h"
#include "Phy.h"

#include "xflat.h"
#include "xf_ref_params.h"

namespace nbgrp {

namespace nmr {

using namespace nbm;

int length;             // beams x particles of the host array (phy::beamLen)
NBeam* nubm_0;          // the state the IO module reads (initial state, then each dump)

namespace {

double wall_s()
{
    timeval t;
    gettimeofday(&t, NULL);
    return t.tv_sec + t.tv_usec * 1e-6;
}

// host NBeam array <-> psi[beam][prtcl][comp][ebin] (the C-ABI's layout)
void gather(std::vector<double>& psi)
{
    const size_t ne = nbm_ebins;
    psi.resize((size_t)length * nbm_cmpn * ne);
    for (int i = 0; i < length; ++i)
        for (int c = 0; c < nbm_cmpn; ++c) std::memcpy(&psi[((size_t)i * nbm_cmpn + c) * ne], nubm_0[i].psi(c), ne * sizeof(double));
}
void scatter(const std::vector<double>& psi)
{
    const size_t ne = nbm_ebins;
    for (int i = 0; i < length; ++i)
        for (int c = 0; c < nbm_cmpn; ++c) std::memcpy(nubm_0[i].psi(c), &psi[((size_t)i * nbm_cmpn + c) * ne], ne * sizeof(double));
}

// One dump stream of the IO module (dumpMode bit), as far as the device can
// decide it: the radius of its last write and its radius / iteration steps.
struct Stream {
    bool on;
    double last_r, rstep, tstep, last_t;
    int itrstep;
    // IO_f.cpp:402 / :413 for the state after an accepted step
    bool due(double r, double rn, long long it, double t) const
    {
        return on && (r == rn || r - last_r > rstep || (tstep > 0 && t - last_t > tstep) ||
                      (itrstep > 0 && it % itrstep == 0));
    }
};

}  // namespace

int init(int len)
{
    printf("Allocating memory for beams...\n");
    if (posix_memalign((void**)&nubm_0, ALIGN_LEN, len * sizeof(NBeam)) != 0) {
        printf("xflat: cannot allocate the beam array\n");
        exit(-10);
    }
    length = len;
    phy::initBeam(nubm_0);     // NBeam(prtcl) for every beam, fio::fillInitData on resume
    return 3;                  // the step's three points (r, r + dr/2, r + dr), Nmr.cpp:39
}

void freemem()
{
    nbm::freemem();
    phy::freeBeam(nubm_0);
    free(nubm_0);
}

int evolutionLoop() throw()
{
    xf_params p;
    xf_ref::params_from_reference(&p);
    xf_opts o;
    std::memset(&o, 0, sizeof o);
    o.precision = xf_ref::precision_from_env();
    o.renorm = -1;
    o.resident = -1;
    xf_ctx* ctx = NULL;
    xf_ref::check(xf_create(&p, &o, &ctx), "xf_create");
    if (xf_beam_len(ctx) != length) {
        printf("xflat: the Phy module holds %d beams, the whole grid is %d (one process drives the whole grid)\n",
               length, xf_beam_len(ctx));
        exit(-10);
    }
    std::vector<double> psi;
    gather(psi);
    double r = Util::R0(), dr = Util::dr();
    xf_ref::check(xf_state_upload(ctx, &psi[0], r, dr), "xf_state_upload");

    const char* ch = getenv("XFLAT_GPU_CHUNK");
    const long long chunk = ch ? std::max(1, atoi(ch)) : 2000;
    const double t0 = wall_s();
    Stream s1 = {(Util::dumpMode() & 1) != 0, r, Util::rstep1(), (double)Util::tstep1(), t0, Util::itrstep1()};
    Stream s2 = {(Util::dumpMode() & 2) != 0, r, Util::rstep2(), (double)Util::tstep2(), t0, Util::itrstep2()};
    long long counter_tot = 0, counter = 0;
    while (true) {
        // Nmr.cpp:130 (the benchmark mode runs for Tn seconds)
        const double t = wall_s();
        const bool go = Util::isBench() ? (t - t0 < Util::Tn())
                                        : (r < Util::Rn() && (long long)t - (long long)t0 < Util::Tn() && counter_tot < Util::Ts());
        if (!go) break;
        // stop after the next accepted step at which a stream's radius or
        // iteration rule can fire
        xf_stop st;
        std::memset(&st, 0, sizeof st);
        double next_r = 1e300;
        int itr = 0;
        for (Stream* s : {&s1, &s2}) {
            if (!s->on) continue;
            if (s->last_r + s->rstep < next_r) { next_r = s->last_r + s->rstep; st.dump_last_r = s->last_r; st.dump_rstep = s->rstep; }
            if (s->itrstep > 0) itr = itr ? (itr < s->itrstep ? itr : s->itrstep) : s->itrstep;
        }
        // the smallest iteration step stops the device; each stream re-checks its own
        if (s1.on && s2.on && s1.itrstep > 0 && s2.itrstep > 0 && (s1.itrstep % itr || s2.itrstep % itr)) itr = 1;
        st.dump_itrstep = itr;
        if ((s1.on || s2.on) && st.dump_rstep <= 0.) st.dump_rstep = 1e-300;
        const long long budget = Util::isBench() ? chunk : std::min(chunk, (long long)Util::Ts() - counter_tot);
        xf_run_info info;
        xf_ref::check(xf_run_steps(ctx, (int)budget, &st, NULL, &info), "xf_run_steps");
        counter_tot += info.steps;
        counter += info.accepted;
        r = info.r;
        dr = info.dr;
        const long long it = counter_tot - 1;       // attempt index of the last attempt
        const double tn = wall_s();
        const bool stopped = info.stop_reason == 2; // the last attempt was accepted and a rule can fire
        bool call = false;
        if (stopped) {
            for (Stream* s : {&s1, &s2})
                if (s->due(r, Util::Rn(), it, tn)) { call = true; s->last_r = r; s->last_t = tn; }
        } else if (info.accepted > 0) {
            // wall-clock rules and the no-IO progress line, at the launch's end
            for (Stream* s : {&s1, &s2})
                if (s->on && ((s->tstep > 0 && tn - s->last_t > s->tstep) || r == Util::Rn())) {
                    call = true;
                    s->last_r = r;
                    s->last_t = tn;
                }
            if (!Util::dumpMode()) call = true;
        }
        if (call) {
            double rr, dd;
            xf_ref::check(xf_state_download(ctx, &psi[0], &rr, &dd), "xf_state_download");
            scatter(psi);
            fio::dumpToFile(nubm_0, (int)(stopped ? it : counter_tot - 1), r, info.last_dr);
        }
        if (info.steps == 0) break;
    }
    // the final state back in the host array (the IO module's freemem and any
    // caller of the module see the evolved beams)
    {
        double rr, dd;
        xf_ref::check(xf_state_download(ctx, &psi[0], &rr, &dd), "xf_state_download");
        scatter(psi);
    }
    xf_ref::check(xf_destroy(ctx), "xf_destroy");
    printf("total time:%f\n", wall_s() - t0);
    printf("Total iterations: %lld\tCorrect iterations: %lld\t at radius(Km): %f\n", counter_tot, counter, r);
    return (int)counter_tot;
}

}  // namespace nmr

}  // namespace nbgrp

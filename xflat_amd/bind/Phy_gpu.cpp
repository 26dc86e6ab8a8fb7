/*
 * Phy_gpu.cpp -- the reference's physics-module interface (include/Phy.h:24-159,
 * namespace nbgrp::PHY_MODULE) over the xflat_amd C-ABI's Phy mirror
 * (include/xflat.h: xf_beams_*, xf_calc_*, xf_evolve*, xf_avg_beam).
 *
 * A source file for the reference's tree, compiled instead of src/Phy_MA.cpp,
 * src/Phy_MAA.cpp or src/Phy_SA.cpp with the same -DMA / -DMAA / -DSA switch
 * (include/global.h:5-19 names the namespace), against the reference's own
 * include/ and with every other translation unit unchanged, src/Nmr.cpp
 * included; linked with libxflat.so.  It is the fine-grained binding: every
 * Phy.h call of the radial loop (Nmr.cpp:153-191) runs as one or two device
 * kernels on device-resident beam sets, for step-by-step comparison.
 *
 * Host and device.  Nmr.cpp:61-65 hands the module its four NBeam arrays
 * through initBeam (nubm_0, nubm_1, nubm_3, nubm_4); each array's address is
 * bound to one device beam set, so the pointer swap of an accepted step
 * (Nmr.cpp:231-233) carries over by itself, and each hvv array of newHvv to
 * one device Hamiltonian slot.  Nmr.cpp reads host data between the module's
 * calls (the error loop, NBeam::calcErrHvv on the host arrays, Nmr.cpp:195-206,
 * and fio::dumpToFile), so every call copies its device outputs back into the
 * host arrays: correct for the unchanged Nmr.cpp, and slow by design.  The
 * host error loop also re-caches the energy sums of psi4 on the host objects
 * (NBeam.cpp:903-938); the device set of the in-place average computes its
 * own (xf_calc_esum) for the next step's calc_Hvv.
 *
 * Dimensions and ownership follow the reference's modules: beamLen, getDim,
 * getDimInfo, startDim, countDim, startBeamIdx, endBeamIdx and firstDimLen
 * describe the same snapshot layout (Phy_MA.cpp:72-136, Phy_MAA.cpp:70-138,
 * Phy_SA.cpp:51-116), newHvv returns a caller-owned host array.  One process
 * drives the whole angle grid (start_beam .. end_beam must cover it).
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <new>
#include <string>
#include <vector>

#include "global.h"
#include "Util.h"
#include "NBGroup.h"
#include "Nmr.h"
#include "NBeam.h"
#include "Fio.h"
#include "Phy.h"

#include "xflat.h"
#include "xf_ref_params.h"

namespace nbgrp {

namespace PHY_MODULE {

using namespace nbm;

namespace {

#if defined(MAA) || (!defined(MA) && !defined(SA))
const int Ndims = 6;                       // [r, theta, phi, prtcl, comp, ebin]
#elif defined(SA)
const int Ndims = 4;                       // [r, prtcl, comp, ebin]
#else
const int Ndims = 5;                       // [r, theta, prtcl, comp, ebin]
#endif

int start_beam, end_beam;                  // the angle bins of this process
int abins, pbins, tetbins, beam_len;
size_t start[Ndims], count[Ndims];

xf_ctx* ctx = NULL;
const NBeam* set_of[4];                    // host array -> device beam set
double* hvv_of[4];                         // host hvv array -> device Hamiltonian slot
std::vector<double> buf;                   // psi[beam][prtcl][comp][ebin]

int S(const NBeam* b)
{
    for (int k = 0; k < 4; ++k)
        if (set_of[k] == b) return k;
    printf("xflat: a beam array the module did not initialise\n");
    exit(-10);
}
int H(const double* h)
{
    for (int k = 0; k < 4; ++k)
        if (hvv_of[k] == h) return k;
    printf("xflat: an hvv array the module did not allocate\n");
    exit(-10);
}

// host NBeam array <-> device set
void push(const NBeam* b)
{
    const size_t ne = nbm_ebins;
    buf.resize((size_t)beam_len * nbm_cmpn * ne);
    for (int i = 0; i < beam_len; ++i)
        for (int c = 0; c < nbm_cmpn; ++c) std::memcpy(&buf[((size_t)i * nbm_cmpn + c) * ne], b[i].psi(c), ne * sizeof(double));
    xf_ref::check(xf_beams_upload(ctx, S(b), &buf[0]), "xf_beams_upload");
}
void pull(NBeam* b)
{
    const size_t ne = nbm_ebins;
    buf.resize((size_t)beam_len * nbm_cmpn * ne);
    xf_ref::check(xf_beams_download(ctx, S(b), &buf[0]), "xf_beams_download");
    for (int i = 0; i < beam_len; ++i)
        for (int c = 0; c < nbm_cmpn; ++c) std::memcpy(b[i].psi(c), &buf[((size_t)i * nbm_cmpn + c) * ne], ne * sizeof(double));
}

}  // namespace

int beamLen() { return beam_len; }
int getDim() { return Ndims; }

void getDimInfo(std::string str[])
{
    int k = 0;
    str[k++] = "r";
    if (Ndims != 4) str[k++] = "theta";
    if (Ndims == 6) str[k++] = "phi";
    str[k++] = "prtcl";
    str[k++] = "comp";
    str[k++] = "ebin";
}

size_t* startDim() { return start; }
size_t* countDim() { return count; }
int& startBeamIdx() { return start_beam; }
int& endBeamIdx() { return end_beam; }
int& firstDimLen() { return Util::Abins(); }

void init()
{
    if (Util::isBench()) {                 // the reference's benchmark grids (Phy_MA.cpp:235-241, Phy_MAA.cpp:249-256)
        abins = (Ndims == 6) ? 720 : (Ndims == 4 ? 1 : 1800);
        pbins = (Ndims == 6) ? 100 : 1;
        start_beam = 0;
        end_beam = abins - 1;
        Util::Abins() = abins;
        Util::Pbins() = pbins;
    } else {
        abins = Util::Abins();
        pbins = (Ndims == 6) ? Util::Pbins() : 1;
    }
    tetbins = (Ndims == 4) ? 1 : end_beam - start_beam + 1;
    if (Ndims != 4 && (start_beam != 0 || tetbins != abins)) {
        printf("xflat: beams %d..%d of %d: one process drives the whole angle grid\n", start_beam, end_beam, abins);
        exit(-10);
    }
    int k = 0;
    for (int d = 0; d < Ndims; ++d) start[d] = 0;
    count[k++] = 1;
    if (Ndims != 4) count[k++] = tetbins;
    if (Ndims == 6) count[k++] = pbins;
    count[k++] = nbm_prtcls;
    count[k++] = nbm_cmpn;
    count[k++] = nbm_ebins;
    beam_len = tetbins * pbins * nbm_prtcls;

    xf_params p;
    xf_ref::params_from_reference(&p);
    xf_opts o;
    std::memset(&o, 0, sizeof o);
    o.precision = XF_FP64;                 // the mirror's beam sets are FP64
    o.species = 4;
    o.renorm = 0;
    o.resident = -1;
    xf_ref::check(xf_create(&p, &o, &ctx), "xf_create");
    if (xf_beam_len(ctx) != beam_len) {
        printf("xflat: the device grid holds %d beams, the module %d\n", xf_beam_len(ctx), beam_len);
        exit(-10);
    }
    for (int s = 0; s < 4; ++s) {
        set_of[s] = NULL;
        hvv_of[s] = NULL;
    }
    nmr::init(beam_len);                   // calls initBeam on the four arrays (Nmr.cpp:61-65)
}

void freemem()
{
    nmr::freemem();
    if (ctx) xf_ref::check(xf_destroy(ctx), "xf_destroy");
    ctx = NULL;
}

void newHvv(double*& hvv)
{
    for (int k = 0; k < 4; ++k)
        if (!hvv_of[k]) {
            if (posix_memalign((void**)&hvv, ALIGN_LEN, (size_t)beam_len / nbm_prtcls * nbm_size * sizeof(double)) != 0) {
                printf("xflat: cannot allocate an hvv array\n");
                exit(-10);
            }
            hvv_of[k] = hvv;
            return;
        }
    printf("xflat: more than four hvv arrays\n");
    exit(-10);
}

void deleteHvv(double*& hvv)
{
    for (int k = 0; k < 4; ++k)
        if (hvv_of[k] == hvv) hvv_of[k] = NULL;
    free(hvv);
    hvv = NULL;
}

void initBeam(NBeam* beam)
{
    int k = -1;
    for (int s = 0; s < 4 && k < 0; ++s)
        if (set_of[s] == beam || set_of[s] == NULL) k = s;
    if (k < 0) {
        printf("xflat: more than four beam arrays\n");
        exit(-10);
    }
    set_of[k] = beam;
    for (int i = 0; i < beam_len / nbm_prtcls; ++i)
        for (int j = 0; j < nbm_prtcls; ++j) new (&beam[i * nbm_prtcls + j]) NBeam(j);   // NBeam(int), NBeam.cpp:236-277
#if !defined(MAA) && (defined(MA) || defined(SA))
    // resume (MA and SA; the reference's MAA module never reads a data file)
    if (Util::inVals() != NULL && !Util::isBench()) fio::fillInitData(beam);
#endif
    push(beam);
    xf_ref::check(xf_calc_esum(ctx, k), "xf_calc_esum");   // the cached energy sums (Nmr.cpp:117-121)
}

void freeBeam(NBeam* beam)
{
    for (int i = 0; i < beam_len; ++i) beam[i].~NBeam();
    for (int s = 0; s < 4; ++s)
        if (set_of[s] == beam) set_of[s] = NULL;
}

void calcAngleBins(const double r, const int step_num) throw()
{
    xf_ref::check(xf_calc_angle_bins(ctx, r, step_num), "xf_calc_angle_bins");
}

void calcDeltaLs(const double dr, const int cur_pnt, const int s_pnt, const int e_pnt) throw()
{
    xf_ref::check(xf_calc_delta_ls(ctx, dr, cur_pnt, s_pnt, e_pnt), "xf_calc_delta_ls");
}

void calc_Hvv(const NBeam* REST beam, const int step_num, double* REST hvv) throw()
{
    const int h = H(hvv);
    xf_ref::check(xf_calc_hvv(ctx, S(beam), step_num, h), "xf_calc_hvv");
    xf_ref::check(xf_get_hvv(ctx, h, hvv), "xf_get_hvv");            // the host copy the caller owns
}

void evolve(const NBeam* REST ibeam, const int pnt, const double* REST hvv, const double hmatt, NBeam* REST obeam) throw()
{
    xf_ref::check(xf_evolve(ctx, S(ibeam), pnt, H(hvv), hmatt, S(obeam)), "xf_evolve");
    pull(obeam);
}

void evolveHvv(const NBeam* REST ibeam, const int pnt, const double* REST hvv, const double hmatt, NBeam* REST obeam) throw()
{
    xf_ref::check(xf_evolve_hvv(ctx, S(ibeam), pnt, H(hvv), hmatt, S(obeam)), "xf_evolve_hvv");
    pull(obeam);
}

void evolveAvg(const NBeam* REST ibeam, const int pnt, const double* REST hvv, const double hmatt, NBeam* REST obeam,
               NBeam* REST obeamAvg) throw()
{
    xf_ref::check(xf_evolve_avg(ctx, S(ibeam), pnt, H(hvv), hmatt, S(obeam), S(obeamAvg)), "xf_evolve_avg");
    pull(obeam);
    pull(obeamAvg);
}

// obeam = U ibeam with its energy sums, then obeamAvg <- (obeam + obeamAvg) / 2
// (NBeam::evolveBinsHvvAvg, NBeam.cpp:516-586)
void evolveHvvAvg(const NBeam* REST ibeam, const int pnt, const double* REST hvv, const double hmatt, NBeam* REST obeam,
                  NBeam* REST obeamAvg) throw()
{
    xf_ref::check(xf_evolve_hvv(ctx, S(ibeam), pnt, H(hvv), hmatt, S(obeam)), "xf_evolve_hvv");
    xf_ref::check(xf_avg_beam(ctx, S(obeam), S(obeamAvg)), "xf_avg_beam");
    pull(obeam);
    pull(obeamAvg);
}

// evolveAvg, then each obeamErr object's error sum_e fv |err - avg|^2
// (NBeam::evolveBinsAvgErr, NBeam.cpp:395-460), formed on the host copies
void evolveAvgErr(const NBeam* REST ibeam, const int pnt, const double* REST hvv, const double hmatt, NBeam* REST obeam,
                  NBeam* REST obeamAvg, NBeam* REST obeamErr) throw()
{
    evolveAvg(ibeam, pnt, hvv, hmatt, obeam, obeamAvg);
    pull(obeamErr);
    for (int i = 0; i < beam_len; ++i) obeamErr[i].getErr() = obeamAvg[i].calcErr(obeamErr[i]);
}

void evolveAvg(const int pnt, const double* REST hvv, const double hmatt, NBeam* REST iobeam, NBeam* REST obeamAvg) throw()
{
    xf_ref::check(xf_evolve_avg_inplace(ctx, pnt, H(hvv), hmatt, S(iobeam), S(obeamAvg)), "xf_evolve_avg_inplace");
    // the energy sums of the average, which the host error loop re-caches on
    // its own objects and the next step's calc_Hvv reads from the device
    xf_ref::check(xf_calc_esum(ctx, S(obeamAvg)), "xf_calc_esum");
    pull(iobeam);
    pull(obeamAvg);
}

void avgBeam(const NBeam* REST ibeam, NBeam* REST obeam) throw()
{
    xf_ref::check(xf_avg_beam(ctx, S(ibeam), S(obeam)), "xf_avg_beam");
    pull(obeam);
}

}  // namespace PHY_MODULE

}  // namespace nbgrp

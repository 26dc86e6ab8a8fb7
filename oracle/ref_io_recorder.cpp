/*
 * ref_io_recorder.cpp -- TEST INFRASTRUCTURE. An IO plugin for the
 * reference's own Fio.h plugin interface (include/Fio.h:16-40, selected by
 * IO_MODULE = io_f, include/global.h:22-28), used only to build the
 * reference's evolution path from its unmodified sources into oracle/_ref/
 * for golden vectors and the CPU baseline.
 *
 * The reference's NetCDF plugin (src/IO_f.cpp) needs libnetcdf, which this
 * image lacks, so it is not compiled; this file implements the same four
 * functions with raw binary files instead:
 *   fillInitData  -- resume semantics of src/IO_f.cpp:197-250: sets R0 and
 *                    dr and the wave functions from $XFLAT_REF_INIT
 *                    (int64 nbeams, int64 ncomp, int64 nebins, double r,
 *                    double dr, double psi[nbeams][ncomp][nebins]).
 *   dumpToFile    -- called by Nmr.cpp:238 after every accepted step; appends
 *                    (int64 it, double r, double dr, psi[...] when the record
 *                    index is a multiple of $XFLAT_REF_PSI_EVERY, else
 *                    nothing) to $XFLAT_REF_OUT.
 */
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <cstdint>
#include <vector>

#include "global.h"
#include "Util.h"
#include "NBeam.h"
#include "Phy.h"
#include "Fio.h"

namespace nbgrp {
namespace io_f {

static FILE* g_out = 0;
static long g_psi_every = 1;
static long g_rec = 0;

void init(int)
{
    const char* path = getenv("XFLAT_REF_OUT");
    const char* ev = getenv("XFLAT_REF_PSI_EVERY");
    g_psi_every = ev ? atol(ev) : 1;
    g_rec = 0;
    if (path) {
        g_out = fopen(path, "wb");
        if (!g_out) { fprintf(stderr, "recorder: cannot open %s\n", path); exit(-10); }
        int64_t hdr[4] = { (int64_t)phy::beamLen(), (int64_t)nbm::nbm_cmpn,
                           (int64_t)nbm::nbm_ebins, (int64_t)g_psi_every };
        fwrite(hdr, sizeof(int64_t), 4, g_out);
    }
}

void freemem()
{
    if (g_out) { fclose(g_out); g_out = 0; }
}

void fillInitData(nbm::NBeam* nubeam)
{
    const char* path = getenv("XFLAT_REF_INIT");
    if (!path) path = Util::inVals();
    FILE* f = fopen(path, "rb");
    if (!f) { fprintf(stderr, "recorder: cannot open %s\n", path); exit(-10); }
    int64_t hdr[3];
    double rdr[2];
    if (fread(hdr, sizeof(int64_t), 3, f) != 3 || fread(rdr, sizeof(double), 2, f) != 2) exit(-10);
    if (hdr[0] != phy::beamLen() || hdr[1] != nbm::nbm_cmpn || hdr[2] != nbm::nbm_ebins) {
        fprintf(stderr, "recorder: init file shape mismatch\n");
        exit(-10);
    }
    Util::R0() = rdr[0];
    Util::dr() = rdr[1];
    const size_t ne = nbm::nbm_ebins;
    std::vector<double> buf(ne);
    for (int i = 0; i < phy::beamLen(); ++i)
        for (int c = 0; c < nbm::nbm_cmpn; ++c) {
            if (fread(&buf[0], sizeof(double), ne, f) != ne) exit(-10);
            memcpy(nubeam[i].psi(c), &buf[0], ne * sizeof(double));
        }
    fclose(f);
}

void dumpToFile(const nbm::NBeam* nubeam, const int it, const double r, const double dr, int, int) throw()
{
    if (!g_out) return;
    int64_t itv = it;
    int64_t has_psi = (g_psi_every > 0 && (g_rec % g_psi_every) == 0) ? 1 : 0;
    fwrite(&itv, sizeof(int64_t), 1, g_out);
    fwrite(&r, sizeof(double), 1, g_out);
    fwrite(&dr, sizeof(double), 1, g_out);
    fwrite(&has_psi, sizeof(int64_t), 1, g_out);
    if (has_psi) {
        const size_t ne = nbm::nbm_ebins;
        for (int i = 0; i < phy::beamLen(); ++i)
            for (int c = 0; c < nbm::nbm_cmpn; ++c)
                fwrite(nubeam[i].psi(c), sizeof(double), ne, g_out);
    }
    ++g_rec;
}

} // namespace io_f
} // namespace nbgrp

// xf_setup.cpp -- host-side set-up of the evolution step: config reader,
// spectra, H_vac, angle grids, coefficients, matter profile, error strings.
// All FP64 with the host libm, like the reference's one-time init.
#include <cmath>
#include <cstdarg>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <fstream>
#include <sstream>
#include <string>

#include "xf_internal.h"

static thread_local std::string g_err;

int xf_fail(int code, const char* fmt, ...) {
    char buf[1024];
    va_list ap;
    va_start(ap, fmt);
    vsnprintf(buf, sizeof buf, fmt, ap);
    va_end(ap);
    g_err = buf;
    return code;
}

extern "C" const char* xf_last_error(void) { return g_err.c_str(); }
extern "C" int xf_version(void) { return 100; }

// ---- config reader: Parser::readLines, src/Parser.cpp:25-127 ---------------
// Each line is read as "key= value" with the same keywords and the same unit
// scaling applied at parse time.  Lines whose first token is empty or starts
// with '#' are skipped; unknown keys are ignored.
static void parse_stream(std::istream& fs, xf_params* p) {
    using namespace xfc;
    std::string line;
    while (std::getline(fs, line)) {
        std::stringstream ss;
        ss << line;
        std::string k;
        ss >> k;
        if (k.empty() || k[0] == '#') continue;
        if (k == "filePrefix=") {
            std::string s;
            ss >> s;
            std::snprintf(p->file_prefix, sizeof p->file_prefix, "%s", s.c_str());
        }
        else if (k == "dumpMode=")       ss >> p->dump_mode;
        else if (k == "newFile_step=")   ss >> p->new_file_step;
        else if (k == "sync_step=")      ss >> p->sync_step;
        else if (k == "r_step1=")        ss >> p->r_step1;
        else if (k == "t_step1=")        ss >> p->t_step1;
        else if (k == "itr_step1=")      ss >> p->itr_step1;
        else if (k == "r_step2=")        ss >> p->r_step2;
        else if (k == "t_step2=")        ss >> p->t_step2;
        else if (k == "itr_step2=")      ss >> p->itr_step2;
        else if (k == "start_beam=")     ss >> p->start_beam;
        else if (k == "end_beam=")       ss >> p->end_beam;
        else if (k == "multiNodeBench=") ss >> p->multi_node_bench;
        else if (k == "minNodes=")       ss >> p->min_nodes;
        else if (k == "maxNodes=")       ss >> p->max_nodes;
        else if (k == "hasMatter=")      ss >> p->has_matter;
        else if (k == "eps0=")           ss >> p->eps0;
        else if (k == "ksi=")            ss >> p->ksi;
        else if (k == "mu=")             ss >> p->mu;
        else if (k == "Tn=")             ss >> p->Tn;
        else if (k == "Ts=")             ss >> p->Ts;
        else if (k == "R0=")             ss >> p->R0;
        else if (k == "Rn=")             ss >> p->Rn;
        else if (k == "dr=")             ss >> p->dr;
        else if (k == "max_dr=")         ss >> p->max_dr;
        else if (k == "E0=")           { ss >> p->E0; p->E0 *= MEV; }
        else if (k == "E1=")           { ss >> p->E1; p->E1 *= MEV; }
        else if (k == "SPoints=")        ss >> p->spoints;
        else if (k == "Pbins=")          ss >> p->pbins;
        else if (k == "Abins=")          ss >> p->abins;
        else if (k == "Ebins=")          ss >> p->ebins;
        else if (k == "Flvs=")           ss >> p->flvs;
        else if (k == "Ye=")             ss >> p->Ye;
        else if (k == "Rv=")             ss >> p->Rv;
        else if (k == "Mns=")            ss >> p->Mns;
        else if (k == "gs=")             ss >> p->gs;
        else if (k == "S=")              ss >> p->S;
        else if (k == "hNS=")            ss >> p->hNS;
        else if (k == "nb0=")          { ss >> p->nb0; p->nb0 *= 1.E+15; }
        else if (k == "theta=")        { ss >> p->theta; p->theta2 = p->theta * 2; }
        else if (k == "dm2=")          { ss >> p->dm2; p->dm2 *= EV2; }
        else if (k == "Lve=")          { ss >> p->L[0]; p->L[0] *= ERG_S; }
        else if (k == "Lv_e=")         { ss >> p->L[1]; p->L[1] *= ERG_S; }
        else if (k == "Lvt=")          { ss >> p->L[2]; p->L[2] *= ERG_S; }
        else if (k == "Lv_t=")         { ss >> p->L[3]; p->L[3] *= ERG_S; }
        else if (k == "Tve=")          { ss >> p->T[0]; p->T[0] *= MEV; }
        else if (k == "Tv_e=")         { ss >> p->T[1]; p->T[1] *= MEV; }
        else if (k == "Tvt=")          { ss >> p->T[2]; p->T[2] *= MEV; }
        else if (k == "Tv_t=")         { ss >> p->T[3]; p->T[3] *= MEV; }
        else if (k == "eta_ve=")         ss >> p->eta[0];
        else if (k == "eta_v_e=")        ss >> p->eta[1];
        else if (k == "eta_vt=")         ss >> p->eta[2];
        else if (k == "eta_v_t=")        ss >> p->eta[3];
    }
    // The reference picks the geometry at compile time (global.h:5-19);
    // here a config with Pbins > 1 selects MAA, otherwise MA.
    p->module = (p->pbins > 1) ? XF_MAA : XF_MA;
}

extern "C" int xf_parse_config(const char* path, xf_params* out) {
    if (!path || !out) return xf_fail(XF_EINVAL, "xf_parse_config: null argument");
    std::ifstream fs(path);
    if (!fs.is_open()) return xf_fail(XF_EIO, "cannot open the input file %s", path);
    std::memset(out, 0, sizeof *out);
    parse_stream(fs, out);
    return XF_OK;
}

extern "C" int xf_parse_config_text(const char* text, xf_params* out) {
    if (!text || !out) return xf_fail(XF_EINVAL, "xf_parse_config_text: null argument");
    std::memset(out, 0, sizeof *out);
    std::istringstream fs{std::string(text)};
    parse_stream(fs, out);
    return XF_OK;
}

// ---- spectra: Fenergy.cpp:16-23 (fq), :25-41 (avgEnergy) -----------------
static double fq(double q, double Tv, double eta_v) {
    return (1. / 18.9686) * (1. / (Tv * Tv * Tv)) * ((q * q) / (std::exp(q / Tv - eta_v) + 1.));
}
static double avg_energy(const xf_params* p, int s) {
    const int Eb = 800;
    double de = (p->E1 - p->E0) / Eb;
    double frst = p->E0 + de * .5;
    double avg = 0., Eavg = 0.;
    for (int e = 0; e < Eb; ++e) {
        double q = frst + e * de;
        double fe = fq(q, p->T[s], p->eta[s]);
        avg += fe;
        Eavg += fe * q;
    }
    return Eavg / avg;
}

int xf_build_tables(const xf_params* p, XfTables* t) {
    using namespace xfc;
    if (p->flvs != 2) return xf_fail(XF_EINVAL, "Flvs= %d: only two flavours are supported", p->flvs);
    if (p->module != XF_MA && p->module != XF_MAA && p->module != XF_SA)
        return xf_fail(XF_EINVAL, "unknown module %d", p->module);
    if ((p->module != XF_SA && p->abins <= 0) || p->ebins <= 0) return xf_fail(XF_EINVAL, "Abins=/Ebins= must be positive");
    if (p->module == XF_MAA && p->pbins <= 0) return xf_fail(XF_EINVAL, "Pbins= must be positive for MAA");
    if (p->Rv <= 0) return xf_fail(XF_EINVAL, "Rv= must be positive");
    t->module = p->module;
    // SA evolves one trajectory: beam_len = nbm_prtcls (Phy_SA.cpp:177)
    t->nth = (p->module == XF_SA) ? 1 : p->abins;
    t->nph = (p->module == XF_MAA) ? p->pbins : 1;
    t->ne = p->ebins;
    t->nbeam = t->nth * t->nph;
    const int ne = t->ne;

    // nbm::init (NBeam.cpp:138-209)
    t->dE = (p->E1 - p->E0) / ne;
    for (int s = 0; s < NPR; ++s) t->Ev[s] = avg_energy(p, s);
    t->fv.assign(4 * ne, 0.);
    t->hvac.assign(2 * ne, 0.);
    double sum[NPR] = {0., 0., 0., 0.};
    double frst = p->E0 + t->dE * .5;
    for (int e = 0; e < ne; ++e) {
        double q = frst + e * t->dE;
        for (int s = 0; s < NPR; ++s) t->fv[s * ne + e] = fq(q, p->T[s], p->eta[s]);
        for (int s = 0; s < NPR; ++s) sum[s] += t->fv[s * ne + e];
        double delta = .5 * p->dm2 / q;
        double theta = p->theta2;
        t->hvac[2 * e] = -std::cos(theta) * delta * .5;
        t->hvac[2 * e + 1] = std::sin(theta) * delta * .5;
    }
    for (int s = 0; s < NPR; ++s) sum[s] = 1. / sum[s];
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < NPR; ++s) t->fv[s * ne + e] *= sum[s];

    // angle grid uniform in cos^2 at R_nu, by repeated subtraction
    // (Phy_MA.cpp:141-164, Phy_MAA.cpp:159-174)
    t->cos2R.assign(t->nth, 0.);
    t->dcos2R.assign(t->nth, 0.);
    t->sinR.assign(t->nth, 0.);
    double dct = (1. - 0.) / t->nth;
    t->cos2R[0] = 1. - dct * .5;
    t->sinR[0] = std::sqrt(1. - t->cos2R[0]);
    t->dcos2R[0] = dct;
    double ct = t->cos2R[0] - dct;
    for (int i = 1; i < t->nth; ++i, ct -= dct) {
        t->cos2R[i] = ct;
        t->sinR[i] = std::sqrt(1. - t->cos2R[i]);
        t->dcos2R[i] = dct;
    }
    t->cosphi.assign(t->nph, 1.);
    t->sinphi.assign(t->nph, 0.);
    t->dPhi = 0.;
    if (p->module == XF_MAA) {
        // phi grid (Phy_MAA.cpp:177-185) and coefficient (:297-298)
        double dphi = (PI * 2.) / t->nph;
        t->dPhi = dphi;
        double phi = dphi * .5;
        for (int q = 0; q < t->nph; ++q, phi += dphi) {
            t->sinphi[q] = std::sin(phi);
            t->cosphi[q] = std::cos(phi);
        }
        double hvv_coeff = (SQRT2 * GF) / (4. * PI * PI * p->Rv * p->Rv);
        t->integ_coef = hvv_coeff * t->dPhi;
    } else {
        // Phy_MA.cpp:271-272; SA: the same coefficient (Phy_SA.cpp:166-168),
        // times D(r) per step point
        t->integ_coef = (SQRT2 * GF) / (2. * PI * p->Rv * p->Rv);
    }
    for (int s = 0; s < NPR; ++s) t->L_E[s] = p->L[s] / t->Ev[s];   // Phy_MA.cpp:280-282

    // matt::init (Matt.cpp:22-29)
    t->mat_coeff = SQRT2 * GF * p->Ye;
    double M = p->Mns / 1.4;
    double sS = 100. / p->S;
    t->nbfar_coeff = 4.2E+45 * p->gs * (M * M * M) * (sS * sS * sS * sS);
    return XF_OK;
}

// matt::getHm, Matt.cpp:31-64
double xf_host_hm(const xf_params* p, const XfTables* t, double r) {
    if (!p->has_matter) return 0;
    double R = 10. / r;
    double far = t->nbfar_coeff * (R * R * R);
    double near = p->nb0 * std::exp((p->Rv - r) / p->hNS);
    return t->mat_coeff * (far + near) * .5;
}

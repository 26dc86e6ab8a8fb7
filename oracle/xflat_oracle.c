/*
 * xflat_oracle.c -- CPU FP64 restatement of XFLAT's multi-angle evolution
 * step.  TEST INFRASTRUCTURE ONLY: it is the checker that the HIP path is
 * compared against, and the "port" CPU baseline.  Nothing in the product
 * library links or calls it.
 *
 * It restates, in plain C with the reference's loop nests (OpenMP over
 * angles, `omp simd` over energy bins), the FP64 algorithm of
 *   src/NBeam.cpp   (spectra, H_vac, U, density, evolveBins*, calcESum, calcErrHvv)
 *   src/Phy_MA.cpp  (cos^2 grid, angle bins, delta-l, 2-moment H_vv)
 *   src/Phy_MAA.cpp (phi grid, trapezoidal delta-l, 4-moment H_vv)
 *   src/Phy_SA.cpp  (one trajectory: H_vv = integ_coef * D(r) * source, dl = dr)
 *   src/Nmr.cpp     (the adaptive step sequence S1..S11, error, accept, dr)
 *   src/Matt.cpp, src/Fenergy.cpp, include/Util.h, include/global.h
 * Every function names the reference lines it follows.
 *
 * State layout: psi[set][((beam*4 + prtcl)*4 + comp)*NE + e] -- the same
 * ordering as the reference's NetCDF snapshot buffer (src/IO_f.cpp:280-288),
 * beam = theta*Pbins + phi.
 */
#include <math.h>
#include <stdlib.h>
#include <string.h>
#include <stdint.h>
#include <sys/time.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#include "../include/xflat.h"

/* ---- constants: include/Util.h:15-22 and include/global.h:43-46 --------- */
static const double O_MEV_2 = 3.893793036626876e-32;
static const double O_SQRT2 = 1.41421356237;
static const double O_PI    = 3.14159265359;
#define O_GF (1.16637E-11 * 3.893793036626876e-32)   /* Util::GF */

#define NPR 4   /* nbm_prtcls for Flvs = 2 (NBeam.cpp:149-151) */
#define NSZ 4   /* nbm_size (NBeam.h:17)                       */

typedef struct oracle {
    xf_params p;
    int module, nth, nph, ne, nbeam, nnb;     /* nnb = nbeam * NPR NBeam objects */
    double dE;
    double *fq[NPR];                          /* normalised spectra fq_*           */
    double *hvac;                             /* 2*NE                              */
    double Ev[NPR], L_E[NPR];
    double *cos2R, *dcos2R, *sinR;            /* per theta at R_nu                 */
    double *sinphi, *cosphi, dPhi;            /* MAA                               */
    double hvv_coeff, integ_coef;
    double mat_coeff, nbfar_coeff;
    double *cos_t, *dcos_t, *sin_t, *dl;      /* [3][nth]                          */
    double *psi[4];                           /* 4 beam sets                       */
    double *hsum[4];                          /* [nnb][4] cached energy sums       */
    double *hvv[4];                           /* [nbeam][4]                        */
    /* stepper state */
    int cur[4];                               /* indices of nubm_0,1,3,4           */
    double r, dr;
    int64_t counter_tot, counter;
    int esum_valid;
} oracle_t;

typedef struct xo_rec { double r, dr, err; int32_t accepted, pad; } xo_rec;

static double* xalloc(size_t n) {
    void* p = NULL;
    if (posix_memalign(&p, 64, (n ? n : 1) * sizeof(double))) return NULL;
    memset(p, 0, (n ? n : 1) * sizeof(double));
    return (double*)p;
}

/* ---- spectra: src/Fenergy.cpp:16-23 (fq), :25-41 (avgEnergy), :44-47 ---- */
static double o_fq(double q, double Tv, double eta_v) {
    return (1. / 18.9686) * (1. / (Tv*Tv*Tv)) * ( (q*q) / (exp(q/Tv - eta_v)+1.) );
}
static double o_avg_energy(const xf_params* p, int s) {
    const int Eb = 800;
    double de = (p->E1 - p->E0) / Eb;
    double frst = p->E0 + de*.5;
    double avg = 0., Eavg = 0.;
    for (int e = 0; e < Eb; ++e) {
        double q = frst + e * de;
        double fe = o_fq(q, p->T[s], p->eta[s]);
        avg += fe;
        Eavg += fe * q;
    }
    return Eavg / avg;
}

/* ---- matter: src/Matt.cpp:22-29 (init), :31-46 (nb), :58-64 (getHm) ----- */
double oracle_get_hm(const oracle_t* o, double radius) {
    if (!o->p.has_matter) return 0;
    double R = 10. / radius;
    double far = o->nbfar_coeff * (R*R*R);
    double near = o->p.nb0 * exp((o->p.Rv - radius) / o->p.hNS);
    return o->mat_coeff * (far + near) * .5;
}

/* ---- set-up: nbm::init NBeam.cpp:138-209, phy_ma::init Phy_MA.cpp:233-284,
 *      init_cos2R Phy_MA.cpp:141-164, phy_maa::init/init_sincos2_R
 *      Phy_MAA.cpp:144-186,247-304, matt::init Matt.cpp:22-29 ------------- */
oracle_t* oracle_create(const xf_params* p) {
    if (p->flvs != 2 || p->ebins <= 0 || (p->module != XF_SA && p->abins <= 0)) return NULL;
    oracle_t* o = (oracle_t*)calloc(1, sizeof(oracle_t));
    o->p = *p;
    o->module = p->module;
    /* SA: one trajectory, beam_len = nbm_prtcls (Phy_SA.cpp:177) */
    o->nth = (p->module == XF_SA) ? 1 : p->abins;
    o->nph = (p->module == XF_MAA) ? p->pbins : 1;
    if (o->nph <= 0) { free(o); return NULL; }
    o->ne = p->ebins;
    o->nbeam = o->nth * o->nph;
    o->nnb = o->nbeam * NPR;
    int ne = o->ne;

    /* fed::init + avgEnergy (NBeam.cpp:141-147) */
    o->dE = (p->E1 - p->E0) / ne;
    for (int s = 0; s < NPR; ++s) o->Ev[s] = o_avg_energy(p, s);
    for (int s = 0; s < NPR; ++s) o->fq[s] = xalloc(ne);
    o->hvac = xalloc(2 * ne);
    double sum[NPR] = {0., 0., 0., 0.};
    double frst = p->E0 + o->dE*.5;
    for (int e = 0; e < ne; ++e) {
        double q = frst + e * o->dE;
        for (int s = 0; s < NPR; ++s) { o->fq[s][e] = o_fq(q, p->T[s], p->eta[s]); }
        for (int s = 0; s < NPR; ++s) sum[s] += o->fq[s][e];
        double delta = .5 * p->dm2 / q;
        double theta = p->theta2;
        o->hvac[2*e]   = -cos(theta) * delta * .5;
        o->hvac[2*e+1] =  sin(theta) * delta * .5;
    }
    for (int s = 0; s < NPR; ++s) sum[s] = 1. / sum[s];
    for (int e = 0; e < ne; ++e)
        for (int s = 0; s < NPR; ++s) o->fq[s][e] *= sum[s];

    /* angle grid uniform in cos^2 at R_nu, built by repeated subtraction */
    o->cos2R = xalloc(o->nth); o->dcos2R = xalloc(o->nth); o->sinR = xalloc(o->nth);
    double dct = (1. - 0.) / o->nth;
    o->cos2R[0] = 1. - dct*.5;
    o->sinR[0] = sqrt(1. - o->cos2R[0]);
    o->dcos2R[0] = dct;
    double ct = o->cos2R[0] - dct;
    for (int i = 1; i < o->nth; ++i, ct -= dct) {
        o->cos2R[i] = ct;
        o->sinR[i] = sqrt(1. - o->cos2R[i]);
        o->dcos2R[i] = dct;
    }
    if (o->module == XF_MAA) {
        o->sinphi = xalloc(o->nph); o->cosphi = xalloc(o->nph);
        double dphi = (O_PI*2.) / o->nph;
        o->dPhi = dphi;
        double phi = dphi*.5;
        for (int q = 0; q < o->nph; ++q, phi += dphi) {
            o->sinphi[q] = sin(phi);
            o->cosphi[q] = cos(phi);
        }
        o->hvv_coeff = (O_SQRT2 * O_GF) / (4. * O_PI * O_PI * p->Rv * p->Rv);
        o->integ_coef = o->hvv_coeff * o->dPhi;
    } else {
        /* Phy_MA.cpp:271-272; SA the same (Phy_SA.cpp:166-168) */
        o->hvv_coeff = (O_SQRT2 * O_GF) / (2. * O_PI * p->Rv * p->Rv);
        o->integ_coef = o->hvv_coeff;
    }
    for (int s = 0; s < NPR; ++s) o->L_E[s] = p->L[s] / o->Ev[s];

    o->mat_coeff = O_SQRT2 * O_GF * p->Ye;
    double M = p->Mns / 1.4;
    double sS = 100. / p->S;
    o->nbfar_coeff = 4.2E+45 * p->gs * (M*M*M) * (sS*sS*sS*sS);

    o->cos_t = xalloc(3 * o->nth); o->dcos_t = xalloc(3 * o->nth);
    o->sin_t = xalloc(3 * o->nth); o->dl = xalloc(3 * o->nth);
    size_t n = (size_t)o->nnb * 4 * ne;
    for (int k = 0; k < 4; ++k) {
        o->psi[k] = xalloc(n);
        o->hsum[k] = xalloc((size_t)o->nnb * NSZ);
        o->hvv[k] = xalloc((size_t)o->nbeam * NSZ);
        if (!o->psi[k] || !o->hsum[k] || !o->hvv[k]) return NULL;
    }
    for (int k = 0; k < 4; ++k) o->cur[k] = k;
    o->r = p->R0; o->dr = p->dr;
    return o;
}

void oracle_destroy(oracle_t* o) {
    if (!o) return;
    for (int s = 0; s < NPR; ++s) free(o->fq[s]);
    free(o->hvac); free(o->cos2R); free(o->dcos2R); free(o->sinR);
    free(o->sinphi); free(o->cosphi);
    free(o->cos_t); free(o->dcos_t); free(o->sin_t); free(o->dl);
    for (int k = 0; k < 4; ++k) { free(o->psi[k]); free(o->hsum[k]); free(o->hvv[k]); }
    free(o);
}

/* accessors used by the tests */
int oracle_ne(const oracle_t* o) { return o->ne; }
int oracle_nbeam(const oracle_t* o) { return o->nbeam; }
void oracle_get_tables(const oracle_t* o, double* fv, double* hvac, double* L_E, double* Ev,
                       double* cos2R, double* coef) {
    for (int s = 0; s < NPR; ++s) {
        if (fv) memcpy(fv + (size_t)s * o->ne, o->fq[s], o->ne * sizeof(double));
        if (L_E) L_E[s] = o->L_E[s];
        if (Ev) Ev[s] = o->Ev[s];
    }
    if (hvac) memcpy(hvac, o->hvac, 2 * o->ne * sizeof(double));
    if (cos2R) memcpy(cos2R, o->cos2R, o->nth * sizeof(double));
    if (coef) *coef = o->integ_coef;
}

static inline double* PSI(oracle_t* o, int set, int nb, int c) {
    return o->psi[set] + ((size_t)nb * 4 + c) * o->ne;
}

/* ---- initial state: NBeam::NBeam(int), NBeam.cpp:236-277 (no PERTERB) --- */
void oracle_init_pure(oracle_t* o, int set) {
    int ne = o->ne;
    for (int nb = 0; nb < o->nnb; ++nb) {
        int prtc = nb % NPR;
        double* ar = PSI(o, set, nb, 0); double* ai = PSI(o, set, nb, 1);
        double* br = PSI(o, set, nb, 2); double* bi = PSI(o, set, nb, 3);
        for (int e = 0; e < ne; ++e) {
            double one_sq = sqrt(1 - 0.*0. - 0.*0.);
            ar[e] = (prtc < 2) ? one_sq : 0.;
            ai[e] = 0.;
            br[e] = (prtc < 2) ? 0. : one_sq;
            bi[e] = 0.;
        }
        o->hsum[set][nb*4+0] = o->hsum[set][nb*4+1] = o->hsum[set][nb*4+2] = o->hsum[set][nb*4+3] = 0.;
    }
}
void oracle_set_psi(oracle_t* o, int set, const double* psi) {
    memcpy(o->psi[set], psi, (size_t)o->nnb * 4 * o->ne * sizeof(double));
}
void oracle_get_psi(const oracle_t* o, int set, double* psi) {
    memcpy(psi, o->psi[set], (size_t)o->nnb * 4 * o->ne * sizeof(double));
}
void oracle_get_hsum(const oracle_t* o, int set, double* out) {
    memcpy(out, o->hsum[set], (size_t)o->nnb * NSZ * sizeof(double));
}
void oracle_get_hvv(const oracle_t* o, int k, double* out) {
    memcpy(out, o->hvv[k], (size_t)o->nbeam * NSZ * sizeof(double));
}

/* ---- per-energy kernels: NBeam.cpp:299-322 (density), :360-393 (U) ------ */
static inline void o_density(double ar, double ai, double br, double bi, double ret[4]) {
    ret[0] = .5 * ( (ar*ar + ai*ai) - (br*br + bi*bi) );
    ret[1] = 0.;
    ret[2] = (ar * br) - (ai * (-bi));
    ret[3] = (ar * (-bi)) + (ai * br);
}

static inline void o_U(double dr, double h_r0, double h_r1, double h_i1,
                       double a_r, double a_i, double b_r, double b_i,
                       double* n_ar, double* n_ai, double* n_br, double* n_bi) {
    double lambda = sqrt( h_r0*h_r0 + (h_r1*h_r1 + h_i1*h_i1) );
    double ldr = lambda * dr;
    double cosCoef = cos(ldr);
    double sinCoef = sin(ldr) / lambda;
    double mat0 =  cosCoef;
    double mat1 = -h_r0 * sinCoef;
    double mat2 =  h_i1 * sinCoef;
    double mat3 = -h_r1 * sinCoef;
    double mat4 = -mat2, mat5 = mat3, mat6 = mat0, mat7 = -mat1;
    double r00r = (mat0 * a_r) - (mat1 * a_i), r00i = (mat0 * a_i) + (mat1 * a_r);
    double r01r = (mat2 * b_r) - (mat3 * b_i), r01i = (mat2 * b_i) + (mat3 * b_r);
    double r10r = (mat4 * a_r) - (mat5 * a_i), r10i = (mat4 * a_i) + (mat5 * a_r);
    double r11r = (mat6 * b_r) - (mat7 * b_i), r11i = (mat6 * b_i) + (mat7 * b_r);
    *n_ar = r00r + r01r;
    *n_ai = r00i + r01i;
    *n_br = r10r + r11r;
    *n_bi = r10i + r11i;
}

/* exported single-cell kernels for the closed-form tests */
void oracle_U(double dl, double h_r0, double h_r1, double h_i1, const double* in, double* out) {
    o_U(dl, h_r0, h_r1, h_i1, in[0], in[1], in[2], in[3], &out[0], &out[1], &out[2], &out[3]);
}
void oracle_density(const double* in, double* out) { o_density(in[0], in[1], in[2], in[3], out); }

/* per-species Hamiltonian from hvv + matter, NBeam.cpp:602-607 */
static inline void o_hmvv(int ptc, const double* hvv, double hmatt, double h[4]) {
    h[0] = (ptc%2) ? -(hvv[0] + hmatt) : (hvv[0] + hmatt);
    h[1] = hvv[1];
    h[2] = (ptc%2) ? -hvv[2] : hvv[2];
    h[3] = hvv[3];
}

/* NBeam::evolveBinsHvv, NBeam.cpp:588-646 */
static void o_evolve_bins_hvv(oracle_t* o, int in, int out, int nb, double dr,
                              const double* hvv, double hmatt) {
    const int ne = o->ne, ptc = nb % NPR;
    double* restrict ar = PSI(o, out, nb, 0); double* restrict ai = PSI(o, out, nb, 1);
    double* restrict br = PSI(o, out, nb, 2); double* restrict bi = PSI(o, out, nb, 3);
    const double* restrict bm_ar = PSI(o, in, nb, 0); const double* restrict bm_ai = PSI(o, in, nb, 1);
    const double* restrict bm_br = PSI(o, in, nb, 2); const double* restrict bm_bi = PSI(o, in, nb, 3);
    const double* restrict fv = o->fq[ptc];
    const double* restrict hv = o->hvac;
    double hm[4]; o_hmvv(ptc, hvv, hmatt, hm);
    double r0 = 0., r1 = 0., r2 = 0., r3 = 0.;
#pragma omp simd reduction(+: r0,r1,r2,r3)
    for (int e = 0; e < ne; ++e) {
        double h0 = hv[2*e], h1 = hv[2*e+1];
        double ha0 = h0 + hm[0], ha2 = h1 + hm[2], ha3 = hm[3];
        o_U(dr, ha0, ha2, ha3, bm_ar[e], bm_ai[e], bm_br[e], bm_bi[e], &ar[e], &ai[e], &br[e], &bi[e]);
        double d[4];
        o_density(ar[e], ai[e], br[e], bi[e], d);
        r0 += d[0] * fv[e];
        r1 += d[1] * fv[e];
        r2 += d[2] * fv[e];
        r3 += d[3] * fv[e];
    }
    double* hs = o->hsum[out] + (size_t)nb * 4;
    hs[0] = r0; hs[1] = r1; hs[2] = r2; hs[3] = r3;
}

/* NBeam::evolveBins(const NBeam&, ...), NBeam.cpp:648-688 */
static void o_evolve_bins(oracle_t* o, int in, int out, int nb, double dr,
                          const double* hvv, double hmatt) {
    const int ne = o->ne, ptc = nb % NPR;
    double* restrict ar = PSI(o, out, nb, 0); double* restrict ai = PSI(o, out, nb, 1);
    double* restrict br = PSI(o, out, nb, 2); double* restrict bi = PSI(o, out, nb, 3);
    const double* restrict bm_ar = PSI(o, in, nb, 0); const double* restrict bm_ai = PSI(o, in, nb, 1);
    const double* restrict bm_br = PSI(o, in, nb, 2); const double* restrict bm_bi = PSI(o, in, nb, 3);
    const double* restrict hv = o->hvac;
    double hm[4]; o_hmvv(ptc, hvv, hmatt, hm);
#pragma omp simd
    for (int e = 0; e < ne; ++e) {
        double ha0 = hv[2*e] + hm[0], ha2 = hv[2*e+1] + hm[2], ha3 = hm[3];
        o_U(dr, ha0, ha2, ha3, bm_ar[e], bm_ai[e], bm_br[e], bm_bi[e], &ar[e], &ai[e], &br[e], &bi[e]);
    }
}

/* NBeam::evolveBinsAvg(const NBeam&, ..., NBeam& avg), NBeam.cpp:462-514 */
static void o_evolve_bins_avg(oracle_t* o, int in, int out, int avg, int nb, double dr,
                              const double* hvv, double hmatt) {
    const int ne = o->ne, ptc = nb % NPR;
    double* restrict ar = PSI(o, out, nb, 0); double* restrict ai = PSI(o, out, nb, 1);
    double* restrict br = PSI(o, out, nb, 2); double* restrict bi = PSI(o, out, nb, 3);
    const double* restrict bm_ar = PSI(o, in, nb, 0); const double* restrict bm_ai = PSI(o, in, nb, 1);
    const double* restrict bm_br = PSI(o, in, nb, 2); const double* restrict bm_bi = PSI(o, in, nb, 3);
    double* restrict av_ar = PSI(o, avg, nb, 0); double* restrict av_ai = PSI(o, avg, nb, 1);
    double* restrict av_br = PSI(o, avg, nb, 2); double* restrict av_bi = PSI(o, avg, nb, 3);
    const double* restrict hv = o->hvac;
    double hm[4]; o_hmvv(ptc, hvv, hmatt, hm);
#pragma omp simd
    for (int e = 0; e < ne; ++e) {
        double ha0 = hv[2*e] + hm[0], ha2 = hv[2*e+1] + hm[2], ha3 = hm[3];
        o_U(dr, ha0, ha2, ha3, bm_ar[e], bm_ai[e], bm_br[e], bm_bi[e], &ar[e], &ai[e], &br[e], &bi[e]);
        av_ar[e] = (ar[e] + av_ar[e]) * .5;
        av_ai[e] = (ai[e] + av_ai[e]) * .5;
        av_br[e] = (br[e] + av_br[e]) * .5;
        av_bi[e] = (bi[e] + av_bi[e]) * .5;
    }
}

/* NBeam::evolveBinsAvg(int, ..., NBeam& avg) in place, NBeam.cpp:690-736 */
static void o_evolve_bins_avg_inplace(oracle_t* o, int io, int avg, int nb, double dr,
                                      const double* hvv, double hmatt) {
    const int ne = o->ne, ptc = nb % NPR;
    double* restrict ar = PSI(o, io, nb, 0); double* restrict ai = PSI(o, io, nb, 1);
    double* restrict br = PSI(o, io, nb, 2); double* restrict bi = PSI(o, io, nb, 3);
    double* restrict av_ar = PSI(o, avg, nb, 0); double* restrict av_ai = PSI(o, avg, nb, 1);
    double* restrict av_br = PSI(o, avg, nb, 2); double* restrict av_bi = PSI(o, avg, nb, 3);
    const double* restrict hv = o->hvac;
    double hm[4]; o_hmvv(ptc, hvv, hmatt, hm);
#pragma omp simd
    for (int e = 0; e < ne; ++e) {
        double ha0 = hv[2*e] + hm[0], ha2 = hv[2*e+1] + hm[2], ha3 = hm[3];
        o_U(dr, ha0, ha2, ha3, ar[e], ai[e], br[e], bi[e], &ar[e], &ai[e], &br[e], &bi[e]);
        av_ar[e] = (ar[e] + av_ar[e]) * .5;
        av_ai[e] = (ai[e] + av_ai[e]) * .5;
        av_br[e] = (br[e] + av_br[e]) * .5;
        av_bi[e] = (bi[e] + av_bi[e]) * .5;
    }
}

/* NBeam::calcESum, NBeam.cpp:812-840 */
static void o_calc_esum_nb(oracle_t* o, int set, int nb) {
    const int ne = o->ne, ptc = nb % NPR;
    const double* restrict ar = PSI(o, set, nb, 0); const double* restrict ai = PSI(o, set, nb, 1);
    const double* restrict br = PSI(o, set, nb, 2); const double* restrict bi = PSI(o, set, nb, 3);
    const double* restrict fv = o->fq[ptc];
    double r0 = 0., r1 = 0., r2 = 0., r3 = 0.;
#pragma omp simd reduction(+: r0,r1,r2,r3)
    for (int e = 0; e < ne; ++e) {
        double d[4];
        o_density(ar[e], ai[e], br[e], bi[e], d);
        r0 += d[0] * fv[e];
        r1 += d[1] * fv[e];
        r2 += d[2] * fv[e];
        r3 += d[3] * fv[e];
    }
    double* hs = o->hsum[set] + (size_t)nb * 4;
    hs[0] = r0; hs[1] = r1; hs[2] = r2; hs[3] = r3;
}

/* NBeam::calcErrHvv, NBeam.cpp:903-938 (this = a, beam = b) */
static double o_calc_err_hvv_nb(oracle_t* o, int a, int b, int nb) {
    const int ne = o->ne, ptc = nb % NPR;
    const double* restrict ar = PSI(o, a, nb, 0); const double* restrict ai = PSI(o, a, nb, 1);
    const double* restrict br = PSI(o, a, nb, 2); const double* restrict bi = PSI(o, a, nb, 3);
    const double* restrict bm_ar = PSI(o, b, nb, 0); const double* restrict bm_ai = PSI(o, b, nb, 1);
    const double* restrict bm_br = PSI(o, b, nb, 2); const double* restrict bm_bi = PSI(o, b, nb, 3);
    const double* restrict fv = o->fq[ptc];
    double eps = 0., r0 = 0., r1 = 0., r2 = 0., r3 = 0.;
#pragma omp simd reduction(+: eps,r0,r1,r2,r3)
    for (int e = 0; e < ne; ++e) {
        double dar = ar[e] - bm_ar[e], dai = ai[e] - bm_ai[e];
        double dbr = br[e] - bm_br[e], dbi = bi[e] - bm_bi[e];
        eps += ( (dar*dar + dai*dai) + (dbr*dbr + dbi*dbi) ) * (fv[e]);
        double d[4];
        o_density(ar[e], ai[e], br[e], bi[e], d);
        r0 += d[0] * fv[e];
        r1 += d[1] * fv[e];
        r2 += d[2] * fv[e];
        r3 += d[3] * fv[e];
    }
    double* hs = o->hsum[a] + (size_t)nb * 4;
    hs[0] = r0; hs[1] = r1; hs[2] = r2; hs[3] = r3;
    return eps;
}

/* NBeam::addAvg, NBeam.cpp:880-901 (obeam = (obeam + ibeam) * .5) */
static void o_add_avg_nb(oracle_t* o, int in, int out, int nb) {
    for (int c = 0; c < 4; ++c) {
        double* restrict x = PSI(o, out, nb, c);
        const double* restrict y = PSI(o, in, nb, c);
#pragma omp simd
        for (int e = 0; e < o->ne; ++e) x[e] = (x[e] + y[e]) * .5;
    }
}

/* ---- geometry: Phy_MA.cpp:172-196 / Phy_MAA.cpp:194-210 (angle bins);
 *      SA keeps the radius of the point (rarr, Phy_SA.cpp:119-122), here in
 *      cos_t[pt] ----------------------------------------------------------- */
void oracle_calc_angle_bins(oracle_t* o, double r, int pt) {
    const int nth = o->nth;
    if (o->module == XF_SA) {
        o->cos_t[pt] = r;
    } else if (o->module == XF_MAA) {
        double ra = o->p.Rv / r;
        double r2 = ra * ra;
#pragma omp parallel for
        for (int a = 0; a < nth; ++a) {
            int idx = pt*nth + a;
            o->cos_t[idx] = sqrt(1. - r2 * (1. - o->cos2R[a]));
            o->sin_t[idx] = ra * o->sinR[a];
            o->dcos_t[idx] = .5 * (1./o->cos_t[idx]) * r2 * o->dcos2R[a];
        }
    } else {
        double r2 = o->p.Rv / r;
        r2 *= r2;
#pragma omp parallel for
        for (int a = 0; a < nth; ++a) {
            int idx = pt*nth + a;
            o->cos_t[idx] = sqrt(1. - r2 * (1. - o->cos2R[a]));
            o->dcos_t[idx] = .5 * (1./o->cos_t[idx]) * r2 * o->dcos2R[a];
        }
    }
}

/* Phy_MA.cpp:218-227 (dl = dr / cos at the start point),
 * Phy_MAA.cpp:232-241 (trapezoid: dr / ((cos_s + cos_e) * .5)) and
 * Phy_SA.cpp:144-147 (dl = dr) */
void oracle_calc_delta_ls(oracle_t* o, double dr, int cur, int s, int e) {
    const int nth = o->nth;
    for (int a = 0; a < nth; ++a) {
        if (o->module == XF_SA)
            o->dl[cur*nth + a] = dr;
        else if (o->module == XF_MAA)
            o->dl[cur*nth + a] = dr / ( (o->cos_t[s*nth + a] + o->cos_t[e*nth + a])*.5 );
        else
            o->dl[cur*nth + a] = dr / ( (o->cos_t[s*nth + a] ) );
    }
}

/* upd_nu_coef, NBeam.h:47-53 */
static inline void o_upd_nu_coef(const double* nu, const double* anu, double n_cf, double an_cf, double* ret) {
    ret[0] += nu[0] * n_cf - anu[0] * an_cf;
    ret[1] += nu[1] * n_cf - anu[1] * an_cf;
    ret[2] += nu[2] * n_cf - anu[2] * an_cf;
    ret[3] += nu[3] * n_cf + anu[3] * an_cf;
}

/* phy_ma::calc_Hvv -> getHvv_partial -> getHvv, Phy_MA.cpp:372-474 */
static void o_calc_hvv_ma(oracle_t* o, int set, int pt, double* hvv) {
    const int nth = o->nth;
    const double* Cos = o->cos_t + pt*nth;
    const double* dCos = o->dcos_t + pt*nth;
    double* np = xalloc((size_t)nth * NSZ);
    double* npc = xalloc((size_t)nth * NSZ);
#pragma omp parallel for
    for (int a = 0; a < nth; ++a) {
        for (int n = 0; n < 2; ++n) {
            int neu = n << 1, aneu = neu + 1;
            const double* rn = o->hsum[set] + ((size_t)a*NPR + neu) * 4;
            const double* ra = o->hsum[set] + ((size_t)a*NPR + aneu) * 4;
            o_upd_nu_coef(rn, ra, o->L_E[neu], o->L_E[aneu], &np[a*NSZ]);
        }
        for (int z = 0; z < NSZ; ++z) {
            np[a*NSZ+z] *= dCos[a];
            npc[a*NSZ+z] = np[a*NSZ+z]*Cos[a];
        }
    }
    double J0[NSZ] = {0, 0, 0, 0}, J1[NSZ] = {0, 0, 0, 0};
    for (int a = 0; a < nth; ++a)
        for (int z = 0; z < NSZ; ++z) {
            J0[z] += np[a*NSZ+z];
            J1[z] += npc[a*NSZ+z];
        }
    free(np); free(npc);
#pragma omp parallel for
    for (int a = 0; a < nth; ++a)
        for (int z = 0; z < NSZ; ++z)
            hvv[a*NSZ+z] = ( J0[z] - (J1[z] * Cos[a]) ) * o->integ_coef;
}

/* phy_maa::calc_Hvv -> getHvv_partial -> getHvv, Phy_MAA.cpp:445-621 */
static void o_calc_hvv_maa(oracle_t* o, int set, int pt, double* hvv) {
    const int nth = o->nth, nph = o->nph;
    double* pPT = xalloc((size_t)nth * NSZ);
    double* pCS = xalloc((size_t)nth * NSZ);
    double* pSS = xalloc((size_t)nth * NSZ);
    double* pPC = xalloc((size_t)nth * NSZ);
#pragma omp parallel for
    for (int a = 0; a < nth; ++a) {
        double cP[NSZ] = {0, 0, 0, 0}, sP[NSZ] = {0, 0, 0, 0};
        for (int q = 0; q < nph; ++q) {
            double pe[NSZ] = {0, 0, 0, 0};
            for (int n = 0; n < 2; ++n) {
                int neu = n << 1, aneu = neu + 1;
                const double* rn = o->hsum[set] + (((size_t)a*nph + q)*NPR + neu) * 4;
                const double* ra = o->hsum[set] + (((size_t)a*nph + q)*NPR + aneu) * 4;
                o_upd_nu_coef(rn, ra, o->L_E[neu], o->L_E[aneu], pe);
            }
            for (int z = 0; z < NSZ; ++z) {
                pPT[a*NSZ+z] += pe[z];
                cP[z] += pe[z] * o->cosphi[q];
                sP[z] += pe[z] * o->sinphi[q];
            }
        }
        int aid = pt*nth + a;
        for (int z = 0; z < NSZ; ++z) {
            int idx = a*NSZ + z;
            pPT[idx] *= o->dcos_t[aid];
            pCS[idx] = cP[z] * o->sin_t[aid] * o->dcos_t[aid];
            pSS[idx] = sP[z] * o->sin_t[aid] * o->dcos_t[aid];
            pPC[idx] = pPT[idx] * o->cos_t[aid];
        }
    }
    double J0[NSZ], Jx[NSZ], Jy[NSZ], Jz[NSZ];
    for (int z = 0; z < NSZ; ++z) {
        for (int a = 1; a < nth; ++a) {
            int idx = a*NSZ + z;
            pPT[z] += pPT[idx];
            pCS[z] += pCS[idx];
            pSS[z] += pSS[idx];
            pPC[z] += pPC[idx];
        }
        J0[z] = pPT[z]; Jx[z] = pCS[z]; Jy[z] = pSS[z]; Jz[z] = pPC[z];
    }
    free(pPT); free(pCS); free(pSS); free(pPC);
#pragma omp parallel for
    for (int a = 0; a < nth; ++a)
        for (int q = 0; q < nph; ++q) {
            int aid = pt*nth + a;
            double st = o->sin_t[aid], ct = o->cos_t[aid], sp = o->sinphi[q], cp = o->cosphi[q];
            double* ret = hvv + ((size_t)a*nph + q) * NSZ;
            for (int z = 0; z < NSZ; ++z)
                ret[z] = ( J0[z] - ( st * cp * Jx[z] + st * sp * Jy[z] + ct * Jz[z] ) ) * o->integ_coef;
        }
}

/* phy_sa::D, Phy_SA.cpp:250-256 */
double oracle_sa_D(double Rv, double r) {
    double R = Rv / r;
    double tm = 1. - sqrt( 1. - (R*R) );
    return 0.5 * (tm*tm);
}

/* phy_sa::calc_Hvv, Phy_SA.cpp:259-283: the four NBeams' energy sums combined
 * by upd_nu_coef, times integ_coef * D(r) of the step point */
static void o_calc_hvv_sa(oracle_t* o, int set, int pt, double* hvv) {
    for (int z = 0; z < NSZ; ++z) hvv[z] = 0.;
    for (int n = 0; n < 2; ++n) {
        int neu = n << 1, aneu = neu + 1;
        o_upd_nu_coef(o->hsum[set] + neu*4, o->hsum[set] + aneu*4, o->L_E[neu], o->L_E[aneu], hvv);
    }
    double coeff = o->integ_coef * oracle_sa_D(o->p.Rv, o->cos_t[pt]);
    for (int z = 0; z < NSZ; ++z) hvv[z] *= coeff;
}

void oracle_calc_hvv(oracle_t* o, int set, int pt, int k) {
    if (o->module == XF_MAA) o_calc_hvv_maa(o, set, pt, o->hvv[k]);
    else if (o->module == XF_SA) o_calc_hvv_sa(o, set, pt, o->hvv[k]);
    else o_calc_hvv_ma(o, set, pt, o->hvv[k]);
}

/* phy::evolve / evolveHvv / evolveAvg / evolveAvg(in place) / avgBeam drivers:
 * Phy_MA.cpp:476-540,580-593 and Phy_MAA.cpp:716-803,855-871.  The hvv of a
 * beam is hvv[(theta*Pbins+phi)*4]; dl depends on theta only.             */
#define BEAM_LOOP(body) do {                                                 \
    const int nth = o->nth, nph = o->nph;                                    \
    _Pragma("omp parallel for")                                              \
    for (int a = 0; a < nth; ++a) {                                          \
        double dl = o->dl[pt*nth + a];                                       \
        for (int q = 0; q < nph; ++q) {                                      \
            const double* hv = o->hvv[k] + ((size_t)a*nph + q) * NSZ;        \
            for (int prc = 0; prc < NPR; ++prc) {                            \
                int nb = (a*nph + q) * NPR + prc;                            \
                body;                                                        \
            }                                                                \
        }                                                                    \
    }                                                                        \
} while (0)

void oracle_evolve(oracle_t* o, int in, int pt, int k, double hm, int out) {
    BEAM_LOOP(o_evolve_bins(o, in, out, nb, dl, hv, hm));
}
void oracle_evolve_hvv(oracle_t* o, int in, int pt, int k, double hm, int out) {
    BEAM_LOOP(o_evolve_bins_hvv(o, in, out, nb, dl, hv, hm));
}
void oracle_evolve_avg(oracle_t* o, int in, int pt, int k, double hm, int out, int avg) {
    BEAM_LOOP(o_evolve_bins_avg(o, in, out, avg, nb, dl, hv, hm));
}
void oracle_evolve_avg_inplace(oracle_t* o, int pt, int k, double hm, int io, int avg) {
    BEAM_LOOP(o_evolve_bins_avg_inplace(o, io, avg, nb, dl, hv, hm));
}
void oracle_avg_beam(oracle_t* o, int in, int out) {
#pragma omp parallel for
    for (int nb = 0; nb < o->nnb; ++nb) o_add_avg_nb(o, in, out, nb);
}
void oracle_calc_esum(oracle_t* o, int set) {
#pragma omp parallel for
    for (int nb = 0; nb < o->nnb; ++nb) o_calc_esum_nb(o, set, nb);
}
/* max over NBeams of calcErrHvv (Nmr.cpp:195-205), before the square root */
double oracle_calc_err_hvv(oracle_t* o, int a, int b) {
    double max_err = -1.;
#pragma omp parallel for reduction(max: max_err)
    for (int nb = 0; nb < o->nnb; ++nb) {
        double e = o_calc_err_hvv_nb(o, a, b, nb);
        max_err = (e > max_err) ? e : max_err;
    }
    return max_err;
}

/* ---- the stepper: nmr::evolutionLoop, Nmr.cpp:89-262 ---------------------
 * state sets: cur[0] = nubm_0 (psi0), cur[1] = nubm_1, cur[2] = nubm_3,
 * cur[3] = nubm_4; hvv ids 0..3 = h0, h1, h3, h4.                          */
void oracle_set_state(oracle_t* o, const double* psi, double r, double dr) {
    for (int k = 0; k < 4; ++k) o->cur[k] = k;
    oracle_set_psi(o, 0, psi);
    for (int k = 1; k < 4; ++k) oracle_set_psi(o, k, psi);
    o->r = r; o->dr = dr;
    o->counter_tot = o->counter = 0;
    o->esum_valid = 0;
}
void oracle_state_init_pure(oracle_t* o) {
    for (int k = 0; k < 4; ++k) { o->cur[k] = k; oracle_init_pure(o, k); }
    o->r = o->p.R0; o->dr = o->p.dr;
    o->counter_tot = o->counter = 0;
    o->esum_valid = 0;
}
void oracle_get_state(const oracle_t* o, double* psi, double* r, double* dr, int64_t* counters) {
    if (psi) oracle_get_psi(o, o->cur[0], psi);
    if (r) *r = o->r;
    if (dr) *dr = o->dr;
    if (counters) { counters[0] = o->counter_tot; counters[1] = o->counter; }
}

/* Attempts up to max_steps steps; stops when r reaches Rn.  Returns the
 * number attempted.  log (optional, max_steps records) gets (r, dr, err,
 * accepted) per attempt.  *loop_seconds receives the loop wall time.      */
int64_t oracle_run(oracle_t* o, int64_t max_steps, xo_rec* log, double* loop_seconds) {
    struct timeval t0, t1;
    gettimeofday(&t0, NULL);
    int b0 = o->cur[0], b1 = o->cur[1], b3 = o->cur[2], b4 = o->cur[3];
    double r = o->r, dr = o->dr;
    const double Rn = o->p.Rn, eps0 = o->p.eps0, ksi = o->p.ksi, maxDr = o->p.max_dr;
    if (!o->esum_valid) { oracle_calc_esum(o, b0); o->esum_valid = 1; }
    int64_t n = 0;
    while (n < max_steps && r < Rn) {
        double dr2 = .5 * dr;
        double rdr = r + dr;
        double rdr2 = r + dr2;
        double hm_r = oracle_get_hm(o, r);
        double hm_rdr2 = oracle_get_hm(o, rdr2);
        double hm_rdr = oracle_get_hm(o, rdr);
        oracle_calc_angle_bins(o, r, 0);
        oracle_calc_angle_bins(o, rdr2, 1);
        oracle_calc_angle_bins(o, rdr, 2);
        oracle_calc_delta_ls(o, dr2, 0, 0, 1);
        oracle_calc_delta_ls(o, dr, 1, 0, 2);
        oracle_calc_delta_ls(o, dr2, 2, 1, 2);

        oracle_calc_hvv(o, b0, 0, 0);                       /* S1        */
        oracle_evolve_hvv(o, b0, 1, 0, hm_r, b1);           /* S2        */
        oracle_evolve_hvv(o, b0, 0, 0, hm_r, b3);           /* S6        */
        oracle_calc_hvv(o, b1, 2, 1);                       /* S3        */
        oracle_evolve_avg(o, b0, 1, 1, hm_rdr, b4, b1);     /* S4/S5     */
        oracle_calc_hvv(o, b3, 1, 2);                       /* S7        */
        oracle_evolve_hvv(o, b0, 1, 2, hm_rdr2, b4);        /* S8        */
        oracle_calc_hvv(o, b4, 2, 3);                       /* S9        */
        oracle_evolve_avg_inplace(o, 2, 3, hm_rdr, b3, b4); /* S10/S11   */

        double max_err = sqrt(oracle_calc_err_hvv(o, b4, b1));
        int acc = (max_err < eps0);
        if (log) { log[n].r = r; log[n].dr = dr; log[n].err = max_err; log[n].accepted = acc; log[n].pad = 0; }
        if (acc) {
            r += dr;
            int tmp = b0; b0 = b4; b4 = tmp;
            o->counter++;
        }
        o->counter_tot++;
        dr *= ksi * sqrt(eps0 / max_err);
        dr = ( dr > maxDr ) ? maxDr : dr;
        dr = ( r+dr > Rn ) ? Rn - r : dr;
        ++n;
    }
    o->cur[0] = b0; o->cur[1] = b1; o->cur[2] = b3; o->cur[3] = b4;
    o->r = r; o->dr = dr;
    gettimeofday(&t1, NULL);
    if (loop_seconds) *loop_seconds = (t1.tv_sec - t0.tv_sec) + 1e-6 * (t1.tv_usec - t0.tv_usec);
    return n;
}

int oracle_max_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}
void oracle_set_threads(int n) {
#ifdef _OPENMP
    if (n > 0) omp_set_num_threads(n);
#else
    (void)n;
#endif
}

// xf_phy.hip -- the fine-grained mirror of the reference's Phy.h plugin
// interface (include/Phy.h:24-159) on the device: four FP64 beam sets with
// all four species (nubm_0, nubm_1, nubm_3, nubm_4 of Nmr.cpp:42-46), four
// hvv buffers, the per-step geometry caches, and one kernel per Phy.h entry
// point.  It is the 1:1 debugging surface; the fused stepper (xf_step_*) is
// the production path.
//
// Device layout: psi[set][prtcl*4 + comp][beam*NE + e]; hsum[set][nbeam*4][4]
// (the per-NBeam energy sums, NBeam::hSumMat); hvv[k][beam][4].
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstring>
#include <vector>

#include "xf_phy.h"

#define HIPCHK(x)                                                                           \
    do {                                                                                    \
        hipError_t e_ = (x);                                                                \
        if (e_ != hipSuccess) return xf_fail(XF_EHIP, "%s: %s", #x, hipGetErrorString(e_)); \
    } while (0)

struct XfPhyDev {
    int module, nth, nph, ne, nbeam;
    const double *hv, *fv, *x, *dx, *sinR, *cphi, *sphi;
    double *cos_t, *dcos_t, *sin_t, *dl;
    double L_E[4];
    double coef, Rv;
};

struct XfPhy {
    XfPhyDev d{};
    double* psi[4] = {};
    double* hsum[4] = {};
    double* hvv[4] = {};
    double* part = nullptr;   // [nth][16] per-theta partials
    double* eps = nullptr;    // [nbeam*4]
    std::vector<void*> owned;
    ~XfPhy() {
        for (void* p : owned) (void)hipFree(p);
    }
};

void xf_phy_destroy(XfPhy* p) { delete p; }

namespace {

enum { K_EVOLVE = 0, K_HVV, K_AVG, K_AVG_INPLACE, K_ESUM, K_ERR, K_ADDAVG };

struct C4 { double ar, ai, br, bi; };

// NBeam::U (NBeam.cpp:360-393)
__device__ __forceinline__ C4 U(double dr, double h_r0, double h_r1, double h_i1, const C4& a) {
    double lambda = sqrt(h_r0 * h_r0 + (h_r1 * h_r1 + h_i1 * h_i1));
    double ldr = lambda * dr;
    double s, c;
    sincos(ldr, &s, &c);
    double sinCoef = s / lambda;
    double m0 = c, m1 = -h_r0 * sinCoef, m2 = h_i1 * sinCoef, m3 = -h_r1 * sinCoef;
    double m4 = -m2, m5 = m3, m6 = m0, m7 = -m1;
    C4 o;
    o.ar = ((m0 * a.ar) - (m1 * a.ai)) + ((m2 * a.br) - (m3 * a.bi));
    o.ai = ((m0 * a.ai) + (m1 * a.ar)) + ((m2 * a.bi) + (m3 * a.br));
    o.br = ((m4 * a.ar) - (m5 * a.ai)) + ((m6 * a.br) - (m7 * a.bi));
    o.bi = ((m4 * a.ai) + (m5 * a.ar)) + ((m6 * a.bi) + (m7 * a.br));
    return o;
}

// NBeam::density (NBeam.cpp:314-322)
__device__ __forceinline__ void dens(const C4& a, double* r) {
    r[0] = .5 * ((a.ar * a.ar + a.ai * a.ai) - (a.br * a.br + a.bi * a.bi));
    r[1] = 0.;
    r[2] = (a.ar * a.br) - (a.ai * (-a.bi));
    r[3] = (a.ar * (-a.bi)) + (a.ai * a.br);
}

__device__ __forceinline__ C4 ld(const double* p, long long nc, int nb_prtcl, long long cell) {
    const double* q = p + (long long)(nb_prtcl * 4) * nc + cell;
    return C4{q[0], q[nc], q[2 * nc], q[3 * nc]};
}
__device__ __forceinline__ void st(double* p, long long nc, int prtcl, long long cell, const C4& v) {
    double* q = p + (long long)(prtcl * 4) * nc + cell;
    q[0] = v.ar; q[nc] = v.ai; q[2 * nc] = v.br; q[3 * nc] = v.bi;
}

__device__ __forceinline__ double block_sum(double v, double* sh) {
    for (int o = 32; o >= 1; o >>= 1) v += __shfl_xor(v, o, 64);
    const int w = threadIdx.x >> 6, nw = blockDim.x >> 6;
    __syncthreads();
    if ((threadIdx.x & 63) == 0) sh[w] = v;
    __syncthreads();
    double s = 0.;
    for (int i = 0; i < nw; ++i) s += sh[i];
    return s;
}

// One workgroup per NBeam (beam, prtcl); lanes over energy bins.  The
// per-kind bodies follow NBeam::evolveBins (NBeam.cpp:648-688),
// evolveBinsHvv (:588-646), evolveBinsAvg (:462-514), evolveBinsAvg in place
// (:690-736), calcESum (:812-840), calcErrHvv (:903-938) and addAvg (:880-901).
template <int KIND>
__global__ void __launch_bounds__(256) k_nbeam(XfPhyDev P, const double* __restrict__ in, double* __restrict__ out,
                                               double* __restrict__ avg, int pt, const double* __restrict__ hvv,
                                               double hmatt, double* __restrict__ hsum, double* __restrict__ eps) {
    __shared__ double sh[16];
    const int nb = blockIdx.x;
    const int beam = nb >> 2, prtcl = nb & 3;
    const int th = beam / P.nph;
    const int ne = P.ne;
    const long long nc = (long long)P.nbeam * ne;
    double dl = 0., hm[4] = {0., 0., 0., 0.};
    if (KIND <= K_AVG_INPLACE) {
        dl = P.dl[pt * P.nth + th];
        const double* hv = hvv + beam * 4;
        hm[0] = (prtcl % 2) ? -(hv[0] + hmatt) : (hv[0] + hmatt);
        hm[1] = hv[1];
        hm[2] = (prtcl % 2) ? -hv[2] : hv[2];
        hm[3] = hv[3];
    }
    const double* fv = P.fv + prtcl * ne;
    double r0 = 0., r2 = 0., r3 = 0., ep = 0.;
    for (int e = threadIdx.x; e < ne; e += blockDim.x) {
        const long long cell = (long long)beam * ne + e;
        if (KIND == K_ESUM) {
            C4 a = ld(in, nc, prtcl, cell);
            double d[4];
            dens(a, d);
            r0 += d[0] * fv[e]; r2 += d[2] * fv[e]; r3 += d[3] * fv[e];
        } else if (KIND == K_ERR) {
            C4 a = ld(in, nc, prtcl, cell), b = ld(out, nc, prtcl, cell);
            double dar = a.ar - b.ar, dai = a.ai - b.ai, dbr = a.br - b.br, dbi = a.bi - b.bi;
            ep += ((dar * dar + dai * dai) + (dbr * dbr + dbi * dbi)) * (fv[e]);
            double d[4];
            dens(a, d);
            r0 += d[0] * fv[e]; r2 += d[2] * fv[e]; r3 += d[3] * fv[e];
        } else if (KIND == K_ADDAVG) {
            C4 a = ld(in, nc, prtcl, cell), o = ld(out, nc, prtcl, cell);
            C4 n{(o.ar + a.ar) * .5, (o.ai + a.ai) * .5, (o.br + a.br) * .5, (o.bi + a.bi) * .5};
            st(out, nc, prtcl, cell, n);
        } else {
            double h0 = P.hv[e] + hm[0];
            double h2 = P.hv[ne + e] + hm[2];
            double h3 = hm[3];
            C4 a = ld(KIND == K_AVG_INPLACE ? out : in, nc, prtcl, cell);
            C4 o = U(dl, h0, h2, h3, a);
            st(out, nc, prtcl, cell, o);
            if (KIND == K_HVV) {
                double d[4];
                dens(o, d);
                r0 += d[0] * fv[e]; r2 += d[2] * fv[e]; r3 += d[3] * fv[e];
            }
            if (KIND == K_AVG || KIND == K_AVG_INPLACE) {
                C4 v = ld(avg, nc, prtcl, cell);
                C4 n{(o.ar + v.ar) * .5, (o.ai + v.ai) * .5, (o.br + v.br) * .5, (o.bi + v.bi) * .5};
                st(avg, nc, prtcl, cell, n);
            }
        }
    }
    if (KIND == K_HVV || KIND == K_ESUM || KIND == K_ERR) {
        r0 = block_sum(r0, sh);
        r2 = block_sum(r2, sh);
        r3 = block_sum(r3, sh);
        if (KIND == K_ERR) ep = block_sum(ep, sh);
        if (threadIdx.x == 0) {
            hsum[nb * 4 + 0] = r0;
            hsum[nb * 4 + 1] = 0.;
            hsum[nb * 4 + 2] = r2;
            hsum[nb * 4 + 3] = r3;
            if (KIND == K_ERR) eps[nb] = ep;
        }
    }
}

// phy::calcAngleBins (Phy_MA.cpp:172-196, Phy_MAA.cpp:194-210); SA keeps the
// radius of the point instead (rarr, Phy_SA.cpp:119-122), here in cos_t.
__global__ void k_angle_bins(XfPhyDev P, double r, int pt) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.nth) return;
    int idx = pt * P.nth + a;
    if (P.module == XF_SA) {
        P.cos_t[idx] = r;
    } else if (P.module == XF_MAA) {
        double ra = P.Rv / r;
        double r2 = ra * ra;
        P.cos_t[idx] = sqrt(1. - r2 * (1. - P.x[a]));
        P.sin_t[idx] = ra * P.sinR[a];
        P.dcos_t[idx] = .5 * (1. / P.cos_t[idx]) * r2 * P.dx[a];
    } else {
        double r2 = P.Rv / r;
        r2 *= r2;
        P.cos_t[idx] = sqrt(1. - r2 * (1. - P.x[a]));
        P.dcos_t[idx] = .5 * (1. / P.cos_t[idx]) * r2 * P.dx[a];
    }
}

// phy::calcDeltaLs (Phy_MA.cpp:218-227, Phy_MAA.cpp:232-241; SA: dl = dr,
// Phy_SA.cpp:144-147)
__global__ void k_delta_ls(XfPhyDev P, double dr, int cur, int s, int e) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.nth) return;
    if (P.module == XF_SA)
        P.dl[cur * P.nth + a] = dr;
    else if (P.module == XF_MAA)
        P.dl[cur * P.nth + a] = dr / ((P.cos_t[s * P.nth + a] + P.cos_t[e * P.nth + a]) * .5);
    else
        P.dl[cur * P.nth + a] = dr / ((P.cos_t[s * P.nth + a]));
}

// upd_nu_coef (NBeam.h:47-53)
__device__ __forceinline__ void upd(const double* nu, const double* anu, double n_cf, double an_cf, double* ret) {
    ret[0] += nu[0] * n_cf - anu[0] * an_cf;
    ret[1] += nu[1] * n_cf - anu[1] * an_cf;
    ret[2] += nu[2] * n_cf - anu[2] * an_cf;
    ret[3] += nu[3] * n_cf + anu[3] * an_cf;
}

// Per-theta partials of getHvv_partial: Phy_MA.cpp:402-428 (2 moments) and
// Phy_MAA.cpp:487-538 (4 moments), one thread per theta, same operations.
__global__ void k_hvv_partial(XfPhyDev P, const double* __restrict__ hsum, int pt, double* __restrict__ part) {
    int a = blockIdx.x * blockDim.x + threadIdx.x;
    if (a >= P.nth) return;
    int aid = pt * P.nth + a;
    if (P.module == XF_MAA) {
        double pPT[4] = {0., 0., 0., 0.}, cP[4] = {0., 0., 0., 0.}, sP[4] = {0., 0., 0., 0.};
        for (int q = 0; q < P.nph; ++q) {
            double pe[4] = {0., 0., 0., 0.};
            for (int n = 0; n < 2; ++n) {
                int neu = n << 1, aneu = neu + 1;
                const double* rn = hsum + (((size_t)a * P.nph + q) * 4 + neu) * 4;
                const double* ra = hsum + (((size_t)a * P.nph + q) * 4 + aneu) * 4;
                upd(rn, ra, P.L_E[neu], P.L_E[aneu], pe);
            }
            for (int z = 0; z < 4; ++z) {
                pPT[z] += pe[z];
                cP[z] += pe[z] * P.cphi[q];
                sP[z] += pe[z] * P.sphi[q];
            }
        }
        for (int z = 0; z < 4; ++z) {
            double t = pPT[z] * P.dcos_t[aid];
            part[a * 16 + z] = t;
            part[a * 16 + 4 + z] = cP[z] * P.sin_t[aid] * P.dcos_t[aid];
            part[a * 16 + 8 + z] = sP[z] * P.sin_t[aid] * P.dcos_t[aid];
            part[a * 16 + 12 + z] = t * P.cos_t[aid];
        }
    } else {
        double np[4] = {0., 0., 0., 0.};
        for (int n = 0; n < 2; ++n) {
            int neu = n << 1, aneu = neu + 1;
            const double* rn = hsum + ((size_t)a * 4 + neu) * 4;
            const double* ra = hsum + ((size_t)a * 4 + aneu) * 4;
            upd(rn, ra, P.L_E[neu], P.L_E[aneu], np);
        }
        if (P.module == XF_SA) {
            // SA: the source itself (Phy_SA.cpp:259-278)
            for (int z = 0; z < 4; ++z) part[a * 16 + z] = np[z];
            return;
        }
        for (int z = 0; z < 4; ++z) {
            np[z] *= P.dcos_t[aid];
            part[a * 16 + z] = np[z];
            part[a * 16 + 4 + z] = np[z] * P.cos_t[aid];
        }
    }
}

// The sequential angle sum (Phy_MA.cpp:431-440, Phy_MAA.cpp:541-557) and
// getHvv for every beam (Phy_MA.cpp:372-377,470-473; Phy_MAA.cpp:445-461,611-619).
__global__ void k_hvv_final(XfPhyDev P, const double* __restrict__ part, int pt, double* __restrict__ hvv) {
    __shared__ double J[16];
    const int nv = (P.module == XF_MAA) ? 16 : (P.module == XF_SA) ? 4 : 8;
    if ((int)threadIdx.x < nv) {
        double s = 0.;
        for (int a = 0; a < P.nth; ++a) s += part[a * 16 + threadIdx.x];
        J[threadIdx.x] = s;
    }
    __syncthreads();
    for (int b = threadIdx.x; b < P.nbeam; b += blockDim.x) {
        int th = b / P.nph;
        int aid = pt * P.nth + th;
        double* ret = hvv + b * 4;
        if (P.module == XF_SA) {
            // hvv *= integ_coef * D(r), D(r) = (1 - sqrt(1 - (Rv/r)^2))^2 / 2
            // (Phy_SA.cpp:250-256,280-281); cos_t holds the point's radius
            const double R = P.Rv / P.cos_t[aid];
            const double tm = 1. - sqrt(1. - (R * R));
            const double coeff = P.coef * (.5 * (tm * tm));
            for (int z = 0; z < 4; ++z) ret[z] = J[z] * coeff;
        } else if (P.module == XF_MAA) {
            int q = b - th * P.nph;
            double st_ = P.sin_t[aid], ct = P.cos_t[aid], sp = P.sphi[q], cp = P.cphi[q];
            for (int z = 0; z < 4; ++z)
                ret[z] = (J[z] - (st_ * cp * J[4 + z] + st_ * sp * J[8 + z] + ct * J[12 + z])) * P.coef;
        } else {
            double Cos = P.cos_t[aid];
            for (int z = 0; z < 4; ++z) ret[z] = (J[z] - (J[4 + z] * Cos)) * P.coef;
        }
    }
}

int dalloc(XfPhy* p, double** q, size_t n) {
    void* v = nullptr;
    hipError_t e = hipMalloc(&v, std::max<size_t>(n, 1) * sizeof(double));
    if (e != hipSuccess) return xf_fail(XF_ENOMEM, "hipMalloc: %s", hipGetErrorString(e));
    p->owned.push_back(v);
    HIPCHK(hipMemset(v, 0, std::max<size_t>(n, 1) * sizeof(double)));
    *q = (double*)v;
    return XF_OK;
}
int dput(XfPhy* p, const double** q, const std::vector<double>& v) {
    double* d = nullptr;
    int rc = dalloc(p, &d, v.size());
    if (rc) return rc;
    HIPCHK(hipMemcpy(d, v.data(), v.size() * 8, hipMemcpyHostToDevice));
    *q = d;
    return XF_OK;
}

int get_phy(xf_ctx* ctx, XfPhy** out) {
    if (!ctx) return xf_fail(XF_EINVAL, "null ctx");
    XfPhy*& ph = xf_ctx_phy(ctx);
    if (!ph) {
        const XfTables& t = xf_ctx_tables(ctx);
        const xf_params& p = xf_ctx_params(ctx);
        XfPhy* q = new XfPhy();
        XfPhyDev& d = q->d;
        d.module = t.module; d.nth = t.nth; d.nph = t.nph; d.ne = t.ne; d.nbeam = t.nbeam;
        for (int s = 0; s < 4; ++s) d.L_E[s] = t.L_E[s];
        d.coef = t.integ_coef;
        d.Rv = p.Rv;
        std::vector<double> hv(2 * t.ne);
        for (int e = 0; e < t.ne; ++e) { hv[e] = t.hvac[2 * e]; hv[t.ne + e] = t.hvac[2 * e + 1]; }
        int rc;
        size_t nst = (size_t)16 * t.nbeam * t.ne;
        if ((rc = dput(q, &d.hv, hv)) || (rc = dput(q, &d.fv, t.fv)) || (rc = dput(q, &d.x, t.cos2R)) ||
            (rc = dput(q, &d.dx, t.dcos2R)) || (rc = dput(q, &d.sinR, t.sinR)) || (rc = dput(q, &d.cphi, t.cosphi)) ||
            (rc = dput(q, &d.sphi, t.sinphi)) || (rc = dalloc(q, &d.cos_t, 3 * t.nth)) ||
            (rc = dalloc(q, &d.dcos_t, 3 * t.nth)) || (rc = dalloc(q, &d.sin_t, 3 * t.nth)) ||
            (rc = dalloc(q, &d.dl, 3 * t.nth)) || (rc = dalloc(q, &q->part, 16 * (size_t)t.nth)) ||
            (rc = dalloc(q, &q->eps, 4 * (size_t)t.nbeam))) {
            delete q;
            return rc;
        }
        for (int k = 0; k < 4; ++k) {
            if ((rc = dalloc(q, &q->psi[k], nst)) || (rc = dalloc(q, &q->hsum[k], 16 * (size_t)t.nbeam)) ||
                (rc = dalloc(q, &q->hvv[k], 4 * (size_t)t.nbeam))) {
                delete q;
                return rc;
            }
        }
        ph = q;
        for (int k = 0; k < 4; ++k) {
            rc = xf_beams_init_pure(ctx, k);
            if (rc) return rc;
        }
    }
    *out = ph;
    return XF_OK;
}

bool bad_set(int k) { return k < 0 || k > 3; }

template <int KIND>
int launch_nbeam(XfPhy* p, const double* in, double* out, double* avg, int pt, const double* hvv, double hm,
                 double* hsum, double* eps) {
    (void)hipGetLastError();    // a launch error, not one left by an earlier handled call
    hipLaunchKernelGGL(k_nbeam<KIND>, dim3(p->d.nbeam * 4), dim3(256), 0, 0, p->d, in, out, avg, pt, hvv, hm, hsum, eps);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return XF_OK;
}

}  // namespace

extern "C" {

int xf_beams_upload(xf_ctx* ctx, int set, const double* psi) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(set) || !psi) return xf_fail(XF_EINVAL, "bad beam set");
    const int ne = p->d.ne, nb = p->d.nbeam;
    const size_t nc = (size_t)nb * ne;
    std::vector<double> pl(16 * nc);
    for (int b = 0; b < nb; ++b)
        for (int s = 0; s < 4; ++s)
            for (int c = 0; c < 4; ++c)
                for (int e = 0; e < ne; ++e) pl[((size_t)s * 4 + c) * nc + (size_t)b * ne + e] = psi[(((size_t)b * 4 + s) * 4 + c) * ne + e];
    HIPCHK(hipMemcpy(p->psi[set], pl.data(), pl.size() * 8, hipMemcpyHostToDevice));
    return XF_OK;
}

int xf_beams_download(xf_ctx* ctx, int set, double* psi) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(set) || !psi) return xf_fail(XF_EINVAL, "bad beam set");
    const int ne = p->d.ne, nb = p->d.nbeam;
    const size_t nc = (size_t)nb * ne;
    std::vector<double> pl(16 * nc);
    HIPCHK(hipMemcpy(pl.data(), p->psi[set], pl.size() * 8, hipMemcpyDeviceToHost));
    for (int b = 0; b < nb; ++b)
        for (int s = 0; s < 4; ++s)
            for (int c = 0; c < 4; ++c)
                for (int e = 0; e < ne; ++e) psi[(((size_t)b * 4 + s) * 4 + c) * ne + e] = pl[((size_t)s * 4 + c) * nc + (size_t)b * ne + e];
    return XF_OK;
}

int xf_beams_init_pure(xf_ctx* ctx, int set) {
    XfPhy* p = xf_ctx_phy(ctx);
    if (!p) {
        int rc = get_phy(ctx, &p);   // get_phy initialises all sets
        return rc;
    }
    if (bad_set(set)) return xf_fail(XF_EINVAL, "bad beam set");
    const int ne = p->d.ne, nb = p->d.nbeam;
    std::vector<double> psi((size_t)nb * 16 * ne, 0.);
    for (int b = 0; b < nb; ++b)
        for (int s = 0; s < 4; ++s)
            for (int e = 0; e < ne; ++e) psi[(((size_t)b * 4 + s) * 4 + (s < 2 ? 0 : 2)) * ne + e] = 1.;
    HIPCHK(hipMemset(p->hsum[set], 0, 16 * (size_t)nb * 8));
    return xf_beams_upload(ctx, set, psi.data());
}

int xf_calc_esum(xf_ctx* ctx, int set) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(set)) return xf_fail(XF_EINVAL, "bad beam set");
    return launch_nbeam<K_ESUM>(p, p->psi[set], nullptr, nullptr, 0, nullptr, 0., p->hsum[set], nullptr);
}

int xf_get_esum(xf_ctx* ctx, int set, double* out) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(set) || !out) return xf_fail(XF_EINVAL, "bad arguments");
    HIPCHK(hipMemcpy(out, p->hsum[set], 16 * (size_t)p->d.nbeam * 8, hipMemcpyDeviceToHost));
    return XF_OK;
}

int xf_calc_angle_bins(xf_ctx* ctx, double r, int pt) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (pt < 0 || pt > 2) return xf_fail(XF_EINVAL, "bad point");
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_angle_bins, dim3((p->d.nth + 255) / 256), dim3(256), 0, 0, p->d, r, pt);
    HIPCHK(hipGetLastError());
    return XF_OK;
}

int xf_calc_delta_ls(xf_ctx* ctx, double dr, int cur, int s, int e) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (cur < 0 || cur > 2 || s < 0 || s > 2 || e < 0 || e > 2) return xf_fail(XF_EINVAL, "bad point");
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_delta_ls, dim3((p->d.nth + 255) / 256), dim3(256), 0, 0, p->d, dr, cur, s, e);
    HIPCHK(hipGetLastError());
    return XF_OK;
}

int xf_calc_hvv(xf_ctx* ctx, int set, int pt, int hvv) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(set) || bad_set(hvv) || pt < 0 || pt > 2) return xf_fail(XF_EINVAL, "bad arguments");
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_hvv_partial, dim3((p->d.nth + 63) / 64), dim3(64), 0, 0, p->d, p->hsum[set], pt, p->part);
    hipLaunchKernelGGL(k_hvv_final, dim3(1), dim3(256), 0, 0, p->d, p->part, pt, p->hvv[hvv]);
    HIPCHK(hipGetLastError());
    HIPCHK(hipDeviceSynchronize());
    return XF_OK;
}

int xf_get_hvv(xf_ctx* ctx, int hvv, double* out) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(hvv) || !out) return xf_fail(XF_EINVAL, "bad arguments");
    HIPCHK(hipMemcpy(out, p->hvv[hvv], 4 * (size_t)p->d.nbeam * 8, hipMemcpyDeviceToHost));
    return XF_OK;
}

int xf_evolve(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(in) || bad_set(out) || bad_set(hvv)) return xf_fail(XF_EINVAL, "bad arguments");
    return launch_nbeam<K_EVOLVE>(p, p->psi[in], p->psi[out], nullptr, pt, p->hvv[hvv], hm, nullptr, nullptr);
}

int xf_evolve_hvv(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(in) || bad_set(out) || bad_set(hvv)) return xf_fail(XF_EINVAL, "bad arguments");
    return launch_nbeam<K_HVV>(p, p->psi[in], p->psi[out], nullptr, pt, p->hvv[hvv], hm, p->hsum[out], nullptr);
}

int xf_evolve_avg(xf_ctx* ctx, int in, int pt, int hvv, double hm, int out, int avg) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(in) || bad_set(out) || bad_set(hvv) || bad_set(avg)) return xf_fail(XF_EINVAL, "bad arguments");
    return launch_nbeam<K_AVG>(p, p->psi[in], p->psi[out], p->psi[avg], pt, p->hvv[hvv], hm, nullptr, nullptr);
}

int xf_evolve_avg_inplace(xf_ctx* ctx, int pt, int hvv, double hm, int io, int avg) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(io) || bad_set(hvv) || bad_set(avg)) return xf_fail(XF_EINVAL, "bad arguments");
    return launch_nbeam<K_AVG_INPLACE>(p, p->psi[io], p->psi[io], p->psi[avg], pt, p->hvv[hvv], hm, nullptr, nullptr);
}

int xf_avg_beam(xf_ctx* ctx, int in, int out) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(in) || bad_set(out)) return xf_fail(XF_EINVAL, "bad arguments");
    return launch_nbeam<K_ADDAVG>(p, p->psi[in], p->psi[out], nullptr, 0, nullptr, 0., nullptr, nullptr);
}

int xf_calc_err_hvv(xf_ctx* ctx, int a, int b, double* max_err) {
    XfPhy* p;
    int rc = get_phy(ctx, &p);
    if (rc) return rc;
    if (bad_set(a) || bad_set(b) || !max_err) return xf_fail(XF_EINVAL, "bad arguments");
    rc = launch_nbeam<K_ERR>(p, p->psi[a], p->psi[b], nullptr, 0, nullptr, 0., p->hsum[a], p->eps);
    if (rc) return rc;
    std::vector<double> eps(4 * (size_t)p->d.nbeam);
    HIPCHK(hipMemcpy(eps.data(), p->eps, eps.size() * 8, hipMemcpyDeviceToHost));
    double m = -1.;
    for (double v : eps) m = (v > m) ? v : m;
    *max_err = m;
    return XF_OK;
}

}  // extern "C"

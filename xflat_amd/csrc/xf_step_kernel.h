#pragma once
// xf_step_kernel.h -- the fused multi-angle evolution step for CDNA4 (gfx950).
// Included by the xf_k_*.hip translation units, each instantiating a subset.
//
// One persistent launch (every workgroup resident) attempts up to max_steps radial steps of
// nmr::evolutionLoop (src/Nmr.cpp:125-252) for one or many independent
// instances.  Each instance owns a contiguous range of workgroups; a
// workgroup owns "tiles" of whole beams (theta[, phi]) with all their energy
// bins, one energy cell per lane slot, so per-beam energy sums stay inside
// the workgroup.  A step is three passes over the cells:
//
//   A: psi1 = U(h0, dl1) psi0, psi3 = U(h0, dl0) psi0   -> moments of S(psi1), S(psi3)
//   B: psi1 <- (U(h1, dl1) psi0 + psi1) / 2,
//      psi4 = U(h3, dl1) psi0                             -> moments of S(psi4)
//   C: psi3 <- U(h4, dl2) psi3, psi4 <- (psi3 + psi4) / 2,
//      err = sum_e fv |psi4 - psi1|^2 per beam            -> moments of S(psi4), max err
//
// (S1..S11 of Nmr.cpp:177-206).  Between passes each workgroup reduces its
// angular-moment partials (2 moments x 3 density components for MA,
// Phy_MA.cpp:388-446; 4 x 3 for MAA, Phy_MAA.cpp:468-562) in a fixed order,
// publishes them as a tagged row, and every workgroup gathers all rows and
// sums them in the same fixed order -- so all workgroups hold bit-identical
// moments, error, accept/reject (Nmr.cpp:227) and dr (Nmr.cpp:247-249)
// without another sync.
//
// State: structure-of-arrays planes psi[species*4 + comp][cell], cell =
// beam*NE + e, double or float.  With 2 species the nu_x / anti-nu_x columns
// are the SU(2) partners (-conj b, conj a) of nu_e / anti-nu_e and are not
// stored.  Modes: 0 keeps psi0 and the intermediate states in registers for
// the whole launch (state never touches HBM), 1 keeps only psi0 in registers
// and recomputes the intermediates in pass C, 2 streams psi0 from HBM in
// every pass (any size) and writes the accepted candidate to the other buffer,
// 4 streams like 2 with the beam records of the workgroup's tiles kept in
// LDS (for the whole step when they fit, else in chunks of tiles), so the
// tiles of a chunk need no set-up or barrier of their own.
//
// Precision: FP64 forms NBeam::U (NBeam.cpp:360-393) as cos(x) I -
// i sin(x)/lambda H, x = lambda dl, with cos x and sin(x)/lambda from their
// Taylor series in x^2 below x^2 = 1/16 (to x^16, and to x^8 / x^9 below
// x^2 = 2^-10; prop_ds) and from sincos above; where one Hamiltonian acts
// over dl and 2 dl the longer step comes from the double-angle identities
// (prop_d2).  The FP32 modes apply U - I in FP32 by FMA chains (delta_acc),
// with cos x - 1 and sin(x)/x - 1 as polynomials in x^2 (prop_f2s: a
// wave-uniform linear tier when every lane has x^2 < 2^-10, Chebyshev
// cubics below x^2 = 1, hardware sin/cos of the half angle above), and
// estimate the step error from second-order remainders plus the first-order
// term -i Hd psi0 with Hd combined per beam in FP64 (pass C); moments, error
// sums and the decision are FP64.
#include <hip/hip_runtime.h>
#include <type_traits>
#include <cstdint>
#include <cstdio>

#include "xf_internal.h"

using XfStepFn = void (*)(const XfInstDev*, const int*, int);

namespace {

constexpr double INV_2PI_D = 0.15915494309189535;
constexpr uint64_t BARRIER_TIMEOUT_TICKS = 200000000ull;   // 2 s at 100 MHz

template <class T> struct Sp { T ar, ai, br, bi; };

// ---------------------------------------------------------------- FP64 path
struct PropD { double m0, m1, m2, m3; };

// lambda^2 = |H|^2 of the traceless 2x2 Hamiltonian (NBeam.cpp:366)
__device__ __forceinline__ double lam_d(double h0, double h1, double h2) {
    return h0 * h0 + (h1 * h1 + h2 * h2);
}
// U = cos(x) I - i sin(x)/lambda H with x = lambda dl (NBeam.cpp:360-393).
// For x^2 < 1/16 both coefficients come from their Taylor series in x^2 =
// lambda^2 dl^2 (cos x = 1 + sum_k (-1)^k x^2k/(2k)!, sin(x)/lambda = dl sum_k
// (-1)^k x^2k/(2k+1)!, truncated after x^16: remainder < 1e-23): FMAs only, no
// square root, division or transcendental call, which on CDNA4 (FP64 FMA at
// the FP32 rate) is the cheap path.  Larger angles use sqrt/sincos/divide.
__device__ __forceinline__ void prop_ds(double lam2, double dl, double& c, double& sc) {
    const double u = lam2 * (dl * dl);
    if (u < 0x1p-10) {
        // small angles (the common case at small dr): the same series
        // truncated after x^8 / x^9; the first omitted terms, u^5/10! on
        // cos x and u^5/11! on sin(x)/x, are below 2.5e-22
        double sn = 1.0 / 362880.0;                        // 1/9!
        sn = fma(u, -sn, 1.0 / 5040.0);
        sn = fma(u, -sn, 1.0 / 120.0);
        sn = fma(u, -sn, 1.0 / 6.0);
        double cs = 1.0 / 40320.0;                         // 1/8!
        cs = fma(u, -cs, 1.0 / 720.0);
        cs = fma(u, -cs, 1.0 / 24.0);
        cs = fma(u, -cs, 0.5);
        c = fma(-u, cs, 1.0);
        sc = dl * fma(-u, sn, 1.0);
    } else if (u < 0.0625) {
        double sn = 1.0 / 355687428096000.0;               // 1/17!
        sn = fma(u, -sn, 1.0 / 1307674368000.0);           // 1/15!
        sn = fma(u, sn, -1.0 / 6227020800.0);              // 1/13!
        sn = fma(u, sn, 1.0 / 39916800.0);                 // 1/11!
        sn = fma(u, sn, -1.0 / 362880.0);                  // 1/9!
        sn = fma(u, sn, 1.0 / 5040.0);                     // 1/7!  (sign folded below)
        sn = fma(u, -sn, 1.0 / 120.0);
        sn = fma(u, -sn, 1.0 / 6.0) ;
        // sn now holds 1/3! - u/5! + u^2/7! - ...; sinc = 1 - u sn
        double cs = 1.0 / 20922789888000.0;                // 1/16!
        cs = fma(u, -cs, 1.0 / 87178291200.0);             // 1/14!
        cs = fma(u, cs, -1.0 / 479001600.0);               // 1/12!
        cs = fma(u, cs, 1.0 / 3628800.0);                  // 1/10!
        cs = fma(u, cs, -1.0 / 40320.0);                   // 1/8!
        cs = fma(u, cs, 1.0 / 720.0);                      // 1/6!  (sign folded below)
        cs = fma(u, -cs, 1.0 / 24.0);
        cs = fma(u, -cs, 0.5);
        // cs = 1/2! - u/4! + u^2/6! - ...; cos x = 1 - u cs
        c = fma(-u, cs, 1.0);
        sc = dl * fma(-u, sn, 1.0);
    } else {
        double lam = sqrt(lam2);
        double ldr = lam * dl;
        double s;
        sincos(ldr, &s, &c);
        sc = s / lam;
    }
}
__device__ __forceinline__ PropD pd_of(double h0, double h1, double h2, double c, double sc) {
    PropD P;
    P.m0 = c;
    P.m1 = -h0 * sc;
    P.m2 = h2 * sc;
    P.m3 = -h1 * sc;
    return P;
}
__device__ __forceinline__ void prop_d(double h0, double h1, double h2, double lam2, double dl, PropD& P) {
    double c, sc;
    prop_ds(lam2, dl, c, sc);
    P = pd_of(h0, h1, h2, c, sc);
}
// U over twice the path length from U over dl (the double-angle identities
// cos 2x = 2 cos^2 x - 1, sin(2x)/lambda = 2 cos x sin(x)/lambda), for the
// sub-steps dr / c0 = 2 (dr / 2) / c0 of the multi-angle and single-angle
// geometries
__device__ __forceinline__ void prop_d2(double h0, double h1, double h2, double lam2, double dl, PropD& P2, PropD& P) {
    double c, sc;
    prop_ds(lam2, dl, c, sc);
    P = pd_of(h0, h1, h2, c, sc);
    P2 = pd_of(h0, h1, h2, fma(2.0 * c, c, -1.0), 2.0 * (sc * c));
}
__device__ __forceinline__ Sp<double> apply_d(const PropD& P, const Sp<double>& a) {
    Sp<double> o;
    o.ar = (P.m0 * a.ar - P.m1 * a.ai) + (P.m2 * a.br - P.m3 * a.bi);
    o.ai = (P.m0 * a.ai + P.m1 * a.ar) + (P.m2 * a.bi + P.m3 * a.br);
    o.br = (-P.m2 * a.ar - P.m3 * a.ai) + (P.m0 * a.br + P.m1 * a.bi);
    o.bi = (-P.m2 * a.ai + P.m3 * a.ar) + (P.m0 * a.bi - P.m1 * a.br);
    return o;
}

// ---------------------------------------------------------------- FP32 path
// (U - I) = [[cm1 + i m1, m2 + i m3], [-m2 + i m3, cm1 - i m1]]
struct PropF { float cm1, m1, m2, m3; };

// The FP32 propagator U - I and the coefficients of its remainder
// R = (U - I) + i dl H, i.e.
// (cos x - 1) I - i (sin(x)/lambda - dl) H, second order in x = lambda dl.
// For x^2 < 1 all coefficients come from polynomials in x^2
// (cos x - 1 = -x^2/2 + x^4/24 - ..., sin(x)/x - 1 = -x^2/6 + x^4/120 - ...):
// each is rounded relative to its own size, the remainders to x^2 rather than
// to x, and no square root or transcendental is needed.  Larger angles take
// the hardware sin/cos of the half angle.
__device__ __forceinline__ void prop_f2s(float h0, float h1, float h2, float dl, float& cm1, float& sc, float& scr) {
    const float lam2 = h0 * h0 + (h1 * h1 + h2 * h2);
    const float x2 = lam2 * (dl * dl);
    if (__all(x2 < 0x1p-10f)) {
        // every lane of the wave at a small angle (a wave-uniform branch):
        // the same functions to first order in t, whose omitted t^2/720 and
        // t^2/5040 are below 2.7e-9 and 1.2e-9 relative for t < 2^-10
        const float pc = fmaf(x2, 0.0416666679f, -0.5f);
        const float ps = fmaf(x2, 0.00833333377f, -0.166666672f);
        cm1 = x2 * pc;
        const float S = x2 * ps;
        scr = dl * S;
        sc = dl + scr;
    } else if (x2 < 1.f) {
        // (cos x - 1)/t and (sin x / x - 1)/t as cubics in t = x^2: the
        // degree-3 Chebyshev interpolants on t in [0, 1] of their Taylor
        // series (Chebyshev economization), coefficients rounded to FP32.
        // Evaluated by FMA Horner they stay within 5e-8 and 8.2e-8 relative
        // of the exact functions over the whole range, i.e. at FP32 rounding,
        // so the remainder coefficients (scr) keep their accuracy, which the
        // hardware sin/cos path cannot give after the cancellation sc - dl.
        const float pc = fmaf(x2, fmaf(x2, fmaf(x2, 2.42562728e-05f, -0.00138854922f), 0.0416665971f), -0.5f);
        const float ps = fmaf(x2, fmaf(x2, fmaf(x2, 2.70607666e-06f, -0.000198381764f), 0.00833332725f), -0.166666672f);
        cm1 = x2 * pc;
        const float S = x2 * ps;
        scr = dl * S;
        sc = dl + scr;
    } else {
        const float rl = __builtin_amdgcn_rsqf(lam2);
        const float xr = lam2 * rl * dl;
        float x = xr * (float)(0.5 * INV_2PI_D);            // half angle, revolutions
        x = x - __builtin_rintf(x);
        const float sh = __builtin_amdgcn_sinf(x);
        const float ch = __builtin_amdgcn_cosf(x);
        cm1 = -2.f * sh * sh;
        sc = 2.f * sh * ch * rl;
        scr = sc - dl;
    }
}
// the coefficient set (cm1, -i c H) of U - I or of a remainder
__device__ __forceinline__ PropF pf_of(float h0, float h1, float h2, float cm1, float c) {
    return PropF{cm1, -h0 * c, h2 * c, -h1 * c};
}
// Double angle.  Where one Hamiltonian acts over dl and over 2 dl (the first
// two sub-steps of a step of the multi-angle and single-angle geometries,
// dr / c0 and (dr / 2) / c0, which differ by exactly a factor of 2),
// U(2 dl) = U(dl)^2, so with D = U(dl) - I
//   U(2 dl) - I = 2 D + D^2,
//   cos 2x - 1 = 2 cm1 (cm1 + 2),  sin(2x)/lambda = 2 sc (1 + cm1),
// and the difference of the remainders R(dl) - R(2 dl) has the coefficients
//   cm1 - (cos 2x - 1) = -cm1 (2 cm1 + 3),
//   scr - (sin(2x)/lambda - 2 dl) = -(scr + 2 sc cm1),
// both second order in x with terms of one sign (no cancellation).
__device__ __forceinline__ PropF pf_rdiff2(float h0, float h1, float h2, float cm1, float sc, float scr) {
    return pf_of(h0, h1, h2, -cm1 * fmaf(2.f, cm1, 3.f), -fmaf(2.f * sc, cm1, scr));
}
// c + (U - I) a with one FMA chain per component: the four products are
// accumulated onto c (four instructions a component, where the paired
// products and their sum take five).  With c = a this is U a itself; with
// c = 0 the increment.
__device__ __forceinline__ Sp<float> delta_acc(const PropF& P, const Sp<float>& a, const Sp<float>& c) {
    Sp<float> o;
    o.ar = fmaf(P.cm1, a.ar, fmaf(-P.m1, a.ai, fmaf(P.m2, a.br, fmaf(-P.m3, a.bi, c.ar))));
    o.ai = fmaf(P.cm1, a.ai, fmaf(P.m1, a.ar, fmaf(P.m2, a.bi, fmaf(P.m3, a.br, c.ai))));
    o.br = fmaf(-P.m2, a.ar, fmaf(-P.m3, a.ai, fmaf(P.cm1, a.br, fmaf(P.m1, a.bi, c.br))));
    o.bi = fmaf(-P.m2, a.ai, fmaf(P.m3, a.ar, fmaf(P.cm1, a.bi, fmaf(-P.m1, a.br, c.bi))));
    return o;
}
// q + sg (cm1, -i scr H): a remainder's coefficient set (pf_of(h, cm1, scr))
// added (sg = 1) or subtracted (sg = -1) with one FMA per component instead
// of its three products and four sums
template <int SG>
__device__ __forceinline__ PropF pf_fma(const PropF& q, float h0, float h1, float h2, float cm1, float scr) {
    if constexpr (SG > 0) return PropF{q.cm1 + cm1, fmaf(-h0, scr, q.m1), fmaf(h2, scr, q.m2), fmaf(-h1, scr, q.m3)};
    else return PropF{q.cm1 - cm1, fmaf(h0, scr, q.m1), fmaf(-h2, scr, q.m2), fmaf(h1, scr, q.m3)};
}
// (U - I) is linear in its coefficients: sums of coefficient sets apply as one.
__device__ __forceinline__ PropF pf_add(const PropF& p, const PropF& q) {
    return PropF{p.cm1 + q.cm1, p.m1 + q.m1, p.m2 + q.m2, p.m3 + q.m3};
}
template <class T> __device__ __forceinline__ Sp<T> sp_add(const Sp<T>& a, const Sp<T>& b) {
    return Sp<T>{a.ar + b.ar, a.ai + b.ai, a.br + b.br, a.bi + b.bi};
}
// (new + old) * .5, the averaging of NBeam.cpp:507-510 and :730-733
template <class T> __device__ __forceinline__ Sp<T> sp_avg(const Sp<T>& nw, const Sp<T>& old) {
    return Sp<T>{(nw.ar + old.ar) * T(.5), (nw.ai + old.ai) * T(.5), (nw.br + old.br) * T(.5), (nw.bi + old.bi) * T(.5)};
}
__device__ __forceinline__ Sp<float> sp_f(const Sp<double>& a) {
    return Sp<float>{(float)a.ar, (float)a.ai, (float)a.br, (float)a.bi};
}
__device__ __forceinline__ Sp<float> sp_f(const Sp<float>& a) { return a; }
__device__ __forceinline__ Sp<double> sp_d(const Sp<float>& a) {
    return Sp<double>{(double)a.ar, (double)a.ai, (double)a.br, (double)a.bi};
}
__device__ __forceinline__ Sp<double> sp_d(const Sp<double>& a) { return a; }

// density matrix components 0, 2, 3 (NBeam.cpp:314-322)
template <class T> __device__ __forceinline__ void dens(const Sp<T>& a, T& d0, T& d2, T& d3) {
    d0 = T(.5) * ((a.ar * a.ar + a.ai * a.ai) - (a.br * a.br + a.bi * a.bi));
    d2 = (a.ar * a.br) - (a.ai * (-a.bi));
    d3 = (a.ar * (-a.bi)) + (a.ai * a.br);
}
template <class T> __device__ __forceinline__ T dnorm2(const Sp<T>& x, const Sp<T>& y) {
    T dar = x.ar - y.ar, dai = x.ai - y.ai, dbr = x.br - y.br, dbi = x.bi - y.bi;
    return (dar * dar + dai * dai) + (dbr * dbr + dbi * dbi);
}

// Tagged row words of the cross-workgroup exchange, 64-bit single-copy-atomic
// stores and loads through the global address space (global, not flat,
// instructions).  The tag is the exchange's generation + 1; a slot holds the
// word of its current exchange or of the one two before it (rows[gen & 1]).
// Two formats:
//  - two words per value (W1 = false, the FP64 stepper, whose moments enter
//    the FP64 parity with the reference unrounded): the value's low and high
//    32-bit halves, each under the 32-bit tag;
//  - one word per value (W1 = true, the FP32-increment steppers): the value
//    rounded to the nearest multiple of 8 units in its last place (relative
//    change below 2^-50, far below their FP32 state), its three low mantissa
//    bits replaced by the tag's low three bits.  Half the words to store and
//    fetch.  A step's exchanges carry 2 NMV, NMV and NMV + 1 values, so a
//    word that only some of them write was last written 2, 4 or 6
//    generations before (never a multiple of 8), and a cleared word (bits 0)
//    is never read before the first write of its slot.
using gword = __attribute__((address_space(1))) unsigned long long;
__device__ __forceinline__ unsigned long long ld_word(const unsigned long long* p) {
    return __hip_atomic_load((const gword*)p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void st_word(unsigned long long* p, unsigned long long w) {
    __hip_atomic_store((gword*)p, w, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}
template <bool W1> constexpr int words_per_value() { return W1 ? 1 : 2; }
// word q of value x
template <bool W1> __device__ __forceinline__ unsigned long long word_of(double x, unsigned tag, int q) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(x);
    if constexpr (W1) return ((u + 4ull) & ~7ull) | (unsigned long long)(tag & 7u);
    else return ((unsigned long long)tag << 32) | (q ? (u >> 32) : (u & 0xffffffffull));
}
// the tag bits of a word (for a 32-bit load, the half that holds them)
template <bool W1> __device__ __forceinline__ bool tag_current(unsigned long long w, unsigned tag) {
    if constexpr (W1) return ((unsigned)w & 7u) == (tag & 7u);
    else return (unsigned)(w >> 32) == tag;
}
template <bool W1> __device__ __forceinline__ unsigned tag_bits(unsigned long long w) {
    return W1 ? ((unsigned)w & 7u) : (unsigned)(w >> 32);
}
// the value from its words w0 (and w1)
template <bool W1> __device__ __forceinline__ double word_value(unsigned long long w0, unsigned long long w1) {
    if constexpr (W1) return __longlong_as_double((long long)(w0 & ~7ull));
    else return __longlong_as_double((long long)((w1 << 32) | (w0 & 0xffffffffull)));
}
// Bounded-wait bookkeeping: 2 s of the 100 MHz real-time counter, or an
// error flagged by another workgroup.  Returns 0 while the wait may go on,
// 1 when this wait expired (the status is set), 2 when another workgroup's
// error ends it.
__device__ __forceinline__ int wait_expired(uint64_t t0, int spins, int* status) {
    if ((spins & 15) != 0) return 0;
    if (__hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) return 2;
    if (__builtin_amdgcn_s_memrealtime() - t0 > BARRIER_TIMEOUT_TICKS) {
        __hip_atomic_store(status, (int)XF_EBARRIER, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return 1;
    }
    return 0;
}
// After this wave's own wait expired: what it saw in the word array w of the
// n rows it waited on (the rows whose tag is not `tag`, and the tag the first
// of them holds), kept by the first workgroup to get here.  One full wave.
template <bool W1>
__device__ __noinline__ void note_expiry(XfRunScalars* rs, int where, int lb, unsigned tag,
                                         const unsigned long long* w, int n) {
    const int lane = threadIdx.x & 63;
    int missing = 0, first = n;
    for (int j0 = 0; j0 < n; j0 += 64) {
        const int j = j0 + lane;
        const bool bad = j < n && !tag_current<W1>(ld_word(w + min(j, n - 1)), tag);
        const unsigned long long m = __ballot(bad);
        missing += __popcll(m);
        if (m && first == n) first = j0 + __ffsll((long long)m) - 1;
    }
    const unsigned seen = first < n ? tag_bits<W1>(ld_word(w + first)) : tag;
    if (lane == 0) {
        int expect = 0;
        if (__hip_atomic_compare_exchange_strong(&rs->diag_where, &expect, where, __ATOMIC_RELAXED, __ATOMIC_RELAXED,
                                                 __HIP_MEMORY_SCOPE_AGENT)) {
            rs->diag_wg = lb;
            rs->diag_tag = tag;
            rs->diag_seen = seen;
            rs->diag_missing = missing;
            rs->diag_row = first;
        }
    }
}

// The run scalars into the pinned host copy, one 64-bit word at a time (one
// thread; a struct temporary would live in scratch memory).
static_assert(sizeof(XfRunScalars) % 8 == 0, "run scalars: whole 64-bit words");
__device__ __forceinline__ void mirror_scalars(const XfRunScalars* src, XfRunScalars* dst) {
    constexpr int NW = (int)(sizeof(XfRunScalars) / 8);
    const unsigned long long* s = reinterpret_cast<const unsigned long long*>(src);
    unsigned long long* d = reinterpret_cast<unsigned long long*>(dst);
    unsigned long long w[NW];
#pragma unroll
    for (int i = 0; i < NW; ++i) w[i] = s[i];
#pragma unroll
    for (int i = 0; i < NW; ++i) d[i] = w[i];
}

// matt::getHm (Matt.cpp:31-64)
__device__ __forceinline__ double hm_at(const XfInstDev& I, double r) {
    if (!I.has_matter) return 0.;
    double R = 10. / r;
    double far = I.nbfar_coeff * (R * R * R);
    double near = I.nb0 * exp((I.Rv - r) / I.hNS);
    return I.mat_coeff * (far + near) * .5;
}

// Per-beam record in LDS (doubles).
template <int GEOM> struct BRec {
    static constexpr int NM = (GEOM == XF_MAA) ? 4 : (GEOM == XF_SA) ? 1 : 2;
    static constexpr int DL = 0;                  // dl[3]
    static constexpr int W = 3;                   // W[3 pts][NM] moment weights
    static constexpr int G = W + 3 * NM;          // G[3 pts][3] hvv geometry
    static constexpr int H = G + 9;               // Hb[4][3]: h0, h1, h3, h4 (+matter)
    static constexpr int HD = H + 12;             // dl0 Hb0 + dl2 Hb4 + dl1 (Hb3 - Hb0 - Hb1), dl0 + dl2 - dl1
    // FP32 copies for the FP32-increment steppers: Hb[4][3] in floats 0..11,
    // dl[3] in floats 12..14, Hd[4] in floats 16..19 (a cell then forms its
    // Hamiltonians with two FP32 adds instead of two FP64 adds and three
    // conversions per propagator)
    static constexpr int HF = HD + 4;
    static constexpr int SIZE = HF + 10;
};

// Geometry of one beam at one of the step's points r, r+dr/2, r+dr (pt 0,
// 1, 2): angle bins (Phy_MA.cpp:172-196, Phy_MAA.cpp:194-210), moment
// weights and the H_vv geometry; the point's cos(theta_R) is stored in the
// geometry slot, from which beam_dl forms the path lengths.
// SA (one trajectory): the moment is the source itself (weight 1), the
// geometry slot holds D(r) = (1 - sqrt(1 - (Rv/r)^2))^2 / 2 (Phy_SA.cpp:250-256).
template <int GEOM>
__device__ __forceinline__ void beam_point(const XfInstDev& I, int b, int pt, double rp, double* rec) {
    using R = BRec<GEOM>;
    if constexpr (GEOM == XF_SA) {
        const double Ra = I.Rv / rp;
        const double tm = 1. - sqrt(1. - (Ra * Ra));
        rec[R::W + pt] = 1.;
        rec[R::G + pt * 3 + 0] = .5 * (tm * tm);
        rec[R::G + pt * 3 + 1] = 0.;
        rec[R::G + pt * 3 + 2] = 0.;
        return;
    }
    const int th = (GEOM == XF_MAA) ? b / I.nph : b;
    const double x = I.x[th], dx = I.dx[th];
    if constexpr (GEOM == XF_MAA) {
        const int ph = b - th * I.nph;
        const double ra = I.Rv / rp;
        const double r2 = ra * ra;
        const double c = sqrt(1. - r2 * (1. - x));
        const double s = ra * I.sinR[th];
        const double dc = .5 * (1. / c) * r2 * dx;
        const double cp = I.cphi[ph], spp = I.sphi[ph];
        rec[R::W + pt * 4 + 0] = dc;
        rec[R::W + pt * 4 + 1] = cp * s * dc;
        rec[R::W + pt * 4 + 2] = spp * s * dc;
        rec[R::W + pt * 4 + 3] = dc * c;
        rec[R::G + pt * 3 + 0] = s * cp;
        rec[R::G + pt * 3 + 1] = s * spp;
        rec[R::G + pt * 3 + 2] = c;
    } else {
        double r2 = I.Rv / rp;
        r2 *= r2;
        const double c = sqrt(1. - r2 * (1. - x));
        const double dc = .5 * (1. / c) * r2 * dx;
        rec[R::W + pt * 2 + 0] = dc;
        rec[R::W + pt * 2 + 1] = dc * c;
        rec[R::G + pt * 3 + 0] = c;
        rec[R::G + pt * 3 + 1] = 0.;
        rec[R::G + pt * 3 + 2] = 0.;
    }
}
// Path lengths of the three sub-steps (Phy_MA.cpp:218-227, Phy_MAA.cpp:232-241;
// SA: the sub-step length itself, Phy_SA.cpp:144-147) from the points'
// cos(theta_R) that beam_point stored.
template <int GEOM>
__device__ __forceinline__ void beam_dl(double* rec, double dr, double dr2) {
    using R = BRec<GEOM>;
    if constexpr (GEOM == XF_SA) {
        rec[R::DL + 0] = dr2;
        rec[R::DL + 1] = dr;
        rec[R::DL + 2] = dr2;
    } else if constexpr (GEOM == XF_MAA) {
        const double c0 = rec[R::G + 2], c1 = rec[R::G + 5], c2 = rec[R::G + 8];
        rec[R::DL + 0] = dr2 / ((c0 + c1) * .5);
        rec[R::DL + 1] = dr / ((c0 + c2) * .5);
        rec[R::DL + 2] = dr2 / ((c1 + c2) * .5);
    } else {
        const double c0 = rec[R::G + 0], c1 = rec[R::G + 3];
        rec[R::DL + 0] = dr2 / c0;
        rec[R::DL + 1] = dr / c0;
        rec[R::DL + 2] = dr2 / c1;
    }
    float* f = reinterpret_cast<float*>(rec + R::HF) + 12;
    f[0] = (float)rec[R::DL + 0];
    f[1] = (float)rec[R::DL + 1];
    f[2] = (float)rec[R::DL + 2];
}

// H_vv of one beam at point pt from the moments M[m*3+z] (z: density comps
// 0,2,3): Phy_MA.cpp:372-377 / Phy_MAA.cpp:445-461; plus the matter term on
// component 0 (NBeam.cpp:604); into Hamiltonian slot `slot` of the record
// (0..3: h0, h1, h3, h4) and its FP32 copy.
template <int GEOM>
__device__ __forceinline__ void beam_h(double* rec, int pt, int slot, const double* M, double coef, double hm) {
    using R = BRec<GEOM>;
    const double* G = rec + R::G + pt * 3;
    double* Hb = rec + R::H + slot * 3;
    double hv[3];
#pragma unroll
    for (int z = 0; z < 3; ++z) {
        if constexpr (GEOM == XF_MAA)
            hv[z] = (M[z] - (G[0] * M[3 + z] + G[1] * M[6 + z] + G[2] * M[9 + z])) * coef;
        else if constexpr (GEOM == XF_SA)
            hv[z] = M[z] * (coef * G[0]);             // hvv *= integ_coef * D(r), Phy_SA.cpp:280-281
        else
            hv[z] = (M[z] - (M[3 + z] * G[0])) * coef;
    }
    Hb[0] = hv[0] + hm;
    Hb[1] = hv[1];
    Hb[2] = hv[2];
    float* f = reinterpret_cast<float*>(rec + R::HF) + slot * 3;
    f[0] = (float)Hb[0];
    f[1] = (float)Hb[1];
    f[2] = (float)Hb[2];
}

// First-order combination of the step's four Hamiltonians that enters
// psi4 - psi1 (see the compensated error of the mixed stepper): per beam the
// part of sum_k c_k dl_k H_k that does not depend on the energy, and the
// coefficient sum_k c_k dl_k of the vacuum part.  FP64, after h4 is known.
// Pass A's set-up in one phase for the geometries whose path lengths come
// from single points (multi-angle: dr/2 / c0, dr / c0, dr/2 / c1; single
// angle: the sub-step lengths): the thread of point pt forms the point, then
// the path lengths that divide by its own cos(theta_R) and, at point 0, the
// Hamiltonian h0 -- the same operations as beam_point, beam_dl and beam_h,
// without a barrier between them.
template <int GEOM>
__device__ __forceinline__ void beam_point_a(const XfInstDev& I, int b, int pt, double rp, double* rec, double dr,
                                             double dr2, const double* M, double coef, double hm) {
    static_assert(GEOM != XF_MAA, "MAA path lengths average two points");
    using R = BRec<GEOM>;
    beam_point<GEOM>(I, b, pt, rp, rec);
    float* f = reinterpret_cast<float*>(rec + R::HF) + 12;
    if constexpr (GEOM == XF_SA) {
        if (pt == 0) {
            rec[R::DL + 0] = dr2;
            rec[R::DL + 1] = dr;
            rec[R::DL + 2] = dr2;
            f[0] = (float)dr2;
            f[1] = (float)dr;
            f[2] = (float)dr2;
        }
    } else {
        const double c = rec[R::G + pt * 3];
        if (pt == 0) {
            rec[R::DL + 0] = dr2 / c;
            rec[R::DL + 1] = dr / c;
            f[0] = (float)rec[R::DL + 0];
            f[1] = (float)rec[R::DL + 1];
        } else if (pt == 1) {
            rec[R::DL + 2] = dr2 / c;
            f[2] = (float)rec[R::DL + 2];
        }
    }
    if (pt == 0) beam_h<GEOM>(rec, 0, 0, M, coef, hm);
}
template <int GEOM>
__device__ __forceinline__ void beam_hd(double* rec) {
    using R = BRec<GEOM>;
    const double dl0 = rec[R::DL + 0], dl1 = rec[R::DL + 1], dl2 = rec[R::DL + 2];
    const double* B0 = rec + R::H + 0;
    const double* B1 = rec + R::H + 3;
    const double* B3 = rec + R::H + 6;
    const double* B4 = rec + R::H + 9;
#pragma unroll
    for (int z = 0; z < 3; ++z) rec[R::HD + z] = dl0 * B0[z] + dl2 * B4[z] + dl1 * ((B3[z] - B0[z]) - B1[z]);
    rec[R::HD + 3] = (dl0 + dl2) - dl1;
    float* f = reinterpret_cast<float*>(rec + R::HF) + 16;
#pragma unroll
    for (int z = 0; z < 4; ++z) f[z] = (float)rec[R::HD + z];
}

// Per-cell data of one thread.
template <class ST, class CT, int NS, int CPT> struct Cells {
    static constexpr bool F64 = std::is_same<CT, double>::value;
    int bl[CPT];         // beam within the tile, -1 if the slot is empty
    int e[CPT];
    Sp<ST> p0[CPT][NS];  // psi0 (resident modes)
    Sp<CT> k1[CPT][NS];  // psi1 (FP64) / its increment (FP32)
    Sp<CT> k3[CPT][NS];  // psi3
    Sp<CT> k4[CPT][NS];  // psi4
};

template <class ST>
__device__ __forceinline__ Sp<ST> ld_sp(const ST* __restrict__ base, long long ncells, int s, long long cell) {
    const ST* p = base + (long long)(s * 4) * ncells + cell;
    return Sp<ST>{p[0], p[ncells], p[2 * ncells], p[3 * ncells]};
}
template <class ST>
__device__ __forceinline__ void st_sp(ST* __restrict__ base, long long ncells, int s, long long cell, const Sp<double>& v) {
    ST* p = base + (long long)(s * 4) * ncells + cell;
    p[0] = (ST)v.ar;
    p[ncells] = (ST)v.ai;
    p[2 * ncells] = (ST)v.br;
    p[3 * ncells] = (ST)v.bi;
}

// Per-energy table in LDS, one record of tab_slots(NS) 8-byte slots per
// energy bin: H_vac (h_r0, h_r1), the error weights fv of the 4 particles,
// the NS source weights (FP32 ones in the low half of their slot), and for
// the FP32-increment steppers H_vac as two floats in slot 6 + NS (the
// padding slot: NS is 2 or 4).  A cell reads every field at an immediate
// offset from one 32-bit address.  The record length is odd in 8-byte units,
// so the lanes of a wave, which hold consecutive energy bins, read 32
// distinct 8-byte banks in each half wave (an even length would fold them
// onto a few banks).
constexpr int tab_slots(int ns) { return (6 + ns) | 1; }

// Hamiltonian of species s at energy e for beam H record Hb (NBeam.cpp:602-625):
// nu: (hv0 + Hb0, hv1 + Hb1, Hb2); anti-nu: (hv0 - Hb0, hv1 - Hb1, Hb2).
// te: the energy's table record.
__device__ __forceinline__ void cell_h(const double* te, int s, const double* Hb, double& h0, double& h1,
                                       double& h2) {
    double hv0 = te[0], hv1 = te[1];
    if (s & 1) {
        h0 = hv0 + (-Hb[0]);
        h1 = hv1 + (-Hb[1]);
    } else {
        h0 = hv0 + Hb[0];
        h1 = hv1 + Hb[1];
    }
    h2 = Hb[2];
}

// The same in FP32 for the FP32-increment steppers, from the FP32 copies of
// H_vac (tef, in the energy's table record) and of Hb (Hf, in the beam's
// record).  The propagators take FP32 Hamiltonians in any case; the error
// estimate's first-order term keeps the FP64 ones (beam_hd).
__device__ __forceinline__ void cell_hf(const float* tef, int s, const float* Hf, float& h0, float& h1, float& h2) {
    const float hv0 = tef[0], hv1 = tef[1];
    if (s & 1) {
        h0 = hv0 + (-Hf[0]);
        h1 = hv1 + (-Hf[1]);
    } else {
        h0 = hv0 + Hf[0];
        h1 = hv1 + Hf[1];
    }
    h2 = Hf[2];
}

// Source contribution of one cell to the angular integral, X(z) for density
// components 0, 2, 3 (upd_nu_coef, NBeam.h:47-53, with L/<E> folded into the
// per-energy source weights).
// The per-energy source weight slot of one species: the FP64 weight, or for
// FP32 cells the FP32 weight and its half (src_add).
template <bool F64>
__device__ __forceinline__ void put_src_weight(double* slot, double w, float wf) {
    if constexpr (F64) {
        *slot = w;
    } else {
        float* f = reinterpret_cast<float*>(slot);
        f[0] = wf;
        f[1] = .5f * wf;
    }
}
// FP32: the slot holds the weight w and w / 2, which takes the density's
// factor 1/2 (d0 = (|a|^2 - |b|^2) / 2 by an FMA chain; w (t / 2) = (w / 2) t
// exactly).
template <class CT>
__device__ __forceinline__ void src_add(const double* te, int s, const Sp<CT>& v, CT& X0, CT& X2, CT& X3) {
    CT d0, d2, d3, w, w0;
    if constexpr (std::is_same<CT, double>::value) {
        dens<CT>(v, d0, d2, d3);
        w = w0 = te[6 + s];
    } else {
        d0 = fmaf(v.ar, v.ar, fmaf(v.ai, v.ai, -fmaf(v.br, v.br, v.bi * v.bi)));
        d2 = (v.ar * v.br) - (v.ai * (-v.bi));
        d3 = (v.ar * (-v.bi)) + (v.ai * v.br);
        const float* wf = reinterpret_cast<const float*>(te + 6 + s);
        w = wf[0];
        w0 = wf[1];
    }
    if (s & 1) { X0 -= w0 * d0; X2 -= w * d2; }
    else       { X0 += w0 * d0; X2 += w * d2; }
    X3 += w * d3;
}

template <int NM>
__device__ __forceinline__ void acc_moments(double* acc, const double* W, double X0, double X2, double X3) {
#pragma unroll
    for (int m = 0; m < NM; ++m) {
        acc[m * 3 + 0] += X0 * W[m];
        acc[m * 3 + 1] += X2 * W[m];
        acc[m * 3 + 2] += X3 * W[m];
    }
}


// psi0 + increment, formed in FP64; with renorm the unitarity of the state
// |a|^2 + |b|^2 = 1 is restored (FP32 modes only).  The norm deviates from 1
// by the FP32 rounding of one increment (~1e-7), so 1/sqrt(1 + d) comes from
// its series 1 - d/2 + 3d^2/8 - 5d^3/16 (remainder < 35 d^4/128 < 3e-17 for
// |d| < 1e-4); a larger deviation takes sqrt and divide.
// The FP32-state candidate: psi0 + t / 2 with t = d3n + d4 (one FMA a
// component), its norm and the renormalisation factor in FP64 (the same
// series), the correction applied by one FP32 FMA per component.
__device__ __forceinline__ Sp<float> candidate_t(const Sp<float>& p0, const Sp<float>& t, int renorm) {
    Sp<float> v{fmaf(.5f, t.ar, p0.ar), fmaf(.5f, t.ai, p0.ai), fmaf(.5f, t.br, p0.br), fmaf(.5f, t.bi, p0.bi)};
    if (renorm) {
        const double n2 = ((double)v.ar * (double)v.ar + (double)v.ai * (double)v.ai) +
                          ((double)v.br * (double)v.br + (double)v.bi * (double)v.bi);
        const double d = n2 - 1.0;
        double sc;
        if (fabs(d) < 1e-4) sc = fma(d, fma(d, fma(d, -0.3125, 0.375), -0.5), 1.0);
        else sc = 1.0 / sqrt(n2);
        const float dm = (float)(sc - 1.0);
        v.ar = fmaf(v.ar, dm, v.ar);
        v.ai = fmaf(v.ai, dm, v.ai);
        v.br = fmaf(v.br, dm, v.br);
        v.bi = fmaf(v.bi, dm, v.bi);
    }
    return v;
}
__device__ __forceinline__ Sp<double> candidate(const Sp<double>& p0, const Sp<float>& df, int renorm) {
    Sp<double> nd{p0.ar + (double)df.ar, p0.ai + (double)df.ai, p0.br + (double)df.br, p0.bi + (double)df.bi};
    if (renorm) {
        const double n2 = (nd.ar * nd.ar + nd.ai * nd.ai) + (nd.br * nd.br + nd.bi * nd.bi);
        const double d = n2 - 1.0;
        double sc;
        if (fabs(d) < 1e-4) sc = fma(d, fma(d, fma(d, -0.3125, 0.375), -0.5), 1.0);
        else sc = 1.0 / sqrt(n2);
        nd.ar *= sc; nd.ai *= sc; nd.br *= sc; nd.bi *= sc;
    }
    return nd;
}

// Wave reductions on the DPP network (VALU moves, no LDS round trip), one
// 32-bit half of a double at a time.  wave_red_dpp leaves the total of all
// 64 lanes in lane 63: rotations by 8, 4, 2, 1 within each row of 16 lanes
// give every lane its row's total, then row 0's and row 2's totals (lanes 15
// and 47) are broadcast into rows 1 and 3, and lane 31's into rows 2 and 3.
// The association order is fixed, so equal inputs give equal bits.  The
// moves leave lanes outside a broadcast's row mask unwritten (no zero
// operand is set up first); those lanes never feed lane 63, the only lane
// whose result is used.
template <int CTRL, int ROWMASK>
__device__ __forceinline__ double dpp_d(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_mov_dpp((int)(unsigned)u, CTRL, ROWMASK, 0xF, false);
    const int hi = __builtin_amdgcn_mov_dpp((int)(unsigned)(u >> 32), CTRL, ROWMASK, 0xF, false);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
template <bool MAX>
__device__ __forceinline__ double dpp_op(double a, double b) { return MAX ? fmax(a, b) : a + b; }
// MAX: maxima of values >= 0.
template <bool MAX>
__device__ __forceinline__ double wave_red_dpp(double v) {
    v = dpp_op<MAX>(v, dpp_d<0x128, 0xF>(v));    // row_ror:8
    v = dpp_op<MAX>(v, dpp_d<0x124, 0xF>(v));    // row_ror:4
    v = dpp_op<MAX>(v, dpp_d<0x122, 0xF>(v));    // row_ror:2
    v = dpp_op<MAX>(v, dpp_d<0x121, 0xF>(v));    // row_ror:1
    v = dpp_op<MAX>(v, dpp_d<0x142, 0xA>(v));    // row_bcast:15 into rows 1, 3
    v = dpp_op<MAX>(v, dpp_d<0x143, 0xC>(v));    // row_bcast:31 into rows 2, 3
    return v;
}
// lane 63's value in every lane
// Sum (MAX: maximum of values >= 0) of p[i] for i = i0, i0 + 64, ... < n, in
// that order, with 8 LDS loads in flight instead of one dependent load per
// term.  The padding terms are +0, which leaves a sum (never -0 here) and a
// maximum of non-negative values unchanged, so the result has the bits of
// the one-at-a-time loop.
template <bool MAX>
__device__ __forceinline__ double lds_strided(const double* p, int i0, int n) {
    double x = 0.;
    for (int i = i0; i < n; i += 8 * 64) {
        double v[8];
#pragma unroll
        for (int u = 0; u < 8; ++u) v[u] = (i + u * 64 < n) ? p[i + u * 64] : 0.;
#pragma unroll
        for (int u = 0; u < 8; ++u) x = MAX ? fmax(x, v[u]) : x + v[u];
    }
    return x;
}
template <int L>
__device__ __forceinline__ double lane_bcast(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readlane((int)(unsigned)u, L);
    const int hi = __builtin_amdgcn_readlane((int)(unsigned)(u >> 32), L);
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ double lane63(double v) { return lane_bcast<63>(v); }
// A wave-uniform value held in scalar registers: the step state (r, dr,
// counters) is read from memory or LDS, which the compiler cannot prove
// uniform; in vector registers it would take ~20 VGPRs for the whole launch.
__device__ __forceinline__ double uni_d(double v) {
    const unsigned long long u = (unsigned long long)__double_as_longlong(v);
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const int hi = __builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return __longlong_as_double((long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo));
}
__device__ __forceinline__ long long uni_ll(long long v) {
    const unsigned long long u = (unsigned long long)v;
    const int lo = __builtin_amdgcn_readfirstlane((int)(unsigned)u);
    const int hi = __builtin_amdgcn_readfirstlane((int)(unsigned)(u >> 32));
    return (long long)(((unsigned long long)(unsigned)hi << 32) | (unsigned)lo);
}

// Cross-workgroup reduction of NV sums and NMAX maxima (maxima of values
// >= 0) by an all-gather of tagged rows:
//  1. the lanes' values are staged in LDS (xcap doubles, in chunks of
//     values when NT * B exceeds it); wave w reduces values w, w + nwaves,
//     ...: lane l sums lanes l, l + 64, ... in order, then the DPP wave
//     reduction, and the wave publishes the total as row lb: value v ->
//     its words (word v, or words 2v and 2v + 1), tagged with gen + 1
//     (word_of), into every replica of rows[gen & 1][replica][word][nblk];
//  2. right after, each wave reads its values (two at a time) for all rows
//     of replica lb % XF_ROW_REPLICAS straight into registers (8 or 16
//     loads per lane in flight) and re-reads only the words whose tag is
//     not yet current until every word is: the wait for the last row and
//     the fetch are the same loads, so an exchange costs one round trip
//     after the last row lands, and only the late rows are read again;
//  3. each wave sums each value over the rows in a fixed order (lane l
//     summing rows l, l + 64, ..., then the DPP reduction), which leaves
//     the same bits in every workgroup.
// Replicas spread the all-gather's reads, which every workgroup makes of the
// same few kilobytes, over more memory lines.  The rows are the only data
// that crosses workgroups, so no fence is needed: each word is one
// single-copy-atomic 64-bit store and its tag names its exchange.  Two
// parities suffice because a workgroup publishes exchange gen + 2 only after
// every row of gen + 1 arrived, and a workgroup publishes gen + 1 only after
// its fetch of gen returned (the published values depend on the fetched
// ones).  All waits are bounded.
// Step 1 (xch_publish) and steps 2-3 (xch_gather_spin, or xch_gather_poll
// for the streaming modes), run back to back by reduce_publish.
template <int NV, int NMAX, bool W1>
__device__ __forceinline__ void xch_publish(const double* acc, double* s_x, int xcap, const XfInstDev& I, int lb,
                                            unsigned gen, long long* tr = nullptr) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int B = blockDim.x;
    const int nblk = I.nblk;
    constexpr int NT = NV + NMAX;
    constexpr int RP = XF_ROW_REPLICAS;
    const unsigned tag = gen + 1u;
    const size_t rep_words = (size_t)XF_ROWW * nblk;
    unsigned long long* rows0 = I.rows + (size_t)(gen & 1u) * RP * rep_words;
    if (tr && tid == 0) tr[0] = (long long)__builtin_amdgcn_s_memrealtime();
    const int vg = max(1, min(NT, xcap / B));          // values staged at once
    for (int v0 = 0; v0 < NT; v0 += vg) {
        const int nv = min(vg, NT - v0);
        if (v0 > 0) __syncthreads();                   // the previous chunk has been read
#pragma unroll
        for (int v = 0; v < NT; ++v)
            if (v >= v0 && v < v0 + nv) s_x[(v - v0) * B + tid] = acc[v];
        __syncthreads();
        // two values per wave at a time, so their load and DPP chains interleave
        for (int c = wave; c < nv; c += 2 * nwaves) {
            const int cc[2] = {c, c + nwaves};
            double x[2];
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = v0 + cc[h];
                const double* src = s_x + (size_t)cc[h] * B;
                x[h] = 0.;
                if (cc[h] < nv) {
                    x[h] = (v < NV) ? lds_strided<false>(src, lane, B) : lds_strided<true>(src, lane, B);
                }
            }
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const int v = v0 + cc[h];
                if (cc[h] >= nv) break;               // wave-uniform
                x[h] = lane63(v < NV ? wave_red_dpp<false>(x[h]) : wave_red_dpp<true>(x[h]));
                constexpr int WV = words_per_value<W1>();
                if (lane >= 64 - WV * RP) {
                    const int q = 63 - lane;                  // replica q / WV, word q % WV
                    st_word(rows0 + (size_t)(q / WV) * rep_words + (size_t)(WV * v + q % WV) * nblk + lb,
                            word_of<W1>(x[h], tag, q % WV));
                }
            }
        }
    }
    if (tr && tid == 0) tr[1] = (long long)__builtin_amdgcn_s_memrealtime();
}

// `idle`: work that needs none of the totals, run by each wave once its
// first loads are in flight.
// Every wave fetches its own values (two at a time) for all rows of its
// replica, and a lane re-reads only the words whose tag is not yet current:
// the wait for the last row and the fetch are one round trip.
template <int NV, int NMAX, bool W1, class Idle>
__device__ __forceinline__ bool xch_gather_spin(double* s_tot, const XfInstDev& I, int lb, unsigned& gen,
                                           int* status, int* s_flag, long long* tr, Idle&& idle) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int nblk = I.nblk;
    constexpr int NT = NV + NMAX;
    constexpr int RP = XF_ROW_REPLICAS;
    const unsigned tag = gen + 1u;
    const unsigned long long tagw = word_of<W1>(0., tag, 0);             // a current word (for absent rows)
    constexpr int WV = words_per_value<W1>();
    const size_t rep_words = (size_t)XF_ROWW * nblk;
    unsigned long long* rows = I.rows + (size_t)(gen & 1u) * RP * rep_words + (size_t)(lb % RP) * rep_words;
    // failure flags alternate with the generation: the next exchange's is
    // cleared here (it was last read before this exchange began)
    if (tid == 0) s_flag[2 + ((gen + 1u) & 1u)] = 0;
    const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
    int wfail = 0;
    bool idled = false;
    for (int v0 = wave; v0 < NT; v0 += 2 * nwaves) {
        const int v1 = v0 + nwaves;
        const bool two = v1 < NT;                             // wave-uniform
        double x[2] = {0., 0.};
        for (int j0 = 0; j0 < nblk && !wfail; j0 += 256) {
            unsigned long long w[2][WV][4];                   // [value][word][row]
#pragma unroll
            for (int h = 0; h < 2; ++h)
#pragma unroll
                for (int q = 0; q < WV; ++q)
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int j = j0 + lane + 64 * u;
                        const int v = h ? v1 : v0;
                        w[h][q][u] = ((h == 0 || two) && j < nblk) ? ld_word(rows + (size_t)(WV * v + q) * nblk + j) : tagw;
                    }
            if (!idled) {
                idle();
                idled = true;
            }
            for (int spins = 1;; ++spins) {
                int bad = 0;
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < WV; ++q)
#pragma unroll
                        for (int u = 0; u < 4; ++u) bad |= !tag_current<W1>(w[h][q][u], tag);
                if (!__any(bad)) break;
                if (const int e = wait_expired(t1, spins, status)) {
                    if (e == 1) note_expiry<W1>(I.rs, 2, lb, tag, rows + (size_t)(WV * v0) * nblk, nblk);
                    wfail = 1;
                    break;
                }
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < WV; ++q)
#pragma unroll
                        for (int u = 0; u < 4; ++u)
                            if (!tag_current<W1>(w[h][q][u], tag)) {
                                const int j = j0 + lane + 64 * u;
                                const int v = h ? v1 : v0;
                                w[h][q][u] = ld_word(rows + (size_t)(WV * v + q) * nblk + j);
                            }
            }
            if (wfail) break;
#pragma unroll
            for (int h = 0; h < 2; ++h) {
                const bool is_sum = (h ? v1 : v0) < NV;
#pragma unroll
                for (int u = 0; u < 4; ++u) {
                    if (j0 + lane + 64 * u >= nblk) continue;
                    const double y = word_value<W1>(w[h][0][u], w[h][WV - 1][u]);
                    x[h] = is_sum ? x[h] + y : fmax(x[h], y);
                }
            }
        }
        x[0] = (v0 < NV) ? wave_red_dpp<false>(x[0]) : wave_red_dpp<true>(x[0]);
        if (lane == 63) s_tot[v0] = x[0];
        if (two) {
            x[1] = (v1 < NV) ? wave_red_dpp<false>(x[1]) : wave_red_dpp<true>(x[1]);
            if (lane == 63) s_tot[v1] = x[1];
        }
    }
    if (!idled) idle();
    if (tr && tid == 0) tr[2] = (long long)__builtin_amdgcn_s_memrealtime();
    if (wfail) s_flag[2 + (gen & 1u)] = 1;
    __syncthreads();
    const int fail = s_flag[2 + (gen & 1u)];
    gen += 1u;
    if (tr && tid == 0) tr[3] = (long long)__builtin_amdgcn_s_memrealtime();
    return !fail;
}

// Whether word 0 of rows 0 .. n-1 (n <= 256) carries the tag: lane l reads
// rows l + 64 u for u < 4, four loads in flight and none waiting for
// another; indexes past the last row read the last row again.
// Only the 32-bit half that holds the tag bits is read (the low one for the
// one-word format, the high one for the two-word format), through a global
// address-space pointer.
template <bool W1>
__device__ __forceinline__ int rows_current(const unsigned long long* rows, int lane, int n, unsigned tag) {
    using gtag = const __attribute__((address_space(1))) unsigned int;
    gtag* t = (gtag*)(reinterpret_cast<const unsigned int*>(rows) + (W1 ? 0 : 1));
    unsigned w[4];
#pragma unroll
    for (int u = 0; u < 4; ++u) w[u] = __hip_atomic_load(t + 2 * min(lane + 64 * u, n - 1), __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
    int mine = 1;
#pragma unroll
    for (int u = 0; u < 4; ++u) mine &= tag_current<W1>(W1 ? (unsigned long long)w[u] : (unsigned long long)w[u] << 32, tag);
    return mine;
}

// The streaming modes' gather: a workgroup can finish its tiles well before
// the others, and every wave re-reading its values for that long would load
// the memory system, so wave 0 polls the tag words alone first (steps 2-3
// above as a poll, then a fetch by every wave).  `idle`: work that needs none
// of the totals, run by the other waves while wave 0 polls (by wave 0 after
// its poll when the block is one wave).
template <int NV, int NMAX, bool W1, class Idle>
__device__ __forceinline__ bool xch_gather_poll(double* s_tot, const XfInstDev& I, int lb, unsigned& gen,
                                           int* status, int* s_flag, long long* tr, Idle&& idle) {
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6, nwaves = blockDim.x >> 6;
    const int B = blockDim.x;
    const int nblk = I.nblk;
    constexpr int NT = NV + NMAX;
    constexpr int RP = XF_ROW_REPLICAS;
    const unsigned tag = gen + 1u;
    const unsigned long long tagw = word_of<W1>(0., tag, 0);             // a current word (for absent rows)
    constexpr int WV = words_per_value<W1>();
    const size_t rep_words = (size_t)XF_ROWW * nblk;
    unsigned long long* rows = I.rows + (size_t)(gen & 1u) * RP * rep_words + (size_t)(lb % RP) * rep_words;
    if (tid < 64) {
        const uint64_t t0 = __builtin_amdgcn_s_memrealtime();
        int fail = 0;
        for (int spins = 1;; ++spins) {
            int mine = 1;
            for (int j0 = 0; j0 < nblk; j0 += 256) mine &= rows_current<W1>(rows + j0, lane, min(nblk - j0, 256), tag);
            if (__all(mine)) break;
            if (const int x = wait_expired(t0, spins, status)) {
                if (x == 1) note_expiry<W1>(I.rs, 1, lb, tag, rows, nblk);
                fail = 1;
                break;
            }
        }
        fail = __any(fail);
        if (tid == 0) {
            s_flag[0] = fail;
            s_flag[1] = 0;                    // set below by any wave whose fetch fails
        }
        if (nwaves == 1) idle();
    } else {
        idle();
    }
    __syncthreads();
    int fail = s_flag[0];
    // fetch and totals: wave w takes values w, w + nwaves, ... two at a
    // time; lane l reads rows l, l + 64, l + 128, l + 192 of each (8 or 16
    // loads in flight per chunk of 256 rows) and adds them in row order, which is the
    // fixed order of the sum over rows.  A stale tag (a word that had not
    // landed when word 0 had) makes the wave read the chunk again.
    const uint64_t t1 = __builtin_amdgcn_s_memrealtime();
    int wfail = fail;
    for (int v0 = wave; v0 < NT; v0 += 2 * nwaves) {
        const int v1 = v0 + nwaves;
        const bool two = v1 < NT;                             // wave-uniform
        double x[2] = {0., 0.};
        for (int j0 = 0; j0 < nblk && !wfail; j0 += 256) {
            for (int spins = 1;; ++spins) {
                unsigned long long w[2][WV][4];               // [value][word][row]
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < WV; ++q)
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            const int j = j0 + lane + 64 * u;
                            const int v = h ? v1 : v0;
                            w[h][q][u] = ((h == 0 || two) && j < nblk) ? ld_word(rows + (size_t)(WV * v + q) * nblk + j) : tagw;
                        }
                int bad = 0;
#pragma unroll
                for (int h = 0; h < 2; ++h)
#pragma unroll
                    for (int q = 0; q < WV; ++q)
#pragma unroll
                        for (int u = 0; u < 4; ++u) bad |= !tag_current<W1>(w[h][q][u], tag);
                if (!__any(bad)) {
#pragma unroll
                    for (int h = 0; h < 2; ++h) {
                        const bool is_sum = (h ? v1 : v0) < NV;
#pragma unroll
                        for (int u = 0; u < 4; ++u) {
                            if (j0 + lane + 64 * u >= nblk) continue;
                            const double y = word_value<W1>(w[h][0][u], w[h][WV - 1][u]);
                            x[h] = is_sum ? x[h] + y : fmax(x[h], y);
                        }
                    }
                    break;
                }
                if (const int x = wait_expired(t1, spins, status)) {
                    if (x == 1) note_expiry<W1>(I.rs, 2, lb, tag, rows + (size_t)(WV * v0) * nblk, nblk);
                    wfail = 1;
                    break;
                }
            }
        }
        x[0] = (v0 < NV) ? wave_red_dpp<false>(x[0]) : wave_red_dpp<true>(x[0]);
        if (lane == 63) s_tot[v0] = x[0];
        if (two) {
            x[1] = (v1 < NV) ? wave_red_dpp<false>(x[1]) : wave_red_dpp<true>(x[1]);
            if (lane == 63) s_tot[v1] = x[1];
        }
    }
    if (tr && tid == 0) tr[2] = (long long)__builtin_amdgcn_s_memrealtime();
    if (wfail) s_flag[1] = 1;
    __syncthreads();
    fail = s_flag[1];
    gen += 1u;
    if (tr && tid == 0) tr[3] = (long long)__builtin_amdgcn_s_memrealtime();
    return !fail;
}

struct NoIdle { __device__ void operator()() const {} };
// SPIN: the resident modes, whose workgroups reach each exchange together
// (xch_gather_spin); the streaming modes poll first (xch_gather_poll).
template <int NV, int NMAX, bool W1, bool SPIN, class Idle = NoIdle>
__device__ __forceinline__ bool reduce_publish(const double* acc, double* s_tot, double* s_x, int xcap, const XfInstDev& I,
                                               int lb, unsigned& gen, int* status, int* s_flag, long long* tr = nullptr,
                                               Idle&& idle = Idle{}) {
    xch_publish<NV, NMAX, W1>(acc, s_x, xcap, I, lb, gen, tr);
    if constexpr (SPIN) return xch_gather_spin<NV, NMAX, W1>(s_tot, I, lb, gen, status, s_flag, tr, idle);
    else return xch_gather_poll<NV, NMAX, W1>(s_tot, I, lb, gen, status, s_flag, tr, idle);
}
}  // namespace

// ============================================================================
// The kernel.
// ============================================================================
template <class ST, class CT, int NS, int GEOM, int CPT, int MODE, int MAXB>
__global__ void __launch_bounds__(MAXB)
k_step(const XfInstDev* __restrict__ insts, const int* __restrict__ blk_inst, int max_steps) {
    constexpr bool F64 = std::is_same<CT, double>::value;
    // FP32 increments (mixed and FP32 state): compensated error estimate
    constexpr bool COMP = !F64;
    // streaming modes: 2 sets up each tile's beam records in every pass; 4
    // keeps the records of all of the workgroup's tiles in LDS for the step
    // (geometry once per step, one set-up per pass, no barrier between the
    // tiles of passes A and B)
    constexpr bool STREAM = (MODE == 2 || MODE == 4);
    constexpr bool TABLE = (MODE == 4);
    // multi-angle and single-angle: the first two sub-steps' path lengths
    // differ by exactly a factor of 2 (U over the longer one by double angle)
    constexpr bool DBL = (GEOM != XF_MAA);
    using R = BRec<GEOM>;
    constexpr int NM = R::NM;
    constexpr int NMV = NM * 3;
    static_assert(2 * NMV + 1 <= XF_NVMAX, "partials");
    static_assert(NS == 2 || NS == 4, "the table's padding slot holds the FP32 H_vac");

    const XfInstDev& I = insts[blk_inst[blockIdx.x]];
    const int B = blockDim.x;
    if ((B & 63) != 0 || B > MAXB) {
        // whole waves only; every workgroup leaves before any barrier
        if (threadIdx.x == 0) __hip_atomic_store(&I.rs->status, (int)XF_EINVAL, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
        return;
    }
    const int tid = threadIdx.x;
    const int lane = tid & 63, wave = tid >> 6, nwaves = B >> 6;
    const int lb = blockIdx.x - I.blk0;
    const int nblk = I.nblk;
    const int ne = I.ne, bpt = I.bpt, ntiles = I.ntiles;
    const long long ncells = I.ncells;
    XfRunScalars* rs = I.rs;
    // an error left by an earlier launch of the same graph: nothing to do
    // (read before any workgroup of this launch can change it)
    if (__hip_atomic_load(&rs->status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0) {
        if (lb == 0 && tid == 0 && I.rs_host) {
            mirror_scalars(rs, I.rs_host);
            I.rs_host->steps_call = 0;
            I.rs_host->acc_call = 0;
        }
        return;
    }

    extern __shared__ double smem[];
    double* s_beam = smem;                                   // [bpt][R::SIZE]
    // mode 4: one record per beam of the workgroup's tiles t = lb + j nblk,
    // tch tiles at a time (all of them when they fit: nch == 1)
    const int ntl = (TABLE && lb < ntiles) ? (ntiles - lb + nblk - 1) / nblk : 1;
    const int tch = TABLE ? max(1, min(I.tchunk, ntl)) : 1;
    const bool chunked = TABLE && tch < ntl;                  // workgroup-uniform
    double* s_errb = s_beam + tch * bpt * R::SIZE;           // [4][B*CPT], between passes the
                                                             // publish staging of the lanes' values
    const int xcap = 4 * B * CPT;                           // doubles of s_errb
    double* s_errs = s_errb + xcap;                          // [bpt][4]
    constexpr int TS = tab_slots(NS);
    double* s_tab = s_errs + 4 * bpt;                        // [ne][TS] per-energy table
    // mode 0 with FP32 increments: the remainder coefficient set Q of every
    // cell and species, kept in LDS from pass A to pass C (16-byte aligned)
    float4* s_keep = reinterpret_cast<float4*>(smem + (((int)(s_tab - smem) + ne * TS + 1) & ~1));   // [CPT][NS][B]
    for (int e = tid; e < ne; e += B) {
        double* t = s_tab + e * TS;
        t[0] = I.hv[e];
        t[1] = I.hv[ne + e];
        // FP32 increments: the error weights carry the 1/4 of |psi4 - psi1|^2 = |E|^2 / 4
#pragma unroll
        for (int q = 0; q < 4; ++q) t[2 + q] = F64 ? I.errw[q * ne + e] : .25 * I.errw[q * ne + e];
        if constexpr (!F64) {
            float* tf = reinterpret_cast<float*>(t + 6 + NS);
            tf[0] = (float)I.hv[e];
            tf[1] = (float)I.hv[ne + e];
        }
#pragma unroll
        for (int q = 0; q < NS; ++q) put_src_weight<F64>(t + 6 + q, I.srcw[q * ne + e], I.srcwf[q * ne + e]);
    }
    __syncthreads();
    __shared__ int s_flag[4];
    if (tid == 0) {
        s_flag[2] = 0;                                        // the exchanges' failure flags
        s_flag[3] = 0;
    }
    __shared__ double s_hm[2];                                // matter term at r + dr/2, r + dr
    // angular moments, uniform over the workgroup, gathered in place by the
    // exchanges: psi1 and psi3 (pass A), psi4 (pass B), and two buffers that
    // take turns as psi0 and the candidate (pass C, with the step error
    // after the moments)
    __shared__ double s_MA[2 * NMV], s_MB[NMV], s_MC[2][NMV + 1];

    // uniform step state
    double r = uni_d(rs->r), dr = uni_d(rs->dr);
    long long total = uni_ll(rs->total), accepted = uni_ll(rs->accepted);
    int cur = __builtin_amdgcn_readfirstlane(rs->cur);
    double last_err = uni_d(rs->last_err);
    double last_dr = uni_d(rs->last_dr);
    unsigned gen = (unsigned)__builtin_amdgcn_readfirstlane((int)rs->gen0);   // generation of the exchanges
    int* status = &rs->status;

    Cells<ST, CT, NS, CPT> C;
    double* M0 = s_MC[0];
    const double* M1 = s_MA;
    const double* M3 = s_MA + NMV;
    const double* M4 = s_MB;
    double* Mn = s_MC[1];

    // ---- tile helpers -------------------------------------------------------
    auto tile_cells = [&](int t) __attribute__((always_inline)) {
        int beam0 = t * bpt;
        int nbt = max(0, min(bpt, I.nbeam - beam0));
        int ncl = nbt * ne;
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            int i = tid + k * B;
            if (t < ntiles && i < ncl) {
                int b = i / ne;
                C.bl[k] = b;
                C.e[k] = i - b * ne;
            } else {
                C.bl[k] = -1;
                C.e[k] = 0;
            }
        }
        return nbt;
    };
    auto cell_index = [&](int t, int k) __attribute__((always_inline)) -> long long {
        return (long long)t * bpt * ne + (long long)C.bl[k] * ne + C.e[k];
    };
    auto load_p0 = [&](int t, int buf) __attribute__((always_inline)) {
        const ST* base = (const ST*)I.psi[buf];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (C.bl[k] >= 0) {
                long long c = cell_index(t, k);
#pragma unroll
                for (int s = 0; s < NS; ++s) C.p0[k][s] = ld_sp<ST>(base, ncells, s, c);
            }
        }
    };
    // geometry of the tile's beams at the step's three points, one (beam,
    // point) per thread; the path lengths follow after a barrier (tile_dl, on
    // the threads that run tile_h), Hb separately
    auto tile_points = [&](int t, int nbt, double rr, double rdr2, double rdr) __attribute__((always_inline)) {
        for (int q = tid; q < 3 * nbt; q += B) {
            const int b = q / 3, pt = q - 3 * b;
            beam_point<GEOM>(I, t * bpt + b, pt, pt == 0 ? rr : (pt == 1 ? rdr2 : rdr), s_beam + b * R::SIZE);
        }
    };
    auto tile_dl = [&](int nbt, double ddr, double dr2) __attribute__((always_inline)) {
        for (int b = tid; b < nbt; b += B) beam_dl<GEOM>(s_beam + b * R::SIZE, ddr, dr2);
    };
    // pass A's points, path lengths and h0 in one phase (multi-angle and
    // single-angle geometries)
    auto tile_setup_a = [&](int t, int nbt, double rr, double rdr2, double rdr, double ddr, double dr2,
                            const double* M, double hm) __attribute__((always_inline)) {
        if constexpr (DBL) {
            const double coef = F64 ? I.coef : I.coef_f;
            for (int q = tid; q < 3 * nbt; q += B) {
                const int b = q / 3, pt = q - 3 * b;
                beam_point_a<GEOM>(I, t * bpt + b, pt, pt == 0 ? rr : (pt == 1 ? rdr2 : rdr), s_beam + b * R::SIZE,
                                   ddr, dr2, M, coef, hm);
            }
        }
    };
    // mode 4: the same over every beam of the workgroup's tiles (slot q: tile
    // lb + (q / bpt) nblk, beam q % bpt of it)
    // (slot q of the chunk that starts at local tile c0: tile
    // lb + (c0 + q / bpt) nblk, beam q % bpt of it; the same thread forms a
    // slot's path lengths, Hamiltonians and Hd, so they need no barrier
    // between them)
    auto table_beam = [&](int c0, int q) __attribute__((always_inline)) -> int {
        const int j = q / bpt;
        return (lb + (c0 + j) * nblk) * bpt + (q - j * bpt);
    };
    auto table_points = [&](int c0, double rr, double rdr2, double rdr) __attribute__((always_inline)) {
        const int nq = min(tch, ntl - c0) * bpt;
        for (int q = tid; q < 3 * nq; q += B) {
            const int bq = q / 3, pt = q - 3 * bq;
            const int beam = table_beam(c0, bq);
            if (beam < I.nbeam)
                beam_point<GEOM>(I, beam, pt, pt == 0 ? rr : (pt == 1 ? rdr2 : rdr), s_beam + bq * R::SIZE);
        }
    };
    auto table_dl = [&](int c0, double ddr, double dr2) __attribute__((always_inline)) {
        const int nq = min(tch, ntl - c0) * bpt;
        for (int q = tid; q < nq; q += B)
            if (table_beam(c0, q) < I.nbeam) beam_dl<GEOM>(s_beam + q * R::SIZE, ddr, dr2);
    };
    auto table_h = [&](int c0, int slot, int pt, const double* M, double hm) __attribute__((always_inline)) {
        const double coef = F64 ? I.coef : I.coef_f;
        const int nq = min(tch, ntl - c0) * bpt;
        for (int q = tid; q < nq; q += B)
            if (table_beam(c0, q) < I.nbeam)
                beam_h<GEOM>(s_beam + q * R::SIZE, pt, slot, M, coef, hm);
    };
    auto table_hd = [&](int c0) __attribute__((always_inline)) {
        const int nq = min(tch, ntl - c0) * bpt;
        for (int q = tid; q < nq; q += B)
            if (table_beam(c0, q) < I.nbeam) beam_hd<GEOM>(s_beam + q * R::SIZE);
    };
    auto tile_h = [&](int nbt, int slot, int pt, const double* M, double hm) __attribute__((always_inline)) {
        const double coef = F64 ? I.coef : I.coef_f;   // float source weights are pre-scaled
        for (int b = tid; b < nbt; b += B) beam_h<GEOM>(s_beam + b * R::SIZE, pt, slot, M, coef, hm);
    };

    const int my_tile0 = lb;                 // resident: exactly this tile
    int nbt_res = 0;
    if (!STREAM) {
        nbt_res = tile_cells(my_tile0);
        load_p0(my_tile0, cur);
    }

    // ---- moments of psi0 at r (the cached energy sums of Nmr.cpp:117-121):
    // left by the previous launch of this shape, or by the moments kernels
    // (xf_moments.hip) the host runs before a launch that has none.  Every
    // workgroup reads the same run scalars, so all of them leave together
    // here when the host did not.
    if (__builtin_amdgcn_readfirstlane(rs->m0_ok) == 0 || uni_ll(rs->m0_key) != I.launch_key || uni_d(rs->m0_r) != r) {
        if (lb == 0 && tid == 0) {
            __hip_atomic_store(status, (int)XF_ESTATE, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            rs->steps_call = 0;
            rs->acc_call = 0;
            if (I.rs_host) {
                mirror_scalars(rs, I.rs_host);
                I.rs_host->status = (int)XF_ESTATE;
            }
        }
        return;
    }
    if (tid < NMV) M0[tid] = rs->m0[tid];
    __syncthreads();
    bool ok = true;

    long long steps_call = 0, acc_call = 0;
    int stop_reason = 0;
    double hm_cur = uni_d(hm_at(I, r));
    // Every exchange ends with a workgroup barrier after its last LDS access,
    // and every reader of the beam records, the staging and the moments comes
    // before one: a step starts without a barrier of its own.
    while (ok) {
        if (!(steps_call < max_steps)) { stop_reason = 0; break; }
        if (!(r < I.Rn)) { stop_reason = 1; break; }
        const double dr2 = .5 * dr;
        const double rdr = r + dr;
        const double rdr2 = r + dr2;
        long long* tr = (I.trace && lb == 0 && steps_call < I.trace_steps) ? I.trace + steps_call * 16 : nullptr;
        if (tr && tid == 0) tr[0] = (long long)__builtin_amdgcn_s_memrealtime();
        // matter term: at r carried from the previous step (after an accepted
        // step r is the same double as its r + dr); at r + dr/2 and r + dr
        // (needed from pass B on) evaluated on lanes 0 and 1 of wave 1 while
        // its loads of pass A's exchange are in flight
        const double hm_r = hm_cur;
        double hm_rdr2 = 0., hm_rdr = 0.;
        auto hm_next = [&]() __attribute__((always_inline)) {
            if (wave == min(1, nwaves - 1) && lane < 2) s_hm[lane] = hm_at(I, lane ? rdr : rdr2);
        };

        // =================================================== pass A (S1, S2, S6)
        {
            double acc[XF_NVMAX];
#pragma unroll
            for (int v = 0; v < 2 * NMV; ++v) acc[v] = 0.;
            for (int t = (STREAM ? lb : my_tile0); t < (STREAM ? ntiles : my_tile0 + 1); t += (STREAM ? nblk : 1)) {
                int nbt = STREAM ? tile_cells(t) : nbt_res;
                const int jt = TABLE ? (t - lb) / nblk : 0;           // local tile index
                const double* tb = TABLE ? s_beam + (jt % tch) * bpt * R::SIZE : s_beam;
                if constexpr (TABLE) {
                    if (jt % tch == 0) {                  // a chunk starts: its records
                        if (jt > 0) __syncthreads();
                        table_points(jt, r, rdr2, rdr);
                        __syncthreads();
                        table_dl(jt, dr, dr2);
                        table_h(jt, 0, 0, M0, hm_r);
                        __syncthreads();
                    }
                    load_p0(t, cur);
                } else {
                    // the previous tile is done with s_beam (resident modes: the
                    // last exchange ended with a barrier after every read of it)
                    if (MODE == 2) __syncthreads();
                    if (MODE == 2) load_p0(t, cur);      // in flight during the set-up
                    if constexpr (DBL) {
                        tile_setup_a(t, nbt, r, rdr2, rdr, dr, dr2, M0, hm_r);
                    } else {
                        tile_points(t, nbt, r, rdr2, rdr);
                        __syncthreads();
                        tile_dl(nbt, dr, dr2);
                        tile_h(nbt, 0, 0, M0, hm_r);
                    }
                    __syncthreads();
                }
#pragma unroll
                for (int k = 0; k < CPT; ++k) {
                    if (C.bl[k] < 0) continue;
                    const double* rec = tb + C.bl[k] * R::SIZE;
                    const int e = C.e[k];
                    const double* te = s_tab + e * TS;
                    const float* tef = reinterpret_cast<const float*>(te + 6 + NS);
                    const float* recf = reinterpret_cast<const float*>(rec + R::HF);
                    CT X0 = CT(0), X2 = CT(0), X3 = CT(0), Y0 = CT(0), Y2 = CT(0), Y3 = CT(0);
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        double h0, h1, h2;
                        cell_h(te, s, rec + R::H + 0, h0, h1, h2);
                        if constexpr (F64) {
                            double lam = lam_d(h0, h1, h2);
                            PropD P1, P0;
                            if constexpr (DBL) {
                                prop_d2(h0, h1, h2, lam, rec[R::DL + 0], P1, P0);
                            } else {
                                prop_d(h0, h1, h2, lam, rec[R::DL + 1], P1);
                                prop_d(h0, h1, h2, lam, rec[R::DL + 0], P0);
                            }
                            Sp<double> a = sp_d(C.p0[k][s]);
                            Sp<double> v1 = apply_d(P1, a), v3 = apply_d(P0, a);
                            if (MODE == 0) { C.k1[k][s] = v1; C.k3[k][s] = v3; }
                            src_add<CT>(te, s, v1, X0, X2, X3);
                            src_add<CT>(te, s, v3, Y0, Y2, Y3);
                        } else if constexpr (DBL) {
                            // dl1 = 2 dl0: U(dl1) = U(dl0)^2, so only the states' densities
                            // need psi3 = U0 psi0 and psi1 = U0 psi3
                            float f0, f1, f2;
                            cell_hf(tef, s, recf + 0, f0, f1, f2);
                            float cm1, sc, scr;
                            prop_f2s(f0, f1, f2, recf[12], cm1, sc, scr);
                            const PropF P0 = pf_of(f0, f1, f2, cm1, sc);
                            const Sp<float> a = sp_f(C.p0[k][s]);
                            const Sp<float> p3 = delta_acc(P0, a, a);
                            src_add<CT>(te, s, delta_acc(P0, p3, p3), X0, X2, X3);
                            src_add<CT>(te, s, p3, Y0, Y2, Y3);
                            if (MODE == 0) {
                                // kept for pass C, formed as pass C forms them in the
                                // other modes: Q = R0 - R1 so far, and delta3
                                const PropF q = pf_rdiff2(f0, f1, f2, cm1, sc, scr);
                                s_keep[(k * NS + s) * B + tid] = make_float4(q.cm1, q.m1, q.m2, q.m3);
                                C.k3[k][s] = delta_acc(P0, a, Sp<float>{-0.f, -0.f, -0.f, -0.f});
                            }
                        } else {
                            float f0, f1, f2, c1, s1, r1, c0, s0, r0;
                            cell_hf(tef, s, recf + 0, f0, f1, f2);
                            prop_f2s(f0, f1, f2, recf[13], c1, s1, r1);
                            prop_f2s(f0, f1, f2, recf[12], c0, s0, r0);
                            const PropF P0 = pf_of(f0, f1, f2, c0, s0);
                            const Sp<float> a = sp_f(C.p0[k][s]);
                            src_add<CT>(te, s, delta_acc(pf_of(f0, f1, f2, c1, s1), a, a), X0, X2, X3);
                            src_add<CT>(te, s, delta_acc(P0, a, a), Y0, Y2, Y3);
                            if (MODE == 0) {
                                const PropF q = pf_fma<1>(pf_fma<-1>(PropF{-0.f, -0.f, -0.f, -0.f}, f0, f1, f2, c1, r1),
                                                          f0, f1, f2, c0, r0);
                                s_keep[(k * NS + s) * B + tid] = make_float4(q.cm1, q.m1, q.m2, q.m3);
                                C.k3[k][s] = delta_acc(P0, a, Sp<float>{-0.f, -0.f, -0.f, -0.f});
                            }
                        }
                    }
                    acc_moments<NM>(acc, rec + R::W + 2 * NM, (double)X0, (double)X2, (double)X3);        // h1 at r+dr
                    acc_moments<NM>(acc + NMV, rec + R::W + 1 * NM, (double)Y0, (double)Y2, (double)Y3);  // h3 at r+dr/2
                }
            }
            ok = reduce_publish<2 * NMV, 0, !F64, !STREAM>(acc, s_MA, s_errb, xcap, I, lb, gen, status, s_flag, tr ? tr + 1 : nullptr,
                                                  hm_next);
            if (!ok) break;
            hm_rdr2 = uni_d(s_hm[0]);
            hm_rdr = uni_d(s_hm[1]);
        }

        // =================================================== pass B (S3, S4/S5, S7, S8)
        {
            double acc[XF_NVMAX];
#pragma unroll
            for (int v = 0; v < NMV; ++v) acc[v] = 0.;
            if (TABLE && !chunked) {                  // the records of pass A stay
                table_h(0, 1, 2, M1, hm_rdr);
                table_h(0, 2, 1, M3, hm_rdr2);
                __syncthreads();
            }
            for (int t = (STREAM ? lb : my_tile0); t < (STREAM ? ntiles : my_tile0 + 1); t += (STREAM ? nblk : 1)) {
                int nbt = STREAM ? tile_cells(t) : nbt_res;
                const int jt = TABLE ? (t - lb) / nblk : 0;           // local tile index
                const double* tb = TABLE ? s_beam + (jt % tch) * bpt * R::SIZE : s_beam;
                if constexpr (TABLE) {
                    if (chunked && jt % tch == 0) {       // a chunk starts: its records again
                        if (jt > 0) __syncthreads();
                        table_points(jt, r, rdr2, rdr);
                        __syncthreads();
                        table_dl(jt, dr, dr2);
                        table_h(jt, 1, 2, M1, hm_rdr);
                        table_h(jt, 2, 1, M3, hm_rdr2);
                        __syncthreads();
                    }
                    load_p0(t, cur);
                } else {
                    if (MODE == 2) {
                        __syncthreads();
                        load_p0(t, cur);                  // in flight during the set-up
                        tile_points(t, nbt, r, rdr2, rdr);
                        __syncthreads();
                        tile_dl(nbt, dr, dr2);
                    }
                    tile_h(nbt, 1, 2, M1, hm_rdr);
                    tile_h(nbt, 2, 1, M3, hm_rdr2);
                    __syncthreads();
                }
#pragma unroll
                for (int k = 0; k < CPT; ++k) {
                    if (C.bl[k] < 0) continue;
                    const double* rec = tb + C.bl[k] * R::SIZE;
                    const int e = C.e[k];
                    const double* te = s_tab + e * TS;
                    const float* tef = reinterpret_cast<const float*>(te + 6 + NS);
                    const float* recf = reinterpret_cast<const float*>(rec + R::HF);
                    CT X0 = CT(0), X2 = CT(0), X3 = CT(0);
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        double h0, h1, h2, g0, g1, g2;
                        cell_h(te, s, rec + R::H + 3, h0, h1, h2);      // h1 (+ Hm(r+dr))
                        cell_h(te, s, rec + R::H + 6, g0, g1, g2);      // h3 (+ Hm(r+dr/2))
                        if constexpr (F64) {
                            Sp<double> a = sp_d(C.p0[k][s]);
                            PropD Pb, P4;
                            prop_d(h0, h1, h2, lam_d(h0, h1, h2), rec[R::DL + 1], Pb);
                            prop_d(g0, g1, g2, lam_d(g0, g1, g2), rec[R::DL + 1], P4);
                            Sp<double> b = apply_d(Pb, a);
                            if (MODE == 0) C.k1[k][s] = sp_avg(b, C.k1[k][s]);
                            Sp<double> v4 = apply_d(P4, a);
                            if (MODE == 0) C.k4[k][s] = v4;
                            src_add<CT>(te, s, v4, X0, X2, X3);
                        } else {
                            Sp<float> a = sp_f(C.p0[k][s]);
                            float f0, f1, f2, cm1, sc, scr;
                            cell_hf(tef, s, recf + 6, f0, f1, f2);
                            prop_f2s(f0, f1, f2, recf[13], cm1, sc, scr);
                            const PropF P4 = pf_of(f0, f1, f2, cm1, sc);
                            src_add<CT>(te, s, delta_acc(P4, a, a), X0, X2, X3);    // psi4 = U4 psi0
                            if (MODE == 0) {
                                // Q -= Rb, Q += R3 and delta4, as pass C forms them in the
                                // other modes
                                float g0, g1, g2, gc, gs, gr;
                                cell_hf(tef, s, recf + 3, g0, g1, g2);
                                prop_f2s(g0, g1, g2, recf[13], gc, gs, gr);
                                const float4 o = s_keep[(k * NS + s) * B + tid];
                                const PropF q = pf_fma<1>(pf_fma<-1>(PropF{o.x, o.y, o.z, o.w}, g0, g1, g2, gc, gr),
                                                          f0, f1, f2, cm1, scr);
                                s_keep[(k * NS + s) * B + tid] = make_float4(q.cm1, q.m1, q.m2, q.m3);
                                C.k4[k][s] = delta_acc(P4, a, Sp<float>{-0.f, -0.f, -0.f, -0.f});
                            }
                        }
                    }
                    acc_moments<NM>(acc, rec + R::W + 2 * NM, (double)X0, (double)X2, (double)X3);   // h4 at r+dr
                }
            }
            ok = reduce_publish<NMV, 0, !F64, !STREAM>(acc, s_MB, s_errb, xcap, I, lb, gen, status, s_flag, tr ? tr + 5 : nullptr);
            if (!ok) break;
        }

        // =================================================== pass C (S9, S10/S11, error)
        double errmax;
        {
            double acc[XF_NVMAX];
#pragma unroll
            for (int v = 0; v < NMV + 1; ++v) acc[v] = 0.;
            double blk_err = 0.;
            if (TABLE && !chunked) {                  // the records of passes A and B stay
                table_h(0, 3, 2, M4, hm_rdr);
                if constexpr (COMP) table_hd(0);
                __syncthreads();
            }
            for (int t = (STREAM ? lb : my_tile0); t < (STREAM ? ntiles : my_tile0 + 1); t += (STREAM ? nblk : 1)) {
                int nbt = STREAM ? tile_cells(t) : nbt_res;
                const int jt = TABLE ? (t - lb) / nblk : 0;           // local tile index
                const double* tb = TABLE ? s_beam + (jt % tch) * bpt * R::SIZE : s_beam;
                if constexpr (TABLE) {
                    if (chunked && jt % tch == 0) {       // a chunk starts: its records again
                        if (jt > 0) __syncthreads();
                        table_points(jt, r, rdr2, rdr);
                        __syncthreads();
                        table_dl(jt, dr, dr2);
                        table_h(jt, 0, 0, M0, hm_r);
                        table_h(jt, 1, 2, M1, hm_rdr);
                        table_h(jt, 2, 1, M3, hm_rdr2);
                        table_h(jt, 3, 2, M4, hm_rdr);
                        if constexpr (COMP) table_hd(jt);
                        __syncthreads();
                    }
                    load_p0(t, cur);
                } else {
                    if (MODE == 2) {
                        __syncthreads();
                        load_p0(t, cur);                  // in flight during the set-up
                        tile_points(t, nbt, r, rdr2, rdr);
                        __syncthreads();
                        tile_dl(nbt, dr, dr2);
                        tile_h(nbt, 0, 0, M0, hm_r);
                        tile_h(nbt, 1, 2, M1, hm_rdr);
                        tile_h(nbt, 2, 1, M3, hm_rdr2);
                        __syncthreads();
                    }
                    tile_h(nbt, 3, 2, M4, hm_rdr);
                    if constexpr (COMP)
                        for (int b = tid; b < nbt; b += B) beam_hd<GEOM>(s_beam + b * R::SIZE);
                    __syncthreads();
                }
#pragma unroll
                for (int k = 0; k < CPT; ++k) {
                    const int slot = k * B + tid;
                    if (C.bl[k] < 0) {
#pragma unroll
                        for (int q = 0; q < 4; ++q) s_errb[q * B * CPT + slot] = 0.;
                        continue;
                    }
                    const double* rec = tb + C.bl[k] * R::SIZE;
                    const int e = C.e[k];
                    const double* te = s_tab + e * TS;
                    const float* tef = reinterpret_cast<const float*>(te + 6 + NS);
                    const float* recf = reinterpret_cast<const float*>(rec + R::HF);
                    CT X0 = CT(0), X2 = CT(0), X3 = CT(0);
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        double dsq;
                        double h0, h1, h2;
                        cell_h(te, s, rec + R::H + 9, h0, h1, h2);     // h4 (+ Hm(r+dr))
                        if constexpr (F64) {
                            Sp<double> a = sp_d(C.p0[k][s]);
                            Sp<double> q1, q3, q4;
                            if (MODE == 0) {
                                q1 = C.k1[k][s]; q3 = C.k3[k][s]; q4 = C.k4[k][s];
                            } else {
                                double x0, x1, x2;
                                cell_h(te, s, rec + R::H + 0, x0, x1, x2);
                                double lam = lam_d(x0, x1, x2);
                                PropD P1, P0, Pb, P4;
                                if constexpr (DBL) {
                                    prop_d2(x0, x1, x2, lam, rec[R::DL + 0], P1, P0);
                                } else {
                                    prop_d(x0, x1, x2, lam, rec[R::DL + 1], P1);
                                    prop_d(x0, x1, x2, lam, rec[R::DL + 0], P0);
                                }
                                q1 = apply_d(P1, a);
                                q3 = apply_d(P0, a);
                                cell_h(te, s, rec + R::H + 3, x0, x1, x2);
                                prop_d(x0, x1, x2, lam_d(x0, x1, x2), rec[R::DL + 1], Pb);
                                q1 = sp_avg(apply_d(Pb, a), q1);
                                cell_h(te, s, rec + R::H + 6, x0, x1, x2);
                                prop_d(x0, x1, x2, lam_d(x0, x1, x2), rec[R::DL + 1], P4);
                                q4 = apply_d(P4, a);
                            }
                            PropD P2;
                            prop_d(h0, h1, h2, lam_d(h0, h1, h2), rec[R::DL + 2], P2);
                            Sp<double> q3n = apply_d(P2, q3);
                            Sp<double> fin = sp_avg(q3n, q4);
                            dsq = dnorm2<double>(fin, q1);
                            src_add<CT>(te, s, fin, X0, X2, X3);
                            if (!STREAM) C.k4[k][s] = fin;
                            if (STREAM) st_sp<ST>((ST*)I.psi[cur ^ 1], ncells, s, cell_index(t, k), fin);
                        } else {
                            Sp<float> a = sp_f(C.p0[k][s]);
                            Sp<float> d3, d4, dn;
                            static_assert(COMP, "FP32 increments always use the compensated error estimate");
                            {
                                // increments of the state as below; the error from the
                                // remainders R = (U - I) + i dl H (second order, so
                                // rounded relative to x^2) plus the first-order part
                                // -i Hd psi0 with Hd combined per beam in FP64:
                                //   psi4 - psi1 = E / 2,
                                //   E = R0 - R1 - Rb + R3 + R4 (psi0 + d3) + L4 d3 - i Hd psi0
                                //     = (R0 - R1 - Rb + R3 + R4 - i Hd) psi0 + P4 d3
                                // with R* and P4 = R4 + L4 the coefficient sets of
                                // (U - I): one application for all remainders, and
                                // P4 d3 shared with the state's d3n.  Mode 0 kept
                                // Q = R0 - R1 - Rb + R3, d3 and d4 from passes A and
                                // B; modes 1/2 recompute them here in the same order.
                                // Remainder sets are accumulated into Q by one FMA per
                                // component (pf_fma), increments by FMA chains
                                // (delta_acc).
                                const Sp<float> Z{-0.f, -0.f, -0.f, -0.f};
                                PropF Q;
                                // h3's propagator (modes 1/2): its increment d4 = D3 psi0
                                // is only needed in d3n + d4 = d3 + u + (D3 + D4) psi0,
                                // so its coefficient set is added to h4's and applied once
                                float k0 = 0.f, k1 = 0.f, k2 = 0.f, kc = 0.f, ks = 0.f;
                                if (MODE == 0) {
                                    const float4 o = s_keep[(k * NS + s) * B + tid];
                                    Q = PropF{o.x, o.y, o.z, o.w};
                                    d3 = C.k3[k][s];
                                    d4 = C.k4[k][s];
                                } else {
                                float f0, f1, f2, cm1, sc, scr;
                                cell_hf(tef, s, recf + 0, f0, f1, f2);                         // h0
                                if constexpr (DBL) {
                                    prop_f2s(f0, f1, f2, recf[12], cm1, sc, scr);
                                    d3 = delta_acc(pf_of(f0, f1, f2, cm1, sc), a, Z);
                                    Q = pf_rdiff2(f0, f1, f2, cm1, sc, scr);
                                } else {
                                    prop_f2s(f0, f1, f2, recf[13], cm1, sc, scr);
                                    Q = pf_fma<-1>(PropF{-0.f, -0.f, -0.f, -0.f}, f0, f1, f2, cm1, scr);
                                    prop_f2s(f0, f1, f2, recf[12], cm1, sc, scr);
                                    d3 = delta_acc(pf_of(f0, f1, f2, cm1, sc), a, Z);
                                    Q = pf_fma<1>(Q, f0, f1, f2, cm1, scr);
                                }
                                cell_hf(tef, s, recf + 3, f0, f1, f2);                         // h1
                                prop_f2s(f0, f1, f2, recf[13], cm1, sc, scr);
                                Q = pf_fma<-1>(Q, f0, f1, f2, cm1, scr);
                                cell_hf(tef, s, recf + 6, f0, f1, f2);                         // h3
                                prop_f2s(f0, f1, f2, recf[13], cm1, sc, scr);
                                k0 = f0; k1 = f1; k2 = f2; kc = cm1; ks = sc;
                                Q = pf_fma<1>(Q, f0, f1, f2, cm1, scr);
                                }
                                float g0, g1, g2, gc, gs, gr;
                                cell_hf(tef, s, recf + 9, g0, g1, g2);
                                prop_f2s(g0, g1, g2, recf[14], gc, gs, gr);                    // h4
                                const PropF P = pf_of(g0, g1, g2, gc, gs);
                                const Sp<float> u = delta_acc(P, d3, Z);
                                // dn = d3n + d4 with d3n = d3 + D4 (psi0 + d3) = d3 + u + D4 psi0
                                if (MODE == 0) dn = delta_acc(P, a, sp_add(sp_add(d3, u), d4));
                                else dn = delta_acc(pf_fma<1>(P, k0, k1, k2, kc, ks), a, sp_add(d3, u));
                                // -i Hd psi0 with Hd's energy-independent part combined
                                // per beam in FP64 (the cancellation of the first-order
                                // terms happens there) and applied in FP32
                                const float* hdf = recf + 16;
                                const float sg = (s & 1) ? -1.f : 1.f;
                                const PropF PD{-0.f, -fmaf(hdf[3], tef[0], sg * hdf[0]), hdf[2],
                                               -fmaf(hdf[3], tef[1], sg * hdf[1])};
                                Q = pf_add(pf_fma<1>(Q, g0, g1, g2, gc, gr), PD);
                                const Sp<float> E = delta_acc(Q, a, u);
                                dsq = (double)fmaf(E.ar, E.ar, fmaf(E.ai, E.ai, fmaf(E.br, E.br, E.bi * E.bi)));   // |E|^2
                            }
                            if constexpr (std::is_same<ST, float>::value) {
                                const Sp<float> vf = candidate_t(a, dn, I.renorm);
                                src_add<CT>(te, s, vf, X0, X2, X3);
                                if (!STREAM) C.k4[k][s] = vf;        // the candidate state itself
                                if (STREAM) st_sp<ST>((ST*)I.psi[cur ^ 1], ncells, s, cell_index(t, k), sp_d(vf));
                            } else {
                                const Sp<float> df{.5f * dn.ar, .5f * dn.ai, .5f * dn.br, .5f * dn.bi};
                                Sp<double> nd = candidate(sp_d(C.p0[k][s]), df, I.renorm);
                                src_add<CT>(te, s, sp_f(nd), X0, X2, X3);
                                if (!STREAM) C.k4[k][s] = df;        // increment; re-formed on accept
                                if (STREAM) st_sp<ST>((ST*)I.psi[cur ^ 1], ncells, s, cell_index(t, k), nd);
                            }
                        }
                        // error weights of the 4 particles (NBeam.cpp:922): with 2 stored
                        // species nu_x / anti-nu_x share |dpsi|^2 with nu_e / anti-nu_e
                        if (NS == 2) {
                            s_errb[s * B * CPT + slot] = te[2 + s] * dsq;
                            s_errb[(s + 2) * B * CPT + slot] = te[4 + s] * dsq;
                        } else {
                            s_errb[s * B * CPT + slot] = te[2 + s] * dsq;
                        }
                    }
                    acc_moments<NM>(acc, rec + R::W + 2 * NM, (double)X0, (double)X2, (double)X3);   // new psi0 at r+dr
                }
                // per-beam energy sums of the error (one wave per beam, fixed order)
                if (tr && tid == 0) tr[14] = (long long)__builtin_amdgcn_s_memrealtime();
                __syncthreads();
                for (int b = wave; b < nbt; b += nwaves) {
                    // lane l sums cells l, l + 64, ... of beam b, in that order, for
                    // the 4 particles together (16 loads in flight, 4 chains)
                    double sq[4] = {0., 0., 0., 0.};
                    const int cend = (b + 1) * ne;
                    for (int c0 = b * ne + lane; c0 < cend; c0 += 4 * 64) {
                        double v[4][4];
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int q = 0; q < 4; ++q) v[u][q] = (c0 + u * 64 < cend) ? s_errb[q * B * CPT + c0 + u * 64] : 0.;
#pragma unroll
                        for (int u = 0; u < 4; ++u)
#pragma unroll
                            for (int q = 0; q < 4; ++q) sq[q] += v[u][q];
                    }
#pragma unroll
                    for (int q = 0; q < 4; ++q) {
                        const double w = wave_red_dpp<false>(sq[q]);
                        if (lane == 63) s_errs[b * 4 + q] = w;
                    }
                }
                __syncthreads();
                for (int b = tid; b < nbt; b += B) {
                    double m = fmax(fmax(s_errs[b * 4 + 0], s_errs[b * 4 + 1]), fmax(s_errs[b * 4 + 2], s_errs[b * 4 + 3]));
                    blk_err = fmax(blk_err, m);
                }
                if (tr && tid == 0) tr[15] = (long long)__builtin_amdgcn_s_memrealtime();
            }
            acc[NMV] = blk_err;
            ok = reduce_publish<NMV, 1, !F64, !STREAM>(acc, Mn, s_errb, xcap, I, lb, gen, status, s_flag, tr ? tr + 9 : nullptr);
            errmax = uni_d(Mn[NMV]);
            if (!ok) break;
        }

        // =================================================== decision (Nmr.cpp:195-249)
        // the stop rule's parameters, read together (one batch of scalar
        // loads instead of a chain of dependent ones)
        const double eps0 = I.eps0, ksi = I.ksi, max_dr = I.max_dr, Rn = I.Rn;
        const double dump_rstep = I.dump_rstep, dump_last_r = I.dump_last_r;
        const int dump_itrstep = I.dump_itrstep;
        const double err = sqrt(errmax);
        const bool acc_ok = err < eps0;
        if (lb == 0 && tid == 0 && I.log) {
            xf_step_rec rec;
            rec.r = r; rec.dr = dr; rec.err = err; rec.accepted = acc_ok ? 1 : 0; rec.pad = 0;
            I.log[steps_call] = rec;
        }
        if (tr && tid == 0) tr[13] = (long long)__builtin_amdgcn_s_memrealtime();
        const long long it = total;
        if (acc_ok) {
            r += dr;
            hm_cur = hm_rdr;
            if (STREAM) {
                cur ^= 1;
            } else {
#pragma unroll
                for (int k = 0; k < CPT; ++k)
#pragma unroll
                    for (int s = 0; s < NS; ++s) {
                        if constexpr (F64 || std::is_same<ST, float>::value) {
                            C.p0[k][s] = Sp<ST>{(ST)C.k4[k][s].ar, (ST)C.k4[k][s].ai, (ST)C.k4[k][s].br, (ST)C.k4[k][s].bi};
                        } else {
                            Sp<double> nd = candidate(sp_d(C.p0[k][s]), C.k4[k][s], I.renorm);
                            C.p0[k][s] = Sp<ST>{(ST)nd.ar, (ST)nd.ai, (ST)nd.br, (ST)nd.bi};
                        }
                    }
            }
            double* m = M0;                      // the candidate's moments become psi0's
            M0 = Mn;
            Mn = m;
            ++accepted;
            ++acc_call;
        }
        ++total;
        ++steps_call;
        last_err = err;
        if (acc_ok) last_dr = dr;
        dr *= ksi * sqrt(eps0 / err);
        dr = (dr > max_dr) ? max_dr : dr;
        dr = (r + dr > Rn) ? Rn - r : dr;
        if (acc_ok) {
            bool due = false;
            if (dump_rstep > 0. && (r == Rn || r - dump_last_r > dump_rstep)) due = true;
            if (dump_itrstep > 0 && (it % dump_itrstep) == 0) due = true;
            if (due) { stop_reason = 2; break; }
        }
    }
    if (!ok) stop_reason = 3;

    // ---- write back
    if (!STREAM && my_tile0 < ntiles) {
        ST* base = (ST*)I.psi[cur];
#pragma unroll
        for (int k = 0; k < CPT; ++k) {
            if (C.bl[k] < 0) continue;
            long long c = cell_index(my_tile0, k);
#pragma unroll
            for (int s = 0; s < NS; ++s) st_sp<ST>(base, ncells, s, c, sp_d(C.p0[k][s]));
        }
    }
    if (lb == 0 && tid == 0) {
        rs->r = r;
        rs->dr = dr;
        rs->total = total;
        rs->accepted = accepted;
        rs->cur = cur;
        rs->last_err = last_err;
        rs->last_dr = last_dr;
        rs->steps_call = steps_call;
        rs->acc_call = acc_call;
        rs->stop_reason = stop_reason;
        rs->gen0 = gen;
        rs->m0_ok = ok ? 1 : 0;
        rs->m0_key = I.launch_key;
        rs->m0_r = r;
#pragma unroll
        for (int v = 0; v < NMV; ++v) rs->m0[v] = M0[v];
        if (I.rs_host) {
            // the host reads this copy after the stream drained; a failed
            // exchange anywhere also fails this workgroup's (it waits for
            // every row), so its view of the status is the launch's
            mirror_scalars(rs, I.rs_host);
            const int st = __hip_atomic_load(status, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
            I.rs_host->status = (!ok && st == 0) ? (int)XF_EBARRIER : st;
        }
    }
}


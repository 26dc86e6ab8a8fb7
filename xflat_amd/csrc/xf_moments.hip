// xf_moments.hip -- psi0's angular moments at the state's radius, for a
// stepper launch that finds none carried from an earlier launch of its shape
// (after an upload, a change of launch shape or a failed launch): the cached
// energy sums that Nmr.cpp:117-121 forms once before the radial loop and then
// carries from step to step.  Two small kernels, so the stepper itself holds
// no code for them: k_moments_part sums the cells of a fixed slice of beams
// per workgroup, k_moments_sum adds the slices in a fixed order and leaves
// the moments in the run scalars, keyed to the launch shape.
#include "xf_step_kernel.h"

namespace {

constexpr int MOM_BLOCK = 256;
constexpr int MOM_BEAMS = 16;          // beam records in LDS at once

// Workgroup (j, i): beams [j * bpw, (j + 1) * bpw) of instance i, bpw =
// ceil(nbeam / gridDim.x); the partial sums of its cells go to
// part[(i * gridDim.x + j) * NMV + v] (zeros for a workgroup without beams).
template <class ST, class CT, int NS, int GEOM>
__global__ void __launch_bounds__(MOM_BLOCK) k_moments_part(const XfInstDev* __restrict__ insts, double* __restrict__ part) {
    constexpr bool F64 = std::is_same<CT, double>::value;
    using R = BRec<GEOM>;
    constexpr int NM = R::NM;
    constexpr int NMV = NM * 3;
    constexpr int TS = tab_slots(NS);
    const XfInstDev& I = insts[blockIdx.y];
    const int tid = threadIdx.x, lane = tid & 63, wave = tid >> 6;
    const int ne = I.ne, nbeam = I.nbeam;
    const long long ncells = I.ncells;
    const XfRunScalars* rs = I.rs;
    const double r = uni_d(rs->r);
    const int cur = __builtin_amdgcn_readfirstlane(rs->cur);
    const ST* base = (const ST*)I.psi[cur];

    extern __shared__ double smem[];
    double* s_tab = smem;                                  // [ne][TS]
    double* s_rec = s_tab + ne * TS;                       // [MOM_BEAMS][R::SIZE]
    double* s_x = s_rec + MOM_BEAMS * R::SIZE;             // [MOM_BLOCK]
    for (int e = tid; e < ne; e += MOM_BLOCK) {
        double* t = s_tab + e * TS;
#pragma unroll
        for (int q = 0; q < NS; ++q) put_src_weight<F64>(t + 6 + q, I.srcw[q * ne + e], I.srcwf[q * ne + e]);
    }
    const int bpw = (nbeam + (int)gridDim.x - 1) / (int)gridDim.x;
    const int b0 = min(nbeam, (int)blockIdx.x * bpw), b1 = min(nbeam, b0 + bpw);
    double acc[NMV];
#pragma unroll
    for (int v = 0; v < NMV; ++v) acc[v] = 0.;
    for (int c0 = b0; c0 < b1; c0 += MOM_BEAMS) {
        const int nb = min(MOM_BEAMS, b1 - c0);
        __syncthreads();                                   // the previous chunk's records are read
        for (int b = tid; b < nb; b += MOM_BLOCK) beam_point<GEOM>(I, c0 + b, 0, r, s_rec + b * R::SIZE);
        __syncthreads();
        for (int q = tid; q < nb * ne; q += MOM_BLOCK) {
            const int b = q / ne, e = q - b * ne;
            const long long cell = (long long)(c0 + b) * ne + e;
            CT X0 = CT(0), X2 = CT(0), X3 = CT(0);
#pragma unroll
            for (int s = 0; s < NS; ++s) {
                const Sp<ST> a = ld_sp<ST>(base, ncells, s, cell);
                if constexpr (F64) src_add<CT>(s_tab + e * TS, s, sp_d(a), X0, X2, X3);
                else src_add<CT>(s_tab + e * TS, s, sp_f(a), X0, X2, X3);
            }
            acc_moments<NM>(acc, s_rec + b * R::SIZE + R::W, (double)X0, (double)X2, (double)X3);
        }
    }
    // the workgroup's sums, value by value in a fixed order
    double* out = part + ((size_t)blockIdx.y * gridDim.x + blockIdx.x) * NMV;
#pragma unroll
    for (int v = 0; v < NMV; ++v) {
        __syncthreads();
        s_x[tid] = acc[v];
        __syncthreads();
        if (wave == 0) {
            const double x = wave_red_dpp<false>(lds_strided<false>(s_x, lane, MOM_BLOCK));
            if (lane == 63) out[v] = x;
        }
    }
}

// One wave per instance: value v = sum over the nparts slices in order
// (lane l adding slices l, l + 64, ..., then the DPP reduction).
template <int GEOM>
__global__ void __launch_bounds__(64) k_moments_sum(const XfInstDev* __restrict__ insts, const double* __restrict__ part,
                                                     int nparts) {
    constexpr int NMV = BRec<GEOM>::NM * 3;
    const XfInstDev& I = insts[blockIdx.x];
    const int lane = threadIdx.x;
    XfRunScalars* rs = I.rs;
    const double* p = part + (size_t)blockIdx.x * nparts * NMV;
#pragma unroll
    for (int v = 0; v < NMV; ++v) {
        double x = 0.;
        for (int j = lane; j < nparts; j += 64) x += p[(size_t)j * NMV + v];
        x = wave_red_dpp<false>(x);
        if (lane == 63) rs->m0[v] = x;
    }
    if (lane == 63) {
        rs->m0_r = rs->r;
        rs->m0_key = I.launch_key;
        rs->m0_ok = 1;
    }
}

template <class ST, class CT, int NS, int GEOM>
int launch_moments(const XfInstDev* d_insts, int ninst, int nparts, int ne_max, double* d_part, hipStream_t st) {
    using R = BRec<GEOM>;
    const size_t smem = sizeof(double) * ((size_t)ne_max * tab_slots(NS) + MOM_BEAMS * R::SIZE + MOM_BLOCK);
    if (smem > XF_LDS_MAX) return xf_fail(XF_EINVAL, "moments: %d energy bins need %zu B of LDS", ne_max, smem);
    auto f = k_moments_part<ST, CT, NS, GEOM>;
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncSetAttribute (moments): %s", hipGetErrorString(e));
    }
    (void)hipGetLastError();
    hipLaunchKernelGGL(f, dim3(nparts, ninst), dim3(MOM_BLOCK), smem, st, d_insts, d_part);
    hipLaunchKernelGGL(k_moments_sum<GEOM>, dim3(ninst), dim3(64), 0, st, d_insts, d_part, nparts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xf_fail(XF_EHIP, "moments launch: %s", hipGetErrorString(e));
    return XF_OK;
}

template <class ST, class CT, int NS>
int launch_geom(int geom, const XfInstDev* d, int ninst, int nparts, int ne_max, double* part, hipStream_t st) {
    switch (geom) {
    case XF_MA: return launch_moments<ST, CT, NS, XF_MA>(d, ninst, nparts, ne_max, part, st);
    case XF_MAA: return launch_moments<ST, CT, NS, XF_MAA>(d, ninst, nparts, ne_max, part, st);
    case XF_SA: return launch_moments<ST, CT, NS, XF_SA>(d, ninst, nparts, ne_max, part, st);
    }
    return xf_fail(XF_EINVAL, "moments: geometry %d", geom);
}

template <class ST, class CT>
int launch_ns(int ns, int geom, const XfInstDev* d, int ninst, int nparts, int ne_max, double* part, hipStream_t st) {
    return ns == 4 ? launch_geom<ST, CT, 4>(geom, d, ninst, nparts, ne_max, part, st)
                   : launch_geom<ST, CT, 2>(geom, d, ninst, nparts, ne_max, part, st);
}

}  // namespace

int xf_moments_launch(const XfLaunch& L, int geom, const XfInstDev* d_insts, int ninst, int nparts, double* d_part,
                      void* stream) {
    if (ninst <= 0 || ninst > 65535 || nparts < 1 || nparts > XF_MOM_PARTS)
        return xf_fail(XF_EINVAL, "moments: %d instances x %d slices", ninst, nparts);
    hipStream_t st = (hipStream_t)stream;
    switch (L.precision) {
    case XF_FP64: return launch_ns<double, double>(L.ns, geom, d_insts, ninst, nparts, L.ne, d_part, st);
    case XF_MIXED: return launch_ns<double, float>(L.ns, geom, d_insts, ninst, nparts, L.ne, d_part, st);
    case XF_FP32: return launch_ns<float, float>(L.ns, geom, d_insts, ninst, nparts, L.ne, d_part, st);
    }
    return xf_fail(XF_EINVAL, "moments: precision %d", L.precision);
}

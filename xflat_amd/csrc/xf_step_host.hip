// xf_step_host.hip -- launch helpers of the fused stepper: kernel choice,
// LDS size, co-residency, launch.
#include <hip/hip_runtime.h>
#include <algorithm>
#include "xf_internal.h"


using XfStepFn = void (*)(const XfInstDev*, const int*, int);
// instantiation tables (xf_k_*.hip): workgroups of up to 512 threads, and
// (FP32 and mixed, 1 or 2 cells per lane) of 513..1024 threads
XfStepFn xf_pick_fp64_ma(int, int, int);
XfStepFn xf_pick_fp64_maa(int, int, int);
XfStepFn xf_pick_fp64_sa(int, int, int);
XfStepFn xf_pick_mixed_ma(int, int, int);
XfStepFn xf_pick_mixed_maa(int, int, int);
XfStepFn xf_pick_mixed_sa(int, int, int);
XfStepFn xf_pick_fp32_ma(int, int, int);
XfStepFn xf_pick_fp32_maa(int, int, int);
XfStepFn xf_pick_fp32_sa(int, int, int);
XfStepFn xf_pick_mixed_ma_wide(int, int, int);
XfStepFn xf_pick_mixed_maa_wide(int, int, int);
XfStepFn xf_pick_mixed_sa_wide(int, int, int);
XfStepFn xf_pick_fp32_ma_wide(int, int, int);
XfStepFn xf_pick_fp32_maa_wide(int, int, int);
XfStepFn xf_pick_fp32_sa_wide(int, int, int);

// The kernel for a shape: the 512-thread instantiation (up to 256 VGPRs per
// wave) when the block fits it, else the 1024-thread one (at most 128 VGPRs).
// A shape whose block is not chosen yet (0) gets the widest kernel, so that
// its thread bound is the one a planner sizes blocks against.
static XfStepFn pick(const XfLaunch& L, int geom) {
    const int ns = L.ns, cpt = L.cpt, mode = L.mode;
    XfStepFn narrow = nullptr, wide = nullptr;
    switch (L.precision) {
        case XF_FP64:
            narrow = geom == XF_SA    ? xf_pick_fp64_sa(ns, cpt, mode)
                     : geom == XF_MAA ? xf_pick_fp64_maa(ns, cpt, mode)
                                      : xf_pick_fp64_ma(ns, cpt, mode);
            break;
        case XF_FP32:
            narrow = geom == XF_SA    ? xf_pick_fp32_sa(ns, cpt, mode)
                     : geom == XF_MAA ? xf_pick_fp32_maa(ns, cpt, mode)
                                      : xf_pick_fp32_ma(ns, cpt, mode);
            wide = geom == XF_SA    ? xf_pick_fp32_sa_wide(ns, cpt, mode)
                   : geom == XF_MAA ? xf_pick_fp32_maa_wide(ns, cpt, mode)
                                    : xf_pick_fp32_ma_wide(ns, cpt, mode);
            break;
        default:
            narrow = geom == XF_SA    ? xf_pick_mixed_sa(ns, cpt, mode)
                     : geom == XF_MAA ? xf_pick_mixed_maa(ns, cpt, mode)
                                      : xf_pick_mixed_ma(ns, cpt, mode);
            wide = geom == XF_SA    ? xf_pick_mixed_sa_wide(ns, cpt, mode)
                   : geom == XF_MAA ? xf_pick_mixed_maa_wide(ns, cpt, mode)
                                    : xf_pick_mixed_ma_wide(ns, cpt, mode);
            break;
    }
    if (L.block > 0 && L.block <= 512) return narrow;
    return wide ? wide : narrow;
}

bool xf_step_has_shape(const XfLaunch& L, int geom) { return pick(L, geom) != nullptr; }

// Doubles per beam record in LDS (at least the kernel's BRec size).
static int record_doubles(int geom) {
    const int nm = geom == XF_MAA ? 4 : geom == XF_SA ? 1 : 2;
    return 3 + 3 * nm + 9 + 12 + 4 + 10;
}

// The kernel's dynamic LDS, in doubles, in its order: beam records, the
// pass-C error buffer (which doubles as the publish staging of the lanes'
// values), the per-beam error sums, the per-energy table of (6 + species) | 1
// slots per bin, and for mode 0 with FP32 increments the kept remainder
// coefficient sets (a float4 per cell slot and species, 16-byte aligned).
size_t xf_step_smem_bytes(const XfLaunch& L, int geom) {
    const size_t bsize = record_doubles(geom);
    const size_t errb = (size_t)4 * L.block * L.cpt;
    const size_t recs = L.mode == 4 ? (size_t)std::max(L.wbeams, L.bpt) : (size_t)L.bpt;   // beam records
    size_t n = recs * bsize + errb + 4 * (size_t)L.bpt + (size_t)std::max(L.ne, 1) * ((6 + L.ns) | 1);
    if (L.mode == 0 && L.precision != XF_FP64) n = ((n + 1) & ~(size_t)1) + (size_t)2 * L.cpt * L.ns * L.block;
    return sizeof(double) * n;
}

// Mode 4: the most beam records that fit the workgroup's LDS beside the rest
// of the layout, with room for the kernel's static LDS.
int xf_step_max_records(const XfLaunch& L, int geom) {
    XfLaunch Z = L;
    Z.wbeams = 0;                               // records = bpt
    const size_t rec = sizeof(double) * (size_t)record_doubles(geom);
    const size_t rest = xf_step_smem_bytes(Z, geom) - (size_t)Z.bpt * rec;
    const size_t budget = XF_LDS_MAX - 2048;
    return rest >= budget ? 0 : (int)((budget - rest) / rec);
}

int xf_step_max_blocks_geom(const XfLaunch& L, int geom, int* max_blocks) {
    XfStepFn f = pick(L, geom);
    if (!f) return xf_fail(XF_EINVAL, "no stepper instantiated for precision %d, %d species, cpt %d, mode %d",
                           L.precision, L.ns, L.cpt, L.mode);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return xf_fail(XF_EHIP, "hipGetDevice failed");
    int ncu = 0;
    if (hipDeviceGetAttribute(&ncu, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess || ncu < 1)
        return xf_fail(XF_EHIP, "hipDeviceGetAttribute(multiprocessor count) failed");
    size_t smem = xf_step_smem_bytes(L, geom);
    if (smem > XF_LDS_MAX) return xf_fail(XF_EINVAL, "shape needs %zu B of LDS, the workgroup limit is %zu", smem, XF_LDS_MAX);
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
    }
    int per = 0;
    hipError_t e = hipOccupancyMaxActiveBlocksPerMultiprocessor(&per, (const void*)f, L.block, smem);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "occupancy query failed: %s", hipGetErrorString(e));
    // The runtime's figure can overstate what LDS holds: a grid of 1-wave
    // workgroups of 26,904 B (dynamic + static) each, six per CU by that
    // figure, started five per CU (the sixth never became resident and the
    // others' exchange waits expired).  Workgroups per CU are also capped by
    // LDS counted in XF_LDS_GRANULE units, static LDS included.
    hipFuncAttributes a;
    e = hipFuncGetAttributes(&a, (const void*)f);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncGetAttributes: %s", hipGetErrorString(e));
    const size_t per_wg = (smem + a.sharedSizeBytes + XF_LDS_GRANULE - 1) / XF_LDS_GRANULE * XF_LDS_GRANULE;
    if (per_wg > 0) per = std::min<int>(per, (int)(XF_LDS_MAX / per_wg));
    *max_blocks = per * ncu;
    return XF_OK;
}

int xf_step_func_info(const XfLaunch& L, int geom, int* vgprs, int* spill_bytes) {
    XfStepFn f = pick(L, geom);
    if (!f) return xf_fail(XF_EINVAL, "no stepper for this shape");
    hipFuncAttributes a;
    hipError_t e = hipFuncGetAttributes(&a, (const void*)f);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncGetAttributes: %s", hipGetErrorString(e));
    if (vgprs) *vgprs = a.numRegs;
    if (spill_bytes) *spill_bytes = (int)a.localSizeBytes;
    return a.maxThreadsPerBlock;
}

int xf_step_launch_geom(const XfLaunch& L, int geom, const XfInstDev* d_insts, const int* d_blk_inst, int max_steps,
                        void* stream) {
    XfStepFn f = pick(L, geom);
    if (!f) return xf_fail(XF_EINVAL, "no stepper for this shape");
    size_t smem = xf_step_smem_bytes(L, geom);
    if (smem > XF_LDS_MAX) return xf_fail(XF_EINVAL, "shape needs %zu B of LDS, the workgroup limit is %zu", smem, XF_LDS_MAX);
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
    }
    // An ordinary launch of a grid that the planner sized to be co-resident
    // (xf_step_max_blocks_geom).  The kernel needs every workgroup resident
    // at once but no device-wide synchronisation resource, so it does not
    // take the runtime's cooperative-launch path, whose device resource one
    // process at a time can hold: a second process (a host program started by
    // a process that has run steppers) would be refused.  If the grid is not
    // resident after all (another kernel holds CUs), the exchanges' bounded
    // waits end the launch with XF_EBARRIER.
    void* args[] = {(void*)&d_insts, (void*)&d_blk_inst, (void*)&max_steps};
    (void)hipGetLastError();
    hipError_t e = hipLaunchKernel((const void*)f, dim3(L.grid), dim3(L.block), args, smem, (hipStream_t)stream);
    if (e != hipSuccess)
        return xf_fail(XF_EHIP, "stepper launch (grid %d x %d, %zu B LDS): %s", L.grid, L.block, smem, hipGetErrorString(e));
    return XF_OK;
}

// Status and exchange rows of every instance to zero and the generation
// count restarted.  A cleared word matches no 32-bit tag (they start at 1);
// its three low bits match the one-word format's tags of generations 7, 15,
// ..., but a slot is first read at the exchange that first writes it, with
// generation 0 to 3, so a cleared word is never taken.
__global__ void k_step_prepare(const XfInstDev* __restrict__ insts) {
    const XfInstDev& I = insts[blockIdx.y];
    if (blockIdx.x == 0 && threadIdx.x == 0) {
        I.rs->status = 0;
        I.rs->gen0 = 0u;
        I.rs->diag_where = 0;
    }
    const long long n = 2ll * XF_ROW_REPLICAS * XF_ROWW * I.nblk;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        I.rows[i] = 0ull;
}

int xf_step_prepare(const XfInstDev* d_insts, int ninst, void* stream) {
    if (ninst <= 0 || ninst > 65535) return xf_fail(XF_EINVAL, "%d instances in one launch", ninst);
    (void)hipGetLastError();    // a launch error, not one left by an earlier handled call
    hipLaunchKernelGGL(k_step_prepare, dim3(8, ninst), dim3(256), 0, (hipStream_t)stream, d_insts);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xf_fail(XF_EHIP, "prepare launch: %s", hipGetErrorString(e));
    return XF_OK;
}

// The radial loop as one graph: `launches` stepper nodes, each after the
// previous one.  The generation count continues from node to node (each
// launch leaves it in the run scalars), so the rows need no clear between
// them; a nonzero status left by one launch makes the later ones exit at once,
// and the error is reported after the graph.  The stepper nodes are ordinary
// kernel nodes, as the direct launch is (the planner sized the grid to be
// co-resident).
int xf_step_build_graph(const XfLaunch& L, int geom, const XfInstDev* d_insts, const int* d_blk_inst, int ninst,
                        int max_steps, int launches, hipGraphExec_t* exec) {
    XfStepFn f = pick(L, geom);
    if (!f) return xf_fail(XF_EINVAL, "no stepper for this shape");
    if (launches < 1 || ninst <= 0 || ninst > 65535) return xf_fail(XF_EINVAL, "graph of %d launches, %d instances", launches, ninst);
    size_t smem = xf_step_smem_bytes(L, geom);
    if (smem > XF_LDS_MAX) return xf_fail(XF_EINVAL, "shape needs %zu B of LDS, the workgroup limit is %zu", smem, XF_LDS_MAX);
    if (smem > 64 * 1024) {
        hipError_t e = hipFuncSetAttribute((const void*)f, hipFuncAttributeMaxDynamicSharedMemorySize, (int)smem);
        if (e != hipSuccess) return xf_fail(XF_EHIP, "hipFuncSetAttribute: %s", hipGetErrorString(e));
    }
    (void)hipGetLastError();
    hipGraph_t g = nullptr;
    hipError_t e = hipGraphCreate(&g, 0);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "hipGraphCreate: %s", hipGetErrorString(e));
    void* step_args[] = {(void*)&d_insts, (void*)&d_blk_inst, (void*)&max_steps};
    hipKernelNodeParams sp{};
    sp.func = (void*)f;
    sp.gridDim = dim3(L.grid);
    sp.blockDim = dim3(L.block);
    sp.sharedMemBytes = (unsigned)smem;
    sp.kernelParams = step_args;
    hipGraphNode_t prev = nullptr;
    for (int k = 0; k < launches && e == hipSuccess; ++k) {
        hipGraphNode_t node = nullptr;
        e = hipGraphAddKernelNode(&node, g, prev ? &prev : nullptr, prev ? 1 : 0, &sp);
        prev = node;
    }
    if (e == hipSuccess) e = hipGraphInstantiate(exec, g, nullptr, nullptr, 0);
    (void)hipGraphDestroy(g);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "graph of the radial loop: %s", hipGetErrorString(e));
    return XF_OK;
}

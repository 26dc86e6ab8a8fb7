// xf_noalloc.hip -- device memory of the stepper entry point that allocates
// nothing (xf_run_steps_noalloc): the launch descriptor, the workgroup ->
// instance map, the per-attempt log, the cross-workgroup exchange rows and
// the partial sums of psi0's moments live in one file-scope device array,
// `scratch`, loaded with the code object (one copy per GPU, held until the
// process ends).
//
// Host code never passes `scratch` itself to a kernel or a copy: in host
// code that name is a host-side stand-in, not a device address.  The array's
// address on the current device comes from hipGetSymbolAddress, queried once
// per device on the first call there and kept per device; every kernel
// argument below is derived from that address.
#include <hip/hip_runtime.h>

#include <mutex>

#include "xf_internal.h"

namespace {

// Layout of the array (bytes, every region 256-byte aligned).
constexpr size_t NA_DESC = 0;                                            // XfInstDev
constexpr size_t NA_DESC_BYTES = 1024;
constexpr size_t NA_BLK = NA_DESC + NA_DESC_BYTES;                       // int[XF_NOALLOC_MAX_WORKGROUPS]
constexpr size_t NA_BLK_BYTES = sizeof(int) * XF_NOALLOC_MAX_WORKGROUPS;
constexpr size_t NA_LOG = NA_BLK + NA_BLK_BYTES;                         // xf_step_rec[XF_NOALLOC_MAX_LOG]
constexpr size_t NA_LOG_BYTES = sizeof(xf_step_rec) * XF_NOALLOC_MAX_LOG;
constexpr size_t NA_ROWS = NA_LOG + NA_LOG_BYTES;                        // exchange rows
constexpr size_t NA_ROWS_BYTES =
    sizeof(unsigned long long) * 2 * XF_ROW_REPLICAS * XF_ROWW * (size_t)XF_NOALLOC_MAX_WORKGROUPS;
constexpr size_t NA_PART = NA_ROWS + NA_ROWS_BYTES;                      // moments' partial sums
constexpr size_t NA_PART_BYTES = sizeof(double) * XF_MOM_PARTS * 12;
constexpr size_t NA_TOTAL = NA_PART + NA_PART_BYTES;
static_assert(sizeof(XfInstDev) <= NA_DESC_BYTES, "descriptor slot");
static_assert(NA_BLK % 256 == 0 && NA_LOG % 256 == 0 && NA_ROWS % 256 == 0 && NA_PART % 256 == 0, "alignment");
static_assert(NA_TOTAL == XF_NOALLOC_SCRATCH_BYTES, "the size stated in the header");

constexpr int NA_MAX_DEVICES = 64;
std::mutex g_mu;
void* g_addr[NA_MAX_DEVICES] = {};

// Writes the launch descriptor (passed by value) into its slot and points
// every workgroup of the launch at instance 0.
__global__ void k_noalloc_setup(XfInstDev d, XfInstDev* __restrict__ dst, int* __restrict__ blk, int grid) {
    if (blockIdx.x == 0 && threadIdx.x == 0) *dst = d;
    for (int i = blockIdx.x * blockDim.x + threadIdx.x; i < grid; i += gridDim.x * blockDim.x) blk[i] = 0;
}

}  // namespace

static __device__ __align__(256) unsigned long long scratch[NA_TOTAL / sizeof(unsigned long long)];

int xf_noalloc_area(XfNoAllocArea* out) {
    int dev = -1;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return xf_fail(XF_EHIP, "hipGetDevice: %s", hipGetErrorString(e));
    if (dev < 0 || dev >= NA_MAX_DEVICES) return xf_fail(XF_EINVAL, "device %d: at most %d devices", dev, NA_MAX_DEVICES);
    char* base = nullptr;
    {
        std::lock_guard<std::mutex> lk(g_mu);
        if (!g_addr[dev]) {
            void* p = nullptr;
            e = hipGetSymbolAddress(&p, HIP_SYMBOL(scratch));
            if (e != hipSuccess || !p)
                return xf_fail(XF_EHIP, "hipGetSymbolAddress(scratch) on device %d: %s", dev, hipGetErrorString(e));
            g_addr[dev] = p;
        }
        base = (char*)g_addr[dev];
    }
    out->desc = (XfInstDev*)(base + NA_DESC);
    out->blk = (int*)(base + NA_BLK);
    out->log = (xf_step_rec*)(base + NA_LOG);
    out->rows = (unsigned long long*)(base + NA_ROWS);
    out->part = (double*)(base + NA_PART);
    return XF_OK;
}

int xf_noalloc_setup(const XfInstDev& d, const XfNoAllocArea& a, int grid, void* stream) {
    if (grid < 1 || grid > XF_NOALLOC_MAX_WORKGROUPS) return xf_fail(XF_EINVAL, "%d workgroups in the scratch map", grid);
    (void)hipGetLastError();
    hipLaunchKernelGGL(k_noalloc_setup, dim3(1), dim3(256), 0, (hipStream_t)stream, d, a.desc, a.blk, grid);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) return xf_fail(XF_EHIP, "scratch set-up launch: %s", hipGetErrorString(e));
    return XF_OK;
}

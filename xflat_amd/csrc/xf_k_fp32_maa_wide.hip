// Instantiations of the fused stepper for workgroups of 513..1024 threads:
// fp32 state/arithmetic, MAA geometry.  The 1024-thread bound limits a
// wave to 128 VGPRs (four waves per SIMD), so these kernels spill; the
// launcher takes them only for blocks above 512 threads.
#include "xf_step_kernel.h"

XfStepFn xf_pick_fp32_maa_wide(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 0) return k_step<float, float, 2, XF_MAA, 1, 0, 1024>;
    if (ns == 2 && cpt == 1 && mode == 1) return k_step<float, float, 2, XF_MAA, 1, 1, 1024>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<float, float, 2, XF_MAA, 1, 2, 1024>;
    if (ns == 2 && cpt == 1 && mode == 4) return k_step<float, float, 2, XF_MAA, 1, 4, 1024>;
    if (ns == 2 && cpt == 2 && mode == 0) return k_step<float, float, 2, XF_MAA, 2, 0, 1024>;
    if (ns == 2 && cpt == 2 && mode == 1) return k_step<float, float, 2, XF_MAA, 2, 1, 1024>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<float, float, 2, XF_MAA, 2, 2, 1024>;
    if (ns == 2 && cpt == 2 && mode == 4) return k_step<float, float, 2, XF_MAA, 2, 4, 1024>;
    return nullptr;
}

// Instantiations of the fused stepper: fp32 state/arithmetic, MAA geometry.
#include "xf_step_kernel.h"

XfStepFn xf_pick_fp32_maa(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 0) return k_step<float, float, 2, XF_MAA, 1, 0, 1024>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<float, float, 2, XF_MAA, 1, 2, 1024>;
    if (ns == 2 && cpt == 2 && mode == 0) return k_step<float, float, 2, XF_MAA, 2, 0, 1024>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<float, float, 2, XF_MAA, 2, 2, 1024>;
    return nullptr;
}

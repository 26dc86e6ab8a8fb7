// Instantiations of the fused stepper for workgroups of up to 512 threads:
// fp64 state, fp32 increments, MA geometry (up to 256 VGPRs per wave; mode 0
// keeps the error remainders and two increments in registers from passes
// A/B to C, modes 1, 2 and 4 recompute them in pass C).
#include "xf_step_kernel.h"

XfStepFn xf_pick_mixed_ma(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 0) return k_step<double, float, 2, XF_MA, 1, 0, 512>;
    if (ns == 2 && cpt == 1 && mode == 1) return k_step<double, float, 2, XF_MA, 1, 1, 512>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<double, float, 2, XF_MA, 1, 2, 512>;
    if (ns == 2 && cpt == 1 && mode == 4) return k_step<double, float, 2, XF_MA, 1, 4, 512>;
    if (ns == 2 && cpt == 2 && mode == 0) return k_step<double, float, 2, XF_MA, 2, 0, 512>;
    if (ns == 2 && cpt == 2 && mode == 1) return k_step<double, float, 2, XF_MA, 2, 1, 512>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<double, float, 2, XF_MA, 2, 2, 512>;
    if (ns == 2 && cpt == 2 && mode == 4) return k_step<double, float, 2, XF_MA, 2, 4, 512>;
    if (ns == 2 && cpt == 4 && mode == 0) return k_step<double, float, 2, XF_MA, 4, 0, 512>;
    if (ns == 2 && cpt == 4 && mode == 1) return k_step<double, float, 2, XF_MA, 4, 1, 512>;
    if (ns == 2 && cpt == 4 && mode == 2) return k_step<double, float, 2, XF_MA, 4, 2, 512>;
    if (ns == 2 && cpt == 4 && mode == 4) return k_step<double, float, 2, XF_MA, 4, 4, 512>;
    if (ns == 4 && cpt == 1 && mode == 1) return k_step<double, float, 4, XF_MA, 1, 1, 512>;
    if (ns == 4 && cpt == 1 && mode == 2) return k_step<double, float, 4, XF_MA, 1, 2, 512>;
    if (ns == 4 && cpt == 1 && mode == 4) return k_step<double, float, 4, XF_MA, 1, 4, 512>;
    return nullptr;
}

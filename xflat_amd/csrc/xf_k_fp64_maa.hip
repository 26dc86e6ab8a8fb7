// Instantiations of the fused stepper for workgroups of up to 512 threads:
// fp64 state/arithmetic, MAA geometry (up to 256 VGPRs per wave; mode 0
// keeps the error remainders and two increments in registers from passes
// A/B to C, modes 1, 2 and 4 recompute them in pass C).
#include "xf_step_kernel.h"

XfStepFn xf_pick_fp64_maa(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 1) return k_step<double, double, 2, XF_MAA, 1, 1, 512>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<double, double, 2, XF_MAA, 1, 2, 512>;
    if (ns == 2 && cpt == 1 && mode == 4) return k_step<double, double, 2, XF_MAA, 1, 4, 512>;
    if (ns == 2 && cpt == 2 && mode == 1) return k_step<double, double, 2, XF_MAA, 2, 1, 512>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<double, double, 2, XF_MAA, 2, 2, 512>;
    if (ns == 2 && cpt == 2 && mode == 4) return k_step<double, double, 2, XF_MAA, 2, 4, 512>;
    if (ns == 2 && cpt == 4 && mode == 1) return k_step<double, double, 2, XF_MAA, 4, 1, 512>;
    if (ns == 2 && cpt == 4 && mode == 2) return k_step<double, double, 2, XF_MAA, 4, 2, 512>;
    if (ns == 2 && cpt == 4 && mode == 4) return k_step<double, double, 2, XF_MAA, 4, 4, 512>;
    if (ns == 4 && cpt == 1 && mode == 1) return k_step<double, double, 4, XF_MAA, 1, 1, 512>;
    if (ns == 4 && cpt == 1 && mode == 2) return k_step<double, double, 4, XF_MAA, 1, 2, 512>;
    if (ns == 4 && cpt == 1 && mode == 4) return k_step<double, double, 4, XF_MAA, 1, 4, 512>;
    if (ns == 4 && cpt == 2 && mode == 1) return k_step<double, double, 4, XF_MAA, 2, 1, 512>;
    if (ns == 4 && cpt == 2 && mode == 2) return k_step<double, double, 4, XF_MAA, 2, 2, 512>;
    if (ns == 4 && cpt == 2 && mode == 4) return k_step<double, double, 4, XF_MAA, 2, 4, 512>;
    if (ns == 4 && cpt == 4 && mode == 1) return k_step<double, double, 4, XF_MAA, 4, 1, 512>;
    if (ns == 4 && cpt == 4 && mode == 2) return k_step<double, double, 4, XF_MAA, 4, 2, 512>;
    if (ns == 4 && cpt == 4 && mode == 4) return k_step<double, double, 4, XF_MAA, 4, 4, 512>;
    return nullptr;
}

// Instantiations of the fused stepper: FP64 state and arithmetic, SA geometry (one
// trajectory, src/Phy_SA.cpp; the whole state is one tile of one workgroup).
#include "xf_step_kernel.h"

XfStepFn xf_pick_fp64_sa(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 1) return k_step<double, double, 2, XF_SA, 1, 1, 512>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<double, double, 2, XF_SA, 1, 2, 512>;
    if (ns == 2 && cpt == 1 && mode == 4) return k_step<double, double, 2, XF_SA, 1, 4, 512>;
    if (ns == 2 && cpt == 2 && mode == 1) return k_step<double, double, 2, XF_SA, 2, 1, 512>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<double, double, 2, XF_SA, 2, 2, 512>;
    if (ns == 2 && cpt == 2 && mode == 4) return k_step<double, double, 2, XF_SA, 2, 4, 512>;
    if (ns == 2 && cpt == 4 && mode == 1) return k_step<double, double, 2, XF_SA, 4, 1, 512>;
    if (ns == 2 && cpt == 4 && mode == 2) return k_step<double, double, 2, XF_SA, 4, 2, 512>;
    if (ns == 2 && cpt == 4 && mode == 4) return k_step<double, double, 2, XF_SA, 4, 4, 512>;
    if (ns == 4 && cpt == 1 && mode == 1) return k_step<double, double, 4, XF_SA, 1, 1, 512>;
    if (ns == 4 && cpt == 1 && mode == 2) return k_step<double, double, 4, XF_SA, 1, 2, 512>;
    if (ns == 4 && cpt == 1 && mode == 4) return k_step<double, double, 4, XF_SA, 1, 4, 512>;
    return nullptr;
}

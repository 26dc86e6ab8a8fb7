// Instantiations of the fused stepper: fp32 state/arithmetic, MA geometry.
#include "xf_step_kernel.h"

XfStepFn xf_pick_fp32_ma(int ns, int cpt, int mode) {
    if (ns == 2 && cpt == 1 && mode == 0) return k_step<float, float, 2, XF_MA, 1, 0, 1024>;
    if (ns == 2 && cpt == 1 && mode == 2) return k_step<float, float, 2, XF_MA, 1, 2, 1024>;
    if (ns == 2 && cpt == 2 && mode == 0) return k_step<float, float, 2, XF_MA, 2, 0, 1024>;
    if (ns == 2 && cpt == 2 && mode == 2) return k_step<float, float, 2, XF_MA, 2, 2, 1024>;
    return nullptr;
}

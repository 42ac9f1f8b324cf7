// loo_m0.hip -- leave-one-out kernels for covariance model 0 (0 spherical, 1 exponential, 2 gaussian).
#include "loo_kernels.h"

namespace hpglx {

void launch_loo_model_0(const LooArgs& a, bool backward, hipStream_t st) {
    switch (nmax_for(a.k)) {
        case 8: launch_loo_m<8, 0>(a, backward, st); break;
        case 16: launch_loo_m<16, 0>(a, backward, st); break;
        case 24: launch_loo_m<24, 0>(a, backward, st); break;
        default: launch_loo_m<32, 0>(a, backward, st); break;
    }
}

}  // namespace hpglx

// loo_m1.hip -- leave-one-out kernels for covariance model 1 (0 spherical, 1 exponential, 2 gaussian).
#include "loo_kernels.h"

namespace hpglx {

void launch_loo_model_1(const LooArgs& a, bool backward, hipStream_t st) {
    switch (nmax_for(a.k)) {
        case 8: launch_loo_m<8, 1>(a, backward, st); break;
        case 16: launch_loo_m<16, 1>(a, backward, st); break;
        case 24: launch_loo_m<24, 1>(a, backward, st); break;
        default: launch_loo_m<32, 1>(a, backward, st); break;
    }
}

}  // namespace hpglx

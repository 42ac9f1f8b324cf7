// solve_n24_k0.hip -- solve kernels for max_neighbours <= 24, kind 0 (0 OK, 1 SK, 2 median IK).
#include "solve_dispatch_impl.h"

namespace hpglx {

void dispatch_solve_24_0(const SolveArgs& sa, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                           hipStream_t st) {
    dispatch_kind<24, 4, 4, 0>(sa, mixed, n, nelem, ev, st);
}

void sgs_level_24_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<24, 4, 0>(sa, la, st);
}

}  // namespace hpglx

// solve_n32_k2.hip -- solve kernels for max_neighbours <= 32, kind 2 (0 OK, 1 SK, 2 median IK).
#include "solve_dispatch_impl.h"

namespace hpglx {

void dispatch_solve_32_2(const SolveArgs& sa, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                           hipStream_t st) {
    dispatch_kind<32, 8, 8, 2>(sa, mixed, n, nelem, ev, st);
}

void sgs_level_32_2(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<32, 8, 2>(sa, la, st);
}

}  // namespace hpglx

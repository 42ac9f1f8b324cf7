// solve_n16_k0.hip -- ordinary kriging: the bucketed solve of calls with
// max_neighbours 9..16 over the buckets up to 16 neighbours (4 lanes per cell) and the
// simulation level kernel with 16 slots.
#include "solve_dispatch_impl.h"

namespace hpglx {

bool solve_sorted_16_0(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    return launch_sorted<16, 0>(sa, l, st);
}

void sgs_level_16_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<16, 4, 0>(sa, la, st);
}

}  // namespace hpglx

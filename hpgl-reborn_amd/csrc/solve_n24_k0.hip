// solve_n24_k0.hip -- ordinary kriging: the bucketed solve of calls with
// max_neighbours above 16 over the buckets up to 24 neighbours (4 lanes per cell) and the
// simulation level kernel with 24 slots.
#include "solve_dispatch_impl.h"

namespace hpglx {

bool solve_sorted_24_0(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    return launch_sorted<24, 0>(sa, l, st);
}

void sgs_level_24_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<24, 4, 0>(sa, la, st);
}

}  // namespace hpglx

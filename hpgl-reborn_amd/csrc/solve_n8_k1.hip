// solve_n8_k1.hip -- simple kriging: the bucketed solve of calls with
// max_neighbours <= 8 over the buckets up to 8 neighbours (4 lanes per cell) and the
// simulation level kernel with 8 slots;
// the FP64 re-solve of mixed-precision fallbacks up to 24 neighbours.
#include "solve_dispatch_impl.h"

namespace hpglx {

bool solve_sorted_8_1(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    return launch_sorted<8, 1>(sa, l, st);
}

void sgs_level_8_1(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<8, 4, 1>(sa, la, st);
}

void solve_fallback_b24_1(const SolveArgs& sa, long long max_cells, hipStream_t st) {
    launch_fallback<false, 1>(sa, max_cells, st);
}

}  // namespace hpglx

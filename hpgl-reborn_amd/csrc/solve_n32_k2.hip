// solve_n32_k2.hip -- median indicator kriging: the bucketed solve of the 25-32 neighbour
// bucket (4 lanes per cell), the simulation level kernel with 32 slots and
// the FP64 re-solve of mixed-precision fallbacks in that bucket.
#include "solve_dispatch_impl.h"

namespace hpglx {

bool solve_sorted_32_2(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    return launch_sorted<32, 2>(sa, l, st);
}

void sgs_level_32_2(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<32, 8, 2>(sa, la, st);
}

void solve_fallback_b32_2(const SolveArgs& sa, long long max_cells, hipStream_t st) {
    launch_fallback<true, 2>(sa, max_cells, st);
}

}  // namespace hpglx

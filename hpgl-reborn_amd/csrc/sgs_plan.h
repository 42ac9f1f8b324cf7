// sgs_plan.h -- host side of sequential Gaussian simulation: the reference's
// random path and uniform draws, and the grouping of the simulated cells into
// levels that the GPU runs in parallel.  Built with g++ without FMA
// contraction, so that every double it produces rounds as the reference's own
// g++-built host code does.
#pragma once

#include <string>
#include <vector>

namespace hpglx {

struct SgsSchedule {
    // The cells to simulate, level by level, in path order within a level.
    std::vector<int> cells;
    // Standard normal deviate of each entry of `cells`: the reference's
    // gaussian_cdf_t().inverse of the cell's uniform draw; NaN where the draw
    // lies outside (0, 1), for which the reference samples the mean itself.
    std::vector<double> z;
    // Start of each level in `cells`; back() == cells.size().
    std::vector<int> level_off;
    int max_level = 0;  // cells in the largest level
};

// Path, draws and levels for a grid of `dims` with search radiuses `r`.
// informed (N bytes, nonzero = data) and simmask (N bytes or null; a cell is
// simulated only where it is exactly 1) are host arrays.  Two cells whose
// offset lies inside the +-r box never share a level, and the earlier one on
// the path comes first, so running a level's cells in parallel gives the
// results of the one-by-one walk.  draw_thread: compute the draws and their
// deviates on a second thread while the walk assigns levels.
bool sgs_schedule(const int dims[3], const int r[3], const unsigned char* informed, const unsigned char* simmask,
                  long seed, SgsSchedule& s, std::string& err, bool draw_thread = false);

// Normal-score transforms of the informed values in place (cdf_utils.h:14-25
// with non_parametric_cdf_2_t and gaussian_cdf_t): forward before the
// simulation, back after it.
void sgs_to_normal_scores(float* v, const unsigned char* informed, long long n, const float* cdf_values,
                          const float* cdf_probs, long long cdf_size);
void sgs_from_normal_scores(float* v, const unsigned char* informed, long long n, const float* cdf_values,
                            const float* cdf_probs, long long cdf_size);
// The locally varying mean of every cell through the forward transform
// (cdf_utils.h:44-50).
void sgs_means_to_normal_scores(float* m, long long n, const float* cdf_values, const float* cdf_probs,
                                long long cdf_size);

}  // namespace hpglx

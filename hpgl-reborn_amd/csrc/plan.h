// plan.h -- host-side, data-independent part of a grid-kriging call.
//
// A plan depends only on (kind, covariance model, radiuses, max_neighbours):
// the ordered search template and the covariance table over integer offsets.
// It is built once on the host (std::sort for the template order, glibc for
// the covariance values, so both round exactly as the reference's do) and
// then reused for every grid and data batch with the same parameters.
#pragma once

#include <cstddef>
#include <cstdint>
#include <string>
#include <vector>

namespace hpglx {

enum Kind { KIND_OK = 0, KIND_SK = 1, KIND_MIK = 2 };

struct CovParams {
    int type;  // 0 spherical, 1 exponential, 2 gaussian
    double ranges[3];
    double angles[3];
    double sill;
    double nugget;
};

// Largest octant table staged in LDS (one workgroup's LDS on gfx950).
constexpr size_t kOctantMaxBytes = 160 * 1024;
// Largest +-W octant table built (entries).
constexpr size_t kOctantWideMaxEntries = size_t(32) << 20;
// Largest +-radius box the template is scanned over, and largest full
// covariance table (entries, 8 bytes each).
constexpr long long kTemplateBoxMaxEntries = 1LL << 31;
constexpr long long kCovTableMaxEntries = 1LL << 27;

struct PlanHost {
    Kind kind;
    CovParams cov;
    double transform[9];
    int r[3];  // search radiuses
    int kmax;  // max_neighbours
    // Search template: offsets from the cell, in the reference's order.
    std::vector<int> tdx, tdy, tdz;
    // Extents per axis (all <= the radius-derived bounds, so large radiuses
    // cost only what the covariance reaches):
    //   H  largest |offset| of a template entry: the padded mask's halo and
    //      the template-path neighbour reach
    //   D  largest |offset| any neighbour of either search path can have:
    //      cluster-path candidates come from the 27 clusters around the cell
    //      (|d| <= 2r - 1) and are kept only where C(d) > C(0)/100, which
    //      bounds |d| by the covariance's own reach; D >= H
    int H[3] = {0, 0, 0};
    int D[3] = {0, 0, 0};
    // Covariance of the kind's provider over the box [-W, W] (x fastest):
    // OK unclipped, SK / MIK zero outside the +-radius box.  W = 2D per axis
    // covers every neighbour pair of either search path.
    int W[3];
    std::vector<double> ctab;
    // Octant tables: with all angles zero the covariance is even in each
    // offset component separately, so C(dx,dy,dz) = C(|dx|,|dy|,|dz|) bit for
    // bit (checked entry by entry over +-W when the plan is built; octant is
    // false and the tables empty otherwise).  All are x fastest.
    //   otab_w  |d| <= W: every pair either search path can produce
    //   rank_w  |d| <= W: 0 where C <= C(0)/100, else 1 + the dense rank of C
    //           among the table's values above C(0)/100
    //   otab    |d| <= E = 2H: every pair of template-path neighbours; empty
    //           when it exceeds kOctantMaxBytes (it is staged in LDS)
    bool octant = false;
    int E[3] = {0, 0, 0};
    std::vector<double> otab_w;
    std::vector<unsigned> rank_w;
    unsigned rank_max = 0;  // largest entry of rank_w
    std::vector<double> otab;
    // Clusterizer capacity: |template| / 250.
    long long limit;
    double c0() const;  // C(0) = the sill
};

// Builds the plan; returns false with a message on invalid parameters.
bool build_plan(PlanHost& p, Kind kind, const CovParams& cov, const int radiuses[3], int kmax, std::string& err);

// Evaluates the unclipped covariance model at offset v (used for scattered
// points); returns NaN when the transform is invalid.
double model_cov(const CovParams& cov, const double t[9], double v0, double v1, double v2);
bool make_transform(const CovParams& cov, double t[9]);

}  // namespace hpglx

// variogram_plan.h -- host side of the experimental variogram
// (_cvariogram/variograms.cpp, ellipsoid.cpp): search ellipsoid, tunnel
// test, and the grid offsets of the search window with the lags each one
// feeds.  Built with g++ -ffp-contract=off so that every comparison rounds
// like the reference's.
#pragma once

#include <vector>

namespace hpglx {

struct VarioTemplate {
    double lag_width = 0, lag_separation = 0, tol_distance = 0, first_lag_distance = 0;
    int num_lags = 0;
    double dir[3][3] = {{0}};  // direction1..3
    double R[3] = {0, 0, 0};
};

// ellipsoid.cpp:37-72: direction columns of Rz(azimuth) Ry(dip) Rx(rotation).
void vario_fill_directions(double azimuth, double dip, double rotation, double dir[3][3]);

// variograms.cpp:27-40.
bool vario_in_tunnel(const VarioTemplate& t, double v0, double v1, double v2);

struct VarioOffset {
    int d[3];           // grid offset
    int lag_min, lag_max;
};

// The offsets calc_variograms visits (variograms.cpp:366-378: the search
// window's bounds converted to int, visited i, j, k ascending) that lie in
// the tunnel and feed at least one of the first `lag_count` lags
// (update_lags, variograms.cpp:326-351), in visiting order.
std::vector<VarioOffset> vario_grid_offsets(const VarioTemplate& t, int lag_count);

// init_lag_list (variograms.cpp:135-146): lag l covers [start[l], end[l]).
void vario_lag_bounds(const VarioTemplate& t, int lag_count, std::vector<double>& start, std::vector<double>& end);

}  // namespace hpglx

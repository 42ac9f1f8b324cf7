// variogram_plan.cpp -- see variogram_plan.h.
#include "variogram_plan.h"

#include <cmath>

namespace hpglx {

namespace {

// 3x3 row-major product, each entry summed from 0.0 over k ascending
// (ellipsoid.cpp:6-28).
void mat3(const double a[9], const double b[9], double c[9]) {
    double r[9];
    for (int row = 0; row < 3; ++row)
        for (int col = 0; col < 3; ++col) {
            double s = 0.0;
            for (int k = 0; k < 3; ++k) s += a[row * 3 + k] * b[k * 3 + col];
            r[row * 3 + col] = s;
        }
    for (int i = 0; i < 9; ++i) c[i] = r[i];
}

double dot3(const double v[3], const double d[3]) {
    double s = 0.0;
    for (int i = 0; i < 3; ++i) s += v[i] * d[i];
    return s;
}

}  // namespace

void vario_fill_directions(double azimuth, double dip, double rotation, double dir[3][3]) {
    const double coef = 3.14159265358 / 180.0;  // the reference's own constant
    const double ra = azimuth * coef, rd = dip * coef, rr = rotation * coef;
    const double A[9] = {std::cos(ra), -std::sin(ra), 0, std::sin(ra), std::cos(ra), 0, 0, 0, 1};
    const double B[9] = {std::cos(rd), 0, -std::sin(rd), 0, 1, 0, std::sin(rd), 0, std::cos(rd)};
    const double C[9] = {1, 0, 0, 0, std::cos(rr), -std::sin(rr), 0, std::sin(rr), std::cos(rr)};
    double M[9];
    mat3(A, B, M);
    mat3(M, C, M);
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) dir[c][i] = M[i * 3 + c];
}

bool vario_in_tunnel(const VarioTemplate& t, double v0, double v1, double v2) {
    const double v[3] = {v0, v1, v2};
    const double ss1 = dot3(v, t.dir[0]), ss2 = dot3(v, t.dir[1]), ss3 = dot3(v, t.dir[2]);
    const double s2 = ss2 / t.R[1], s3 = ss3 / t.R[2];
    const double dist = std::sqrt(s2 * s2 + s3 * s3);
    return dist <= 1.0 && t.tol_distance * dist <= ss1;
}

std::vector<VarioOffset> vario_grid_offsets(const VarioTemplate& t, int lag_count) {
    // calc_search_template_window (variograms.cpp:83-126)
    double mn[3] = {1e10, 1e10, 1e10}, mx[3] = {-1e10, -1e10, -1e10};
    for (int i = 0; i < 2; ++i)
        for (int j = -1; j < 3; j += 2)
            for (int k = -1; k < 3; k += 2) {
                double DI[3], DJ[3], DK[3], V[3];
                for (int c = 0; c < 3; ++c) {
                    DI[c] = t.dir[0][c] * t.lag_separation;
                    DI[c] = DI[c] * t.num_lags;
                    DI[c] = DI[c] * i;
                    DJ[c] = t.dir[1][c] * t.R[1];
                    DJ[c] = DJ[c] * j;
                    DK[c] = t.dir[2][c] * t.R[2];
                    DK[c] = DK[c] * k;
                    V[c] = DI[c] + DJ[c];
                    V[c] = V[c] + DK[c];
                }
                for (int c = 0; c < 3; ++c) {
                    if (V[c] < mn[c]) mn[c] = V[c];
                    if (V[c] > mx[c]) mx[c] = V[c];
                }
            }
    std::vector<VarioOffset> out;
    for (int i2 = (int)mn[0]; i2 <= mx[0]; ++i2)
        for (int j2 = (int)mn[1]; j2 <= mx[1]; ++j2)
            for (int k2 = (int)mn[2]; k2 <= mx[2]; ++k2) {
                if (!vario_in_tunnel(t, i2, j2, k2)) continue;
                const double dist = std::sqrt((double)i2 * i2 + (double)j2 * j2 + (double)k2 * k2);
                int lo = (int)std::ceil((dist - (t.lag_width / 2) - t.first_lag_distance) / t.lag_separation);
                if (lo >= lag_count) continue;
                int hi = (int)std::floor((dist + (t.lag_width / 2) - t.first_lag_distance) / t.lag_separation);
                if (hi < 0) continue;
                if (hi >= lag_count) hi = lag_count - 1;
                if (lo < 0) lo = 0;
                if (lo > hi) continue;
                out.push_back(VarioOffset{{i2, j2, k2}, lo, hi});
            }
    return out;
}

void vario_lag_bounds(const VarioTemplate& t, int lag_count, std::vector<double>& start, std::vector<double>& end) {
    start.resize(lag_count);
    end.resize(lag_count);
    for (int i = 0; i < lag_count; ++i) {
        const double d = i * t.lag_separation + t.first_lag_distance;
        start[i] = d - t.lag_width / 2;
        end[i] = d + t.lag_width / 2;
    }
}

}  // namespace hpglx

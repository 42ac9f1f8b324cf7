// plan.cpp -- builds the search template and covariance table of a plan.
//
// Compiled with g++ (not hipcc) and -ffp-contract=off so that every
// covariance value is the bit-identical double the reference computes:
// the ordering of the search template and of cluster-path candidates is a
// sort on these doubles, with frequent exact ties.
//
// Reference behaviour followed (paths relative to the reference repository):
//   transform  src/geo_bsd/hpgl/cov_model.h:16-61  (scale*((rx*ry)*rz), TNT product order)
//   norm       src/geo_bsd/hpgl/cov_model.h:63-77
//   models     src/geo_bsd/hpgl/cov_model.h:146-183 (h < 1e-4 -> sill)
//   template   src/geo_bsd/hpgl/covariance_field.h:61-106 (z,y,x push order,
//              threshold C(0)/100, std::sort on covariance descending)
//   clipping   src/geo_bsd/hpgl/precalculated_covariance.h:67-75 (SK),
//              src/geo_bsd/hpgl/covariance_field.h:144-153 (median IK)
//   limit      src/geo_bsd/hpgl/sugarbox_indexed_neighbour_lookup.h:126 (|template| / 250)
#include "plan.h"

#include <cstring>

#include <algorithm>
#include <cmath>
#include <thread>

namespace hpglx {

namespace {
const double kPi = 3.14159265358979323846264338327950288419;

void mul3(const double a[9], const double b[9], double c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s = s + a[3 * i + j] * b[3 * j + k];
            c[3 * i + k] = s;
        }
}
}  // namespace

bool make_transform(const CovParams& p, double t[9]) {
    for (int i = 0; i < 3; ++i)
        if (p.ranges[i] == 0) return false;
    auto rad = [](double g) { return (g * kPi) / 180; };
    double sc[9] = {0}, rz[9] = {0}, ry[9] = {0}, rx[9] = {0};
    sc[0] = 1.0;
    sc[4] = p.ranges[0] / p.ranges[1];
    sc[8] = p.ranges[0] / p.ranges[2];
    rz[0] = cos(rad(p.angles[0]));
    rz[4] = cos(rad(p.angles[0]));
    rz[1] = sin(rad(p.angles[0]));
    rz[3] = -sin(rad(p.angles[0]));
    rz[8] = 1;
    ry[0] = cos(rad(p.angles[1]));
    ry[8] = cos(rad(p.angles[1]));
    ry[2] = -sin(rad(p.angles[1]));
    ry[6] = sin(rad(p.angles[1]));
    ry[4] = 1;
    rx[4] = cos(rad(p.angles[2]));
    rx[8] = cos(rad(p.angles[2]));
    rx[5] = sin(rad(p.angles[2]));
    rx[7] = -sin(rad(p.angles[2]));
    rx[0] = 1;
    double xy[9], xyz[9];
    mul3(rx, ry, xy);
    mul3(xy, rz, xyz);
    mul3(sc, xyz, t);
    return true;
}

double model_cov(const CovParams& p, const double t[9], double v0, double v1, double v2) {
    double acc = 0.0;
    for (int i = 0; i < 3; ++i) {
        double d = 0.0;
        d += v0 * t[3 * i + 0];
        d += v1 * t[3 * i + 1];
        d += v2 * t[3 * i + 2];
        acc += d * d;
    }
    double h = sqrt(acc);
    if (h < 0.0001) return p.sill;
    if (p.type == 2) return (p.sill - p.nugget) * exp(-3 * pow(h / p.ranges[0], 2));
    if (p.type == 1) return (p.sill - p.nugget) * exp(-3 * h / p.ranges[0]);
    if (h > p.ranges[0]) return 0;
    return (p.sill - p.nugget) * (1 - 1.5 * h / p.ranges[0] + 0.5 * pow(h / p.ranges[0], 3));
}

double PlanHost::c0() const { return model_cov(cov, transform, 0, 0, 0); }

namespace {
// Reach of the model: a distance h_r (in the units of h in model_cov) such
// that every offset d with C(d) > C(0)/100 has |T d| < h_r.  The model
// decreases in h beyond the h < 1e-4 branch, so h_r is found by bisection;
// infinity when C stays above C(0)/100 (no bound).
double cov_reach(const CovParams& p) {
    const double thr = p.sill / 100;
    auto f = [&](double h) {
        if (h < 0.0001) return p.sill;
        if (p.type == 2) return (p.sill - p.nugget) * exp(-3 * pow(h / p.ranges[0], 2));
        if (p.type == 1) return (p.sill - p.nugget) * exp(-3 * h / p.ranges[0]);
        if (h > p.ranges[0]) return 0.0;
        return (p.sill - p.nugget) * (1 - 1.5 * h / p.ranges[0] + 0.5 * pow(h / p.ranges[0], 3));
    };
    double lo = 0.0001;
    if (!(f(lo) > thr)) return lo;
    double hi = 2 * lo;
    while (f(hi) > thr) {
        hi *= 2;
        if (hi > 1e15) return HUGE_VAL;
    }
    for (int i = 0; i < 200; ++i) {
        const double mid = 0.5 * (lo + hi);
        if (!(mid > lo && mid < hi)) break;
        if (f(mid) > thr)
            lo = mid;
        else
            hi = mid;
    }
    return hi;
}

// Largest |d_i| over the ellipsoid |T d| < reach: reach * |row i of T^-1|,
// rounded up with a margin; -1 when there is no bound.
void reach_extents(const double t[9], double reach, long long ext[3]) {
    const double det = t[0] * (t[4] * t[8] - t[5] * t[7]) - t[1] * (t[3] * t[8] - t[5] * t[6]) +
                       t[2] * (t[3] * t[7] - t[4] * t[6]);
    double inv[9];
    inv[0] = t[4] * t[8] - t[5] * t[7];
    inv[1] = t[2] * t[7] - t[1] * t[8];
    inv[2] = t[1] * t[5] - t[2] * t[4];
    inv[3] = t[5] * t[6] - t[3] * t[8];
    inv[4] = t[0] * t[8] - t[2] * t[6];
    inv[5] = t[2] * t[3] - t[0] * t[5];
    inv[6] = t[3] * t[7] - t[4] * t[6];
    inv[7] = t[1] * t[6] - t[0] * t[7];
    inv[8] = t[0] * t[4] - t[1] * t[3];
    for (int i = 0; i < 3; ++i) {
        ext[i] = -1;
        if (!(std::isfinite(reach) && std::isfinite(det) && det != 0)) continue;
        const double n = sqrt(inv[3 * i] * inv[3 * i] + inv[3 * i + 1] * inv[3 * i + 1] +
                              inv[3 * i + 2] * inv[3 * i + 2]) / fabs(det);
        const double e = reach * n * (1 + 1e-9);
        if (std::isfinite(e) && e < 1e15) ext[i] = (long long)floor(e) + 1;
    }
}
}  // namespace

bool build_plan(PlanHost& p, Kind kind, const CovParams& cov, const int radiuses[3], int kmax, std::string& err) {
    if (kind != KIND_OK && kind != KIND_SK && kind != KIND_MIK) {
        err = "unknown kriging kind";
        return false;
    }
    if (cov.type < 0 || cov.type > 2) {
        err = "Unknown covariance type";
        return false;
    }
    p.kind = kind;
    p.cov = cov;
    if (!make_transform(cov, p.transform)) {
        err = "All covariance ranges should be non-zero";
        return false;
    }
    for (int i = 0; i < 3; ++i) {
        if (radiuses[i] < 0) {
            err = "search radiuses must be >= 0";
            return false;
        }
        p.r[i] = radiuses[i];
    }
    if ((2LL * p.r[0] + 1) * (2LL * p.r[1] + 1) * (2LL * p.r[2] + 1) > kTemplateBoxMaxEntries) {
        err = "search radiuses too large: the +-radius box exceeds 2^31 offsets";
        return false;
    }
    if (kmax < 1) {
        err = "max_neighbours must be >= 1";
        return false;
    }
    p.kmax = kmax;

    // Template (covariance_field.h:61-106).  Every provider gives the same
    // value inside the +-radius box, so the unclipped model is used.
    const int rx = p.r[0], ry = p.r[1], rz = p.r[2];
    const int xd = 2 * rx + 1, yd = 2 * ry + 1;
    (void)xd;
    (void)yd;
    // Each entry carries its key (the covariance the reference's comparator
    // looks up), so the comparisons, and std::sort's permutation, are the
    // reference's.
    struct V { int x, y, z; double key; };
    std::vector<V> vec;
    const double zero_cov = model_cov(cov, p.transform, 0, 0, 0);
    const double threshold = zero_cov / 100;
    for (int z = -rz; z <= rz; ++z)
        for (int y = -ry; y <= ry; ++y)
            for (int x = -rx; x <= rx; ++x) {
                double v = model_cov(cov, p.transform, 0.0 - x, 0.0 - y, 0.0 - z);
                if (v > threshold) vec.push_back(V{-x, -y, -z, v});
            }
    std::sort(vec.begin(), vec.end(), [](const V& a, const V& b) { return a.key > b.key; });
    // The reference then cuts the sorted list at the first entry below the
    // threshold, evaluating C(-d) (covariance_field.h:93-103); C(-d) rounds
    // as C(d) (the transform is linear and only squares enter), which is the
    // kept key, above the threshold by construction: the cut keeps all.
    for (size_t i = 0; i < vec.size(); ++i)
        if (vec[i].key < threshold) {
            vec.resize(i);
            break;
        }
    p.tdx.resize(vec.size());
    p.tdy.resize(vec.size());
    p.tdz.resize(vec.size());
    for (size_t i = 0; i < vec.size(); ++i) {
        p.tdx[i] = vec[i].x;
        p.tdy[i] = vec[i].y;
        p.tdz[i] = vec[i].z;
    }
    p.limit = (long long)(vec.size() / 250);

    // Extents: H from the template; D adds the cluster path's reach, at most
    // 2r - 1 (the 27 clusters of r cells) and bounded by the covariance.
    for (int i = 0; i < 3; ++i) p.H[i] = 0;
    for (const V& o : vec) {
        p.H[0] = std::max(p.H[0], std::abs(o.x));
        p.H[1] = std::max(p.H[1], std::abs(o.y));
        p.H[2] = std::max(p.H[2], std::abs(o.z));
    }
    long long ext[3];
    reach_extents(p.transform, cov_reach(cov), ext);
    for (int i = 0; i < 3; ++i) {
        long long d = 2LL * p.r[i] - 1;
        if (ext[i] >= 0 && ext[i] < d) d = ext[i];
        if (d < p.H[i]) d = p.H[i];
        p.D[i] = (int)d;
    }
    // neighbour offsets travel packed in 11/11/10 bits (krige_kernels.h)
    if (p.D[0] > 1023 || p.D[1] > 1023 || p.D[2] > 511) {
        err = "the covariance reaches beyond 1023 cells in x or y or 511 in z; not supported";
        return false;
    }

    // Covariance table over [-W, W].
    for (int i = 0; i < 3; ++i) p.W[i] = 2 * p.D[i];
    const int sx = 2 * p.W[0] + 1, sy = 2 * p.W[1] + 1, sz = 2 * p.W[2] + 1;
    if ((long long)sx * sy * sz > kCovTableMaxEntries) {
        err = "the covariance table over the neighbour reach exceeds 2^27 entries; not supported";
        return false;
    }
    p.ctab.assign((size_t)sx * sy * sz, 0.0);
    const bool clip = kind != KIND_OK;
    auto entry = [&](int dx, int dy, int dz) {
        if (clip && (dx < -rx || dx > rx || dy < -ry || dy > ry || dz < -rz || dz > rz)) return 0.0;
        return model_cov(cov, p.transform, (double)dx, (double)dy, (double)dz);
    };
    // Planes [0, nz) split over threads.
    auto threaded = [](int nz, size_t work, auto&& planes) {
        unsigned nt = std::thread::hardware_concurrency();
        nt = nt < 1 ? 1 : (nt > 16 ? 16 : nt);
        if (work < 65536) nt = 1;
        std::vector<std::thread> pool;
        for (unsigned t = 0; t < nt; ++t) {
            const int z0 = (int)((long long)nz * t / nt), z1 = (int)((long long)nz * (t + 1) / nt);
            pool.emplace_back(planes, z0, z1);
        }
        for (auto& th : pool) th.join();
    };
    // With all angles zero the transform is diagonal but for signed zeros:
    // each component of T d is its one nonzero product (or a signed zero) and
    // only squares enter, so model_cov rounds identically for every sign of
    // each offset component (and so does the clipping box).  Then the octant
    // d >= 0 is evaluated and mirrored; it is also the octant table.
    const int wx1 = p.W[0] + 1, wy1 = p.W[1] + 1, wz1 = p.W[2] + 1;
    const size_t wn = (size_t)wx1 * wy1 * wz1;
    const bool even = cov.angles[0] == 0 && cov.angles[1] == 0 && cov.angles[2] == 0;
    std::vector<double> oct;
    if (even) {
        oct.resize(wn);
        threaded(wz1, wn, [&](int z0, int z1) {
            for (int z = z0; z < z1; ++z)
                for (int y = 0; y < wy1; ++y)
                    for (int x = 0; x < wx1; ++x) oct[((size_t)z * wy1 + y) * wx1 + x] = entry(x, y, z);
        });
        threaded(sz, (size_t)sx * sy * sz, [&](int z0, int z1) {
            for (int z = z0; z < z1; ++z) {
                const int az = std::abs(z - p.W[2]);
                for (int y = 0; y < sy; ++y) {
                    const double* orow = oct.data() + ((size_t)az * wy1 + std::abs(y - p.W[1])) * wx1;
                    double* row = p.ctab.data() + ((size_t)z * sy + y) * sx;
                    for (int x = 0; x < sx; ++x) row[x] = orow[std::abs(x - p.W[0])];
                }
            }
        });
    } else {
        threaded(sz, (size_t)sx * sy * sz, [&](int z0, int z1) {
            for (int z = z0; z < z1; ++z)
                for (int y = 0; y < sy; ++y)
                    for (int x = 0; x < sx; ++x)
                        p.ctab[((size_t)z * sy + y) * sx + x] = entry(x - p.W[0], y - p.W[1], z - p.W[2]);
        });
    }

    // Octant tables (all angles zero).
    p.octant = false;
    p.otab.clear();
    p.otab_w.clear();
    p.rank_w.clear();
    for (int i = 0; i < 3; ++i) p.E[i] = 2 * p.H[i];
    if (even && wn <= kOctantWideMaxEntries) {
        p.otab_w = std::move(oct);
        {
            p.octant = true;
            // Dense ranks of the keys above C(0)/100 (0 = filtered out): the
            // cluster path's std::sort comparator (key a > key b) and its
            // filter (key > C(0)/100) give the same answers on ranks.
            const double thr = p.c0() / 100;
            std::vector<double> vals;
            vals.reserve(wn);
            for (double v : p.otab_w)
                if (v > thr) vals.push_back(v);
            std::sort(vals.begin(), vals.end());
            vals.erase(std::unique(vals.begin(), vals.end()), vals.end());
            p.rank_w.resize(wn);
            p.rank_max = (unsigned)vals.size();
            for (size_t i = 0; i < wn; ++i) {
                const double v = p.otab_w[i];
                p.rank_w[i] = v > thr ? (unsigned)(std::lower_bound(vals.begin(), vals.end(), v) - vals.begin()) + 1u
                                      : 0u;
            }
            // The LDS table of the solve: |d| <= 2r, when it fits.
            const int ox = p.E[0] + 1, oy = p.E[1] + 1, oz = p.E[2] + 1;
            const size_t on = (size_t)ox * oy * oz;
            if (on * sizeof(double) <= kOctantMaxBytes) {
                p.otab.resize(on);
                for (int z = 0; z < oz; ++z)
                    for (int y = 0; y < oy; ++y)
                        for (int x = 0; x < ox; ++x)
                            p.otab[((size_t)z * oy + y) * ox + x] = p.otab_w[((size_t)z * wy1 + y) * wx1 + x];
            }
        }
    }
    return true;
}

}  // namespace hpglx

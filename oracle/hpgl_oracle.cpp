// hpgl_oracle.cpp -- FP64 CPU oracle for the HPGL grid-kriging hot path.
//
// TEST INFRASTRUCTURE ONLY.  This file is the checker: tests/, the smoke
// entry point and the CPU-baseline leg of bench.py load it to compare the HIP
// path against; the product library never links, loads or calls it.
//
// It is a clean-room restatement of what the reference computes for
// geo_bsd.ordinary_kriging / simple_kriging / median_ik, written from a reading
// of the reference sources (paths relative to the reference repository):
//   covariance model     src/geo_bsd/hpgl/cov_model.h:17-77, 146-183
//   clipped tables       src/geo_bsd/hpgl/precalculated_covariance.h:36-75,
//                        src/geo_bsd/hpgl/covariance_field.h:144-162
//   search template      src/geo_bsd/hpgl/covariance_field.h:61-106
//   template walk        src/geo_bsd/hpgl/sugarbox_neighbour_lookup.h:37-62
//   cluster path         src/geo_bsd/hpgl/sugarbox_indexed_neighbour_lookup.h:120-190,
//                        src/geo_bsd/hpgl/clusterizer.cpp:14-189
//   weights              src/geo_bsd/hpgl/my_kriging_weights.h:60-361
//   combine              src/geo_bsd/hpgl/combiner.h:10-25
//   drivers              src/geo_bsd/hpgl/cont_kriging.h:39-129,
//                        ordinary_kriging.cpp:14-48, simple_kriging.cpp:40-75,
//                        median_ik.cpp:24-90, calc_mean.cpp:6-30,
//                        lvm_kriging.cpp:16-39, indicator_kriging.h:88-161
//   most probable class  src/geo_bsd/hpgl/cdf_utils.cpp:11-27
//   property files       src/geo_bsd/hpgl/read_inc_file.cpp:8-193,
//                        property_writer.cpp:18-139, api.cpp:94-109
//   variogram            src/geo_bsd/_cvariogram/variograms.cpp:27-506,
//                        ellipsoid.cpp:37-72
//   order relations      src/geo_bsd/hpgl/indicator_kriging.h:34-59
//   sequential Gaussian  src/geo_bsd/hpgl/sequential_simulation.h:51-119,
//   simulation           sequential_gaussian_simulation.cpp:18-78,
//                        path_random_generator.cpp:20-94,
//                        bs_random_generator.cpp:8-24, sample.h:14-25,
//                        gaussian_distribution.h:8-93,
//                        non_parametric_cdf.h (non_parametric_cdf_2_t),
//                        cdf_utils.h:14-25, kriging_interpolation.h:269-313
//
// The reference delegates three pieces of arithmetic to third-party code; the
// oracle calls or restates them as follows:
//   * std::sort (libstdc++, GCC 11.4 here) orders the search template and the
//     cluster-path candidates.  Ties are frequent, so the oracle calls the very
//     same std::sort with the same comparator on the same input sequence.
//   * glibc exp/pow/sqrt produce the covariance keys; the oracle evaluates the
//     same expressions with the same libm (built with -ffp-contract=off).
//   * LAPACK dpotrf_/dpotrs_ (version unpinned upstream) factor and solve the
//     SPD systems; the oracle uses a textbook unblocked Cholesky instead, which
//     agrees with any LAPACK to rounding (checked against numpy in tests).
#include <algorithm>
#include <cmath>
#include <random>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace {

const double kPi = 3.14159265358979323846264338327950288419;

struct CovParams {
    int type;  // 0 spherical, 1 exponential, 2 gaussian (covariance_type.h)
    double ranges[3];
    double angles[3];
    double sill;
    double nugget;
};

// cov_model.h:16-61: t = scale * ((rot_x * rot_y) * rot_z), each product a
// plain i-k-j triple loop with the sum started at 0 (TNT matmult).
void mat3_mul(const double a[9], const double b[9], double c[9]) {
    for (int i = 0; i < 3; ++i)
        for (int k = 0; k < 3; ++k) {
            double s = 0;
            for (int j = 0; j < 3; ++j) s = s + a[3 * i + j] * b[3 * j + k];
            c[3 * i + k] = s;
        }
}

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
    mat3_mul(rx, ry, xy);
    mat3_mul(xy, rz, xyz);
    mat3_mul(sc, xyz, t);
    return true;
}

struct Model {
    CovParams p;
    double t[9];

    // cov_model.h:63-77 (transfrom_and_norm) then :146-183 (the three models).
    double operator()(double v0, double v1, double v2) const {
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
        return (p.sill - p.nugget) *
               (1 - 1.5 * h / p.ranges[0] + 0.5 * pow(h / p.ranges[0], 3));
    }
};

enum Kind { OK = 0, SK = 1, MIK = 2 };

// The three covariance providers of the drivers.  OK: the model itself,
// unclipped (ordinary_kriging.cpp:29).  SK: precalculated table over the
// +-radius box, 0 outside (precalculated_covariance.h:67-75).  MIK: the
// covariance field, also 0 outside the box (covariance_field.h:144-153).
// The three are all even functions of the offset, evaluated bit-identically.
struct Provider {
    Kind kind;
    Model m;
    int r[3];
    double operator()(int ax, int ay, int az, int bx, int by, int bz) const {
        int dx = ax - bx, dy = ay - by, dz = az - bz;
        if (kind != OK) {
            if (dx < -r[0] || dx > r[0] || dy < -r[1] || dy > r[1] || dz < -r[2] || dz > r[2])
                return 0;
        }
        return m(dx, dy, dz);
    }
};

struct Off { int x, y, z; };

// covariance_field.h:61-106.  Offsets -(x,y,z) are pushed for z, y, x
// ascending over the box when C > C(0)/100; std::sort orders them by the
// covariance looked up at the stored vector, descending.
std::vector<Off> build_template(const Provider& cov, const int r[3]) {
    const int xd = 2 * r[0] + 1, yd = 2 * r[1] + 1;
    std::vector<double> data;
    std::vector<Off> vec;
    double threshold = cov(0, 0, 0, 0, 0, 0) / 100;
    for (int z = -r[2]; z <= r[2]; ++z)
        for (int y = -r[1]; y <= r[1]; ++y)
            for (int x = -r[0]; x <= r[0]; ++x) {
                double v = cov(0, 0, 0, x, y, z);
                data.push_back(v);
                if (v > threshold) vec.push_back(Off{-x, -y, -z});
            }
    auto key = [&](const Off& o) {
        long long idx = (long long)(o.z + r[2]) * yd * xd + (long long)(o.y + r[1]) * xd + (o.x + r[0]);
        return data[(size_t)idx];
    };
    std::sort(vec.begin(), vec.end(), [&](const Off& a, const Off& b) { return key(a) > key(b); });
    double thr2 = cov(0, 0, 0, 0, 0, 0) / 100;
    for (size_t i = 0; i < vec.size(); ++i)
        if (cov(0, 0, 0, vec[i].x, vec[i].y, vec[i].z) < thr2) {
            vec.resize(i);
            break;
        }
    return vec;
}

struct Grid {
    int nx, ny, nz;
    long long n() const { return (long long)nx * ny * nz; }
    // sugarbox_grid.h:52-58, sugarbox_grid.cpp:70-74
    int index(int x, int y, int z) const {
        if ((unsigned)x < (unsigned)nx && (unsigned)y < (unsigned)ny && (unsigned)z < (unsigned)nz)
            return z * nx * ny + y * nx + x;
        return -1;
    }
    void loc(int idx, int& x, int& y, int& z) const {
        x = idx % nx;
        y = idx % (nx * ny) / nx;
        z = idx / (nx * ny);
    }
};

// clusterizer.cpp:14-189.  Cluster (x/rx, y/ry, z/rz); cluster box dims
// g/r + 2 (computed in double, truncated); membership test inclusive of the
// upper bound; a cluster is "exceeded" once it holds more than `limit` nodes.
struct Clusters {
    int mx, my, mz;
    int r[3];
    size_t limit;
    std::vector<int> count;
    std::vector<std::vector<int>> nodes;
    int index(int cx, int cy, int cz) const {
        if (!(0 <= cx && cx <= mx && 0 <= cy && cy <= my && 0 <= cz && cz <= mz)) return -1;
        return cz * mx * my + cy * mx + cx;
    }
    bool exceeded(int c) const { return (size_t)count[c] > limit; }
};

// Cholesky A = L L^T of a row-major SPD matrix (lower triangle used), then
// the two triangular solves.  Fails like dpotrf when a pivot is <= 0 or NaN.
bool cholesky(std::vector<double>& a, int n) {
    for (int j = 0; j < n; ++j) {
        double d = a[j * n + j];
        for (int k = 0; k < j; ++k) d -= a[j * n + k] * a[j * n + k];
        if (!(d > 0)) return false;
        d = sqrt(d);
        a[j * n + j] = d;
        for (int i = j + 1; i < n; ++i) {
            double s = a[i * n + j];
            for (int k = 0; k < j; ++k) s -= a[i * n + k] * a[j * n + k];
            a[i * n + j] = s / d;
        }
    }
    return true;
}

void chol_solve(const std::vector<double>& l, int n, std::vector<double>& x) {
    for (int i = 0; i < n; ++i) {
        double s = x[i];
        for (int k = 0; k < i; ++k) s -= l[i * n + k] * x[k];
        x[i] = s / l[i * n + i];
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = x[i];
        for (int k = i + 1; k < n; ++k) s -= l[k * n + i] * x[k];
        x[i] = s / l[i * n + i];
    }
}

struct Setup {
    Kind kind;
    Provider cov;
    Grid g;
    int r[3];
    int kmax;
    std::vector<Off> tmpl;
    Clusters cl;
    const unsigned char* mask;
    bool informed(int idx) const {
        if (idx < 0 || idx >= (int)g.n()) return false;
        return mask[idx] != 0;
    }
};

// clusters = false: the template walk alone (sequential simulation), where a
// radius of 0 is allowed; kriging clusters the grid by the radiuses
// (clusterizer.cpp:87-89), which must then be at least 1.
bool setup(Setup& s, int kind, const CovParams* cp, const int radiuses[3], int kmax, int nx, int ny,
           int nz, const unsigned char* mask, bool clusters = true) {
    s.kind = (Kind)kind;
    s.cov.kind = s.kind;
    s.cov.m.p = *cp;
    if (!make_transform(*cp, s.cov.m.t)) return false;
    for (int i = 0; i < 3; ++i) {
        s.r[i] = radiuses[i];
        s.cov.r[i] = radiuses[i];
        if (radiuses[i] < (clusters ? 1 : 0)) return false;
    }
    s.g = Grid{nx, ny, nz};
    s.kmax = kmax;
    s.mask = mask;
    s.tmpl = build_template(s.cov, s.r);
    if (!clusters) return true;
    Clusters& c = s.cl;
    for (int i = 0; i < 3; ++i) c.r[i] = s.r[i];
    c.mx = (int)(nx / (double)s.r[0] + 2);
    c.my = (int)(ny / (double)s.r[1] + 2);
    c.mz = (int)(nz / (double)s.r[2] + 2);
    // sugarbox_indexed_neighbour_lookup.h:126 (MAGIC_NUMBER_2 = 250)
    c.limit = s.tmpl.size() / 250;
    size_t nc = (size_t)c.mx * c.my * c.mz;
    c.count.assign(nc, 0);
    c.nodes.assign(nc, std::vector<int>());
    for (int idx = 0; idx < (int)s.g.n(); ++idx) {
        if (!mask[idx]) continue;
        int x, y, z;
        s.g.loc(idx, x, y, z);
        int ci = c.index(x / s.r[0], y / s.r[1], z / s.r[2]);
        c.count[ci]++;
        if ((size_t)c.count[ci] <= c.limit) c.nodes[ci].push_back(idx);
    }
    return true;
}

struct Entry {
    int index;
    double cov;
};

// Neighbour search for one cell; returns the neighbour node indices in the
// reference's order and whether the cluster path was taken.
void find(const Setup& s, int node, std::vector<int>& out, bool& cluster_path) {
    out.clear();
    int cx, cy, cz;
    s.g.loc(node, cx, cy, cz);
    const Clusters& c = s.cl;
    int kx = cx / s.r[0], ky = cy / s.r[1], kz = cz / s.r[2];
    bool exceeded = false;
    for (int k = kz - 1; k <= kz + 1; ++k)
        for (int j = ky - 1; j <= ky + 1; ++j)
            for (int i = kx - 1; i <= kx + 1; ++i) {
                int ci = c.index(i, j, k);
                if (ci >= 0 && c.exceeded(ci)) exceeded = true;
            }
    cluster_path = !exceeded;
    if (exceeded) {
        for (size_t t = 0; t < s.tmpl.size() && (int)out.size() < s.kmax; ++t) {
            const Off& o = s.tmpl[t];
            int idx = s.g.index(cx + o.x, cy + o.y, cz + o.z);
            if (s.informed(idx)) out.push_back(idx);
        }
        return;
    }
    std::vector<Entry> cand;
    double thr = s.cov(cx, cy, cz, cx, cy, cz) / 100;
    for (int k = kz - 1; k <= kz + 1; ++k)
        for (int j = ky - 1; j <= ky + 1; ++j)
            for (int i = kx - 1; i <= kx + 1; ++i) {
                int ci = c.index(i, j, k);
                if (ci < 0) continue;
                for (int nd : c.nodes[ci]) {
                    int x, y, z;
                    s.g.loc(nd, x, y, z);
                    double v = s.cov(cx, cy, cz, x, y, z);
                    if (v > thr) cand.push_back(Entry{nd, v});
                }
            }
    std::sort(cand.begin(), cand.end(), [](const Entry& a, const Entry& b) { return a.cov > b.cov; });
    int m = std::min((int)cand.size(), s.kmax);
    for (int i = 0; i < m; ++i) out.push_back(cand[i].index);
}

// Status codes shared with the HIP path.
enum Status { ST_KRIGED = 0, ST_NO_NEIGHBOURS = 1, ST_SINGULAR = 2, ST_INFORMED = 3 };

struct CellResult {
    int status;
    double var;
    std::vector<double> w;  // final weights
};

// my_kriging_weights.h:60-196 (SK) and :198-361 (OK), with variance.
CellResult weights(const Setup& s, int node, const std::vector<int>& nb) {
    CellResult res;
    int n = (int)nb.size();
    res.var = -1;
    if (n <= 0) {
        res.status = ST_NO_NEIGHBOURS;
        return res;
    }
    int cx, cy, cz;
    s.g.loc(node, cx, cy, cz);
    std::vector<int> xs(n), ys(n), zs(n);
    for (int i = 0; i < n; ++i) s.g.loc(nb[i], xs[i], ys[i], zs[i]);
    std::vector<double> a((size_t)n * n), b(n);
    for (int i = 0; i < n; ++i) {
        for (int j = i; j < n; ++j) {
            a[i * n + j] = s.cov(xs[i], ys[i], zs[i], xs[j], ys[j], zs[j]);
            a[j * n + i] = a[i * n + j];
        }
        b[i] = s.cov(xs[i], ys[i], zs[i], cx, cy, cz);
    }
    if (!cholesky(a, n)) {
        res.status = ST_SINGULAR;
        return res;
    }
    double c0 = s.cov(cx, cy, cz, cx, cy, cz);
    if (s.kind != OK) {
        res.w = b;
        chol_solve(a, n, res.w);
        double v = c0;
        for (int i = 0; i < n; ++i) v -= res.w[i] * b[i];
        res.var = v < 0 ? 0 : v;
        res.status = ST_KRIGED;
        return res;
    }
    std::vector<double> sw = b, ones(n, 1.0);
    chol_solve(a, n, sw);
    chol_solve(a, n, ones);
    double sum_sk = 0, sum_ones = 0;
    for (int k = 0; k < n; ++k) {
        sum_sk += sw[k];
        sum_ones += ones[k];
    }
    if (std::fabs(sum_ones) < 1e-15) {
        res.status = ST_SINGULAR;
        return res;
    }
    double mu = (sum_sk - 1) / sum_ones;
    res.w.resize(n);
    for (int k = 0; k < n; ++k) res.w[k] = sw[k] - mu * ones[k];
    double v = c0;
    for (int i = 0; i < n; ++i) v -= res.w[i] * b[i];
    v -= mu;
    res.var = v < 0 ? 0 : v;
    res.status = ST_KRIGED;
    return res;
}

// ------------------------------------------------------------------ SGS
// path_random_generator.cpp:20-60 and :89-94: a full-period linear
// congruential walk over [0, size).  The prime table the reference compiles
// in (primes.txt, not shipped) is the ascending list of primes; the part the
// generator reads (every prime up to size, and the first one beyond) is
// regenerated by a sieve.
struct PathLcg {
    long long A = 1, C = 0, M = 1, R = 0;
};

bool path_init(PathLcg& st, long long size, int seed) {
    if (size < 1 || size > 10000000) return false;
    const long long lim = 2 * size + 16;  // a prime lies in (size, 2 size]
    std::vector<char> comp((size_t)lim + 1, 0);
    std::vector<long long> pr;
    for (long long i = 2; i <= lim; ++i) {
        if (comp[(size_t)i]) continue;
        pr.push_back(i);
        for (long long j = i * i; j <= lim; j += i) comp[(size_t)j] = 1;
    }
    size_t idx = 0;
    st.A = 1;
    st.M = size;
    while (idx < pr.size()) {
        if (pr[idx] > size) break;
        if (size % pr[idx] == 0) st.A *= pr[idx];
        ++idx;
    }
    if (st.M % 4 == 0) st.A *= 2;
    st.A += 1;
    idx = 0;
    while (idx + 1 < pr.size() && pr[idx + 1] <= size / 3) ++idx;
    for (long long k = (long long)idx; k >= 0; k--) {
        st.C = pr[(size_t)k];
        if (floor(double(size) / double(st.C)) != (double(size) / double(st.C))) break;
    }
    st.R = seed % size;
    return true;
}

long long path_next(PathLcg& st) {
    const long long r = st.R;
    st.R = (st.A * st.R + st.C) % st.M;
    return r;
}

// gaussian_distribution.h:8-93: the CDF (Abramowitz-Stegun 26.2.17) and the
// quantile (rational approximation); inverse() returns the mean outside (0, 1).
double gauss_prob(double x, double mean, double var) {
    const double p = 0.2316419;
    const double b[] = {0.319381530, -0.356563782, 1.781477937, -1.821255978, 1.330274429};
    double sx = (x - mean) / sqrt(var);
    const bool neg = sx < 0;
    if (neg) sx = -sx;
    const double t = 1 / (1 + p * sx);
    const double zx = 1 / sqrt(2 * 3.14159265358) * exp(-sx * sx / 2);
    double result = 1 - zx * ((((b[4] * t + b[3]) * t + b[2]) * t + b[1]) * t + b[0]) * t;
    if (neg) result = 1 - result;
    return result;
}

double gauss_inverse(double pr, double mean, double var) {
    if (pr <= 0.0 || pr >= 1.0) return mean;
    const double ps[] = {-0.322232431088, -1.0, -0.342242088547, -0.0204231210245, -0.453642210148e-4};
    const double qs[] = {0.0993484626060, 0.588581570495, 0.531103462366, 0.103537752850, 0.38560700634e-2};
    double p2 = (pr > 0.5) ? (1 - pr) : pr;
    if (p2 < 1e-150) p2 = 1e-150;
    const double y = sqrt(log(1 / (p2 * p2)));
    double xp = y + ((((ps[4] * y + ps[3]) * y + ps[2]) * y + ps[1]) * y + ps[0]) /
                        ((((qs[4] * y + qs[3]) * y + qs[2]) * y + qs[1]) * y + qs[0]);
    if (pr < 0.5) xp = -xp;
    return xp * sqrt(var) + mean;
}

// non_parametric_cdf.h, non_parametric_cdf_2_t: piecewise-linear CDF over
// ascending values (float arithmetic), prob() and inverse().
float np_prob(const float* vals, const float* probs, long long size, float value) {
    const float* it = std::lower_bound(vals, vals + size, value);
    if (it == vals + size) return probs[size - 1];
    if (it == vals) return probs[0];
    const long long i2 = it - vals, i1 = i2 - 1;
    const float x1 = vals[i1], x2 = vals[i2], y1 = probs[i1], y2 = probs[i2];
    return y1 + (y2 - y1) / (x2 - x1) * (value - x1);
}

float np_inverse(const float* vals, const float* probs, long long size, float prob) {
    if (size <= 0) return 0;
    const float* it = std::lower_bound(probs, probs + size, prob);
    if (it == probs) return vals[0];
    if (it == probs + size) return vals[size - 1];
    const long long i2 = it - probs, i1 = i2 - 1;
    const float x1 = vals[i1], x2 = vals[i2], y1 = probs[i1], y2 = probs[i2];
    return x1 + (x2 - x1) / (y2 - y1) * (prob - y1);
}

}  // namespace

extern "C" {

typedef struct {
    int type;
    double ranges[3];
    double angles[3];
    double sill;
    double nugget;
} orc_cov_t;

// Search template for the given model and radiuses.  Returns the number of
// offsets (or -1 on invalid parameters); writes dx,dy,dz triples when `out`
// is non-null (capacity `cap` triples) and the covariance keys when `keys`
// is non-null.
int orc_template(const orc_cov_t* cp, const int radiuses[3], int* out, double* keys, int cap) {
    Setup s;
    static const unsigned char dummy = 0;
    if (!setup(s, OK, (const CovParams*)cp, radiuses, 1, 1, 1, 1, &dummy)) return -1;
    int n = (int)s.tmpl.size();
    for (int i = 0; i < n && i < cap; ++i) {
        if (out) {
            out[3 * i + 0] = s.tmpl[i].x;
            out[3 * i + 1] = s.tmpl[i].y;
            out[3 * i + 2] = s.tmpl[i].z;
        }
        if (keys) keys[i] = s.cov.m(-s.tmpl[i].x, -s.tmpl[i].y, -s.tmpl[i].z);
    }
    return n;
}

// Covariance between two integer grid offsets' difference (dx,dy,dz) for the
// unclipped model; NaN on invalid parameters.
double orc_cov(const orc_cov_t* cp, double dx, double dy, double dz) {
    Model m;
    m.p = *(const CovParams*)cp;
    if (!make_transform(m.p, m.t)) return NAN;
    return m(dx, dy, dz);
}

// Grid kriging over all cells.
//   kind      0 = ordinary, 1 = simple, 2 = median indicator
//   values    N floats (OK/SK: data; MIK: ignored, see ind)
//   ind       MIK only: nind arrays of N bytes, indicator value per cell
//   mask      N bytes, nonzero = informed
//   means     SK: means[0] is the mean (auto_mean: computed as in
//             calc_mean.cpp); MIK: means[k] is marginal P(I_k = 1)
// Outputs (each may be null):
//   out_val   N floats (OK/SK), the reference's float-accumulated estimate
//   out_mask  N bytes
//   out_prob  nind*N doubles (MIK), the double-accumulated probability
//   out_est64 nind*N doubles, the estimate m + sum w (v - m) in double
//   out_var   N doubles, kriging variance (-1 when not kriged)
//   out_count N ints, number of neighbours
//   out_nb    N*kmax ints, neighbour node indices in order, -1 padded
//   out_path  N bytes, 1 where the cluster path was taken
//   out_status N ints, see Status
// Returns 0, or -1 on invalid parameters.
//   lvm_mean  SK only, may be null: one mean per cell (lvm_kriging.cpp:16-39);
//             the centre and every neighbour then use their own mean
int orc_krige(int kind, const orc_cov_t* cp, const int radiuses[3], int kmax, int nx, int ny, int nz,
              const float* values, const unsigned char* ind, int nind, const unsigned char* mask,
              const double* means, int auto_mean, double* mean_used, float* out_val,
              unsigned char* out_mask, double* out_prob, double* out_est64, double* out_var,
              int* out_count, int* out_nb, unsigned char* out_path, int* out_status, const float* lvm_mean) {
    Setup s;
    if (kmax < 1) return -1;
    if (!setup(s, kind, (const CovParams*)cp, radiuses, kmax, nx, ny, nz, mask)) return -1;
    const int N = (int)s.g.n();
    if (kind == MIK && nind < 1) return -1;
    double mean = 0;
    if (kind == SK) {
        if (auto_mean) {
            // calc_mean.cpp:6-30: serial double sum over informed cells.
            int cnt = 0;
            double sum = 0;
            for (int i = 0; i < N; ++i)
                if (mask[i]) {
                    sum += values[i];
                    cnt += 1;
                }
            mean = cnt == 0 ? 0 : sum / cnt;
        } else {
            mean = means[0];
        }
        if (mean_used) *mean_used = mean;
    }
#pragma omp parallel for schedule(dynamic, 256)
    for (int idx = 0; idx < N; ++idx) {
        std::vector<int> nb;
        bool cpath = false;
        if (out_count) out_count[idx] = 0;
        if (out_path) out_path[idx] = 0;
        if (out_nb)
            for (int i = 0; i < kmax; ++i) out_nb[(size_t)idx * kmax + i] = -1;
        if (kind != MIK && mask[idx]) {
            // cont_kriging.h:97-100: informed cells are copied.
            if (out_val) out_val[idx] = values[idx];
            if (out_mask) out_mask[idx] = 1;
            if (out_var) out_var[idx] = 0;
            if (out_est64) out_est64[idx] = values[idx];
            if (out_status) out_status[idx] = ST_INFORMED;
            continue;
        }
        find(s, idx, nb, cpath);
        if (out_count) out_count[idx] = (int)nb.size();
        if (out_path) out_path[idx] = cpath ? 1 : 0;
        if (out_nb)
            for (size_t i = 0; i < nb.size(); ++i) out_nb[(size_t)idx * kmax + i] = nb[i];
        CellResult cr;
        if (nb.empty()) {
            cr.status = ST_NO_NEIGHBOURS;
            cr.var = -1;
        } else {
            cr = weights(s, idx, nb);
        }
        if (out_status) out_status[idx] = cr.status;
        if (out_var) out_var[idx] = cr.status == ST_KRIGED ? cr.var : -1;
        int n = (int)nb.size();
        if (kind == MIK) {
            // median_ik.cpp:60-78 with combine<uchar,double>.  On failure the
            // reference leaves `prob` uninitialised; the oracle defines it as
            // the marginal probability (the SK mean).
            for (int k = 0; k < nind; ++k) {
                const unsigned char* d = ind + (size_t)k * N;
                float mf = (float)means[k];
                double prob = (double)mf;
                double e64 = (double)mf;
                if (cr.status == ST_KRIGED) {
                    for (int i = 0; i < n; ++i) {
                        float iv = d[nb[i]] == 1 ? 1.0f : 0.0f;
                        prob += (double)(cr.w[i] * (double)(iv - mf));
                        e64 += cr.w[i] * ((double)iv - (double)mf);
                    }
                }
                if (out_prob) out_prob[(size_t)k * N + idx] = prob;
                if (out_est64) out_est64[(size_t)k * N + idx] = e64;
            }
            if (out_val) out_val[idx] = 0;
            if (out_mask) out_mask[idx] = 1;
            continue;
        }
        const bool lvm = kind == SK && lvm_mean != nullptr;
        float mf = kind == SK ? (lvm ? lvm_mean[idx] : (float)mean) : 0.0f;
        if (cr.status == ST_KRIGED) {
            // combiner.h:10-25 with result_t = float.
            float res = mf;
            double e64 = (double)mf;
            for (int i = 0; i < n; ++i) {
                float v = values[nb[i]];
                float mi = lvm ? lvm_mean[nb[i]] : mf;
                res += (float)(cr.w[i] * (v - mi));
                e64 += cr.w[i] * ((double)v - (double)mi);
            }
            if (out_val) out_val[idx] = res;
            if (out_mask) out_mask[idx] = 1;
            if (out_est64) out_est64[idx] = e64;
        } else if (kind == SK) {
            // mean_on_failure (cont_kriging.h:80-94)
            if (out_val) out_val[idx] = mf;
            if (out_mask) out_mask[idx] = 1;
            if (out_est64) out_est64[idx] = mf;
        } else {
            // undefined_on_failure: the output keeps the clone of the input.
            if (out_val) out_val[idx] = values[idx];
            if (out_mask) out_mask[idx] = mask[idx];
            if (out_est64) out_est64[idx] = NAN;
        }
    }
    return 0;
}

// indicator_kriging.h:34-59: clamp to [0,1], then one forward pass that
// replaces each inverted adjacent pair by its average.  Operates in place on
// `cells` consecutive CDFs of length k (float, as indicator_probability_t).
void orc_order_relations(float* probs, int k, long long cells) {
    for (long long c = 0; c < cells; ++c) {
        float* p = probs + c * k;
        for (int i = 0; i < k; ++i) {
            if (p[i] < 0.0)
                p[i] = 0.0f;
            else if (p[i] > 1.0)
                p[i] = 1.0f;
        }
        for (int i = 0; i + 1 < k; ++i)
            if (p[i] > p[i + 1]) {
                double avg = (p[i] + p[i + 1]) / 2.0;
                p[i] = (float)avg;
                p[i + 1] = (float)avg;
            }
    }
}

// Indicator kriging over `ncat` categories (indicator_kriging.h:88-161).
// Category c: simple kriging of (data == c) with mean (float)marg[c], the
// clipped table of covs[c] over radiuses[3c..3c+2], kmax[c] neighbours, every
// cell searched (informed ones too); the sum is accumulated in float and a
// cell that cannot be kriged gets the marginal.  Then, per cell, the order
// relations (orc_order_relations) and the first category of largest
// probability (cdf_utils.cpp:11-27).
//   out_cat   N bytes, out_mask N bytes (set to 1)
//   out_raw   ncat*N floats, probabilities before the correction (may be null)
//   out_prob  ncat*N floats, after it (may be null)
//   out_status ncat*N ints (may be null)
// Returns 0, or -1 on invalid parameters.
int orc_indicator_kriging(const unsigned char* data, const unsigned char* mask, int nx, int ny, int nz, int ncat,
                          const orc_cov_t* covs, const int* radiuses, const int* kmax, const double* marg,
                          unsigned char* out_cat, unsigned char* out_mask, float* out_raw, float* out_prob,
                          int* out_status) {
    if (ncat < 1) return -1;
    const long long N = (long long)nx * ny * nz;
    std::vector<float> prob((size_t)ncat * N);
    for (int c = 0; c < ncat; ++c) {
        Setup s;
        if (kmax[c] < 1) return -1;
        if (!setup(s, MIK, (const CovParams*)&covs[c], radiuses + 3 * c, kmax[c], nx, ny, nz, mask)) return -1;
        const float mf = (float)marg[c];
#pragma omp parallel for schedule(dynamic, 256)
        for (long long idx = 0; idx < N; ++idx) {
            std::vector<int> nb;
            bool cpath = false;
            find(s, (int)idx, nb, cpath);
            CellResult cr;
            if (nb.empty()) {
                cr.status = ST_NO_NEIGHBOURS;
            } else {
                cr = weights(s, (int)idx, nb);
            }
            float p = mf;
            if (cr.status == ST_KRIGED)
                for (size_t i = 0; i < nb.size(); ++i) {
                    float iv = data[nb[i]] == c ? 1.0f : 0.0f;
                    p += (float)(cr.w[i] * (iv - mf));
                }
            prob[(size_t)c * N + idx] = p;
            if (out_status) out_status[(size_t)c * N + idx] = cr.status;
        }
    }
    if (out_raw) std::memcpy(out_raw, prob.data(), prob.size() * sizeof(float));
#pragma omp parallel for schedule(static)
    for (long long idx = 0; idx < N; ++idx) {
        std::vector<float> p(ncat);
        for (int c = 0; c < ncat; ++c) p[c] = prob[(size_t)c * N + idx];
        orc_order_relations(p.data(), ncat, 1);
        double best = p[0];
        int best_c = 0;
        for (int c = 1; c < ncat; ++c)
            if ((double)p[c] > best) {
                best = p[c];
                best_c = c;
            }
        if (out_prob)
            for (int c = 0; c < ncat; ++c) out_prob[(size_t)c * N + idx] = p[c];
        if (out_cat) out_cat[idx] = (unsigned char)best_c;
        if (out_mask) out_mask[idx] = 1;
    }
    return 0;
}

// Simple kriging weights at arbitrary float coordinates
// (simple_kriging.cpp:18-38 via api.cpp:356-410): unclipped model.
// Returns 0 on success, 1 when the system is singular (weights zeroed),
// -1 on invalid parameters.
int orc_sk_weights(const orc_cov_t* cp, const float center[3], const float* xs, const float* ys,
                   const float* zs, int n, double* weights, double* variance) {
    Model m;
    m.p = *(const CovParams*)cp;
    if (!make_transform(m.p, m.t)) return -1;
    if (n <= 0) return 1;
    std::vector<double> a((size_t)n * n), b(n);
    auto c = [&](double ax, double ay, double az, double bx, double by, double bz) {
        return m(ax - bx, ay - by, az - bz);
    };
    for (int i = 0; i < n; ++i) {
        for (int j = i; j < n; ++j) {
            a[i * n + j] = c(xs[i], ys[i], zs[i], xs[j], ys[j], zs[j]);
            a[j * n + i] = a[i * n + j];
        }
        b[i] = c(xs[i], ys[i], zs[i], center[0], center[1], center[2]);
    }
    if (!cholesky(a, n)) {
        for (int i = 0; i < n; ++i) weights[i] = 0;
        return 1;
    }
    std::vector<double> w = b;
    chol_solve(a, n, w);
    double v = c(center[0], center[1], center[2], center[0], center[1], center[2]);
    for (int i = 0; i < n; ++i) {
        weights[i] = w[i];
        v -= w[i] * b[i];
    }
    if (variance) *variance = v < 0 ? 0 : v;
    return 0;
}

// ---------------------------------------------------------- property files
// The INC reader over an in-memory file text[0, n).  Header: lines are read
// as up to 1023 characters, each read followed by a skip of whitespace; the
// first line starting with a letter names the property.  Data: tokens of up
// to 255 non-space characters; a token starting with "--" drops the rest of
// its line, one starting with '/' ends the data (an error before `size`
// values), any other goes through sscanf "%f" (as_int = 0) or "%d"
// (as_int = 1; stored as a byte).  At the end of the text without '/', the
// last token read is converted again for every remaining value.  Masks:
// value != undefined.  Returns 0, or -1 with the message in err[0, errcap).
int orc_read_inc(const char* text, long long n, int as_int, float undef_f, int undef_i, int size, float* data_f,
                 unsigned char* data_b, unsigned char* mask, char* err, int errcap) {
    auto space = [](unsigned char c) { return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r'; };
    auto set_err = [&](const std::string& m) {
        if (err && errcap > 0) std::snprintf(err, (size_t)errcap, "%s", m.c_str());
        return -1;
    };
    long long pos = 0;
    // header
    for (;;) {
        while (pos < n && space((unsigned char)text[pos])) ++pos;
        if (pos >= n) return set_err("Property name not found.");
        long long len = 0;
        while (pos + len < n && text[pos + len] != '\n' && len < 1023) ++len;
        const unsigned char first = (unsigned char)text[pos];
        pos += len;
        const bool named = (first >= 'A' && first <= 'Z') || (first >= 'a' && first <= 'z');
        while (len == 1023) {  // the continuation reads of a long line
            while (pos < n && space((unsigned char)text[pos])) ++pos;
            len = 0;
            while (pos + len < n && text[pos + len] != '\n' && len < 1023) ++len;
            if (len == 0) break;
            pos += len;
        }
        while (pos < n && space((unsigned char)text[pos])) ++pos;
        if (named) break;
    }
    // data
    std::string tok, last;
    bool have_last = false;
    int v = 0;
    while (v < size) {
        while (pos < n && space((unsigned char)text[pos])) ++pos;
        if (pos < n) {
            tok.clear();
            while (pos < n && !space((unsigned char)text[pos]) && tok.size() < 255) tok.push_back(text[pos++]);
            last = tok;
            have_last = true;
        } else {
            if (!have_last) return set_err("Unexpected end of file.");
            tok = last;  // fscanf failed and left the buffer as it was
            if (tok.size() >= 2 && tok[0] == '-' && tok[1] == '-') return set_err("Unexpected end of file.");
        }
        if (tok.size() >= 2 && tok[0] == '-' && tok[1] == '-') {
            while (pos < n && text[pos] != '\n') ++pos;
            continue;
        }
        if (!tok.empty() && tok[0] == '/') return set_err("Unexpected end of data.");
        if (as_int) {
            int x;
            if (std::sscanf(tok.c_str(), "%d", &x) != 1) return set_err("Error parsing '" + tok + "' string.");
            data_b[v] = (unsigned char)x;
            mask[v] = x == undef_i ? 0 : 1;
        } else {
            float x;
            if (std::sscanf(tok.c_str(), "%f", &x) != 1) return set_err("Error parsing '" + tok + "' string.");
            data_f[v] = x;
            mask[v] = x == undef_f ? 0 : 1;
        }
        ++v;
    }
    return 0;
}

// Byte values equal to values[j] become j where the mask is set
// (api.cpp:94-109).
void orc_remap_bytes(unsigned char* data, const unsigned char* mask, int size, const unsigned char* values, int nv) {
    for (int i = 0; i < size; ++i) {
        if (!mask[i]) continue;
        for (int j = 0; j < nv; ++j)
            if (data[i] == values[j]) {
                data[i] = (unsigned char)j;
                break;
            }
    }
}

// The value lines of a property file (property_writer.cpp:18-26, 44-85):
// floats as "%E\n" (undefined where the mask is 0), bytes as "%d\n" of the
// signed char remap[value].  Writes at most cap bytes; returns the length.
long long orc_format_values(const float* f, const unsigned char* b, const unsigned char* mask, int size,
                            float undef_f, unsigned char undef_b, const unsigned char* remap, int remap_count,
                            char* out, long long cap) {
    long long len = 0;
    char line[64];
    for (int i = 0; i < size; ++i) {
        int k;
        if (f) {
            k = std::snprintf(line, sizeof line, "%E\n", (double)(mask[i] ? f[i] : undef_f));
        } else {
            unsigned char v = undef_b;
            if (mask[i]) v = b[i] < remap_count ? remap[b[i]] : b[i];
            k = std::snprintf(line, sizeof line, "%d\n", (int)(signed char)v);
        }
        if (len + k <= cap) std::memcpy(out + len, line, (size_t)k);
        len += k;
    }
    return len;
}

// ------------------------------------------------------ experimental variogram
// Directions of the search ellipsoid: columns of Rz(azimuth) * Ry(dip) *
// Rx(rotation), with the reference's pi to 11 digits; products summed from
// 0.0 over the inner index.
void orc_vario_directions(double azimuth, double dip, double rotation, double* dirs9) {
    const double k = 3.14159265358 / 180.0;
    const double a = azimuth * k, d = dip * k, r = rotation * k;
    const double A[9] = {cos(a), -sin(a), 0, sin(a), cos(a), 0, 0, 0, 1};
    const double B[9] = {cos(d), 0, -sin(d), 0, 1, 0, sin(d), 0, cos(d)};
    const double C[9] = {1, 0, 0, 0, cos(r), -sin(r), 0, sin(r), cos(r)};
    double AB[9], M[9];
    mat3_mul(A, B, AB);
    mat3_mul(AB, C, M);
    for (int c = 0; c < 3; ++c)
        for (int i = 0; i < 3; ++i) dirs9[3 * c + i] = M[3 * i + c];
}

namespace {
struct Tunnel {
    const double* dir;  // 3 x 3, row = direction
    double R2, R3, tol;
    bool operator()(double v0, double v1, double v2) const {
        double s[3];
        for (int k = 0; k < 3; ++k) {
            double acc = 0.0;
            acc += v0 * dir[3 * k + 0];
            acc += v1 * dir[3 * k + 1];
            acc += v2 * dir[3 * k + 2];
            s[k] = acc;
        }
        const double a = s[1] / R2, b = s[2] / R3;
        const double dist = sqrt(a * a + b * b);
        return dist <= 1.0 && tol * dist <= s[0];
    }
};
}  // namespace

// Semivariogram of a dense grid (x fastest), every informed pair at every
// in-tunnel offset of the search window, accumulated sequentially in the
// reference's order (offsets i, j, k ascending from the window's bounds
// converted to int; cells z, y, x ascending), each pair added to every lag
// whose [ceil, floor] range its distance falls in.  All pairs are used.
int orc_vario_grid(const double* dirs9, double R2, double R3, double lag_width, double lag_sep, double tol,
                   int num_lags, double first, const float* data, const unsigned char* mask, int nx, int ny, int nz,
                   int lag_count, float* result) {
    Tunnel in{dirs9, R2, R3, tol};
    double mn[3] = {1e10, 1e10, 1e10}, mx[3] = {-1e10, -1e10, -1e10};
    for (int i = 0; i < 2; ++i)
        for (int j = -1; j <= 1; j += 2)
            for (int k = -1; k <= 1; k += 2)
                for (int c = 0; c < 3; ++c) {
                    double di = dirs9[c] * lag_sep;
                    di = di * num_lags;
                    di = di * i;
                    double dj = dirs9[3 + c] * R2;
                    dj = dj * j;
                    double dk = dirs9[6 + c] * R3;
                    dk = dk * k;
                    double v = di + dj;
                    v = v + dk;
                    if (v < mn[c]) mn[c] = v;
                    if (v > mx[c]) mx[c] = v;
                }
    std::vector<double> sum(lag_count, 0.0);
    std::vector<long long> cnt(lag_count, 0);
    for (int ox = (int)mn[0]; ox <= mx[0]; ++ox)
        for (int oy = (int)mn[1]; oy <= mx[1]; ++oy)
            for (int oz = (int)mn[2]; oz <= mx[2]; ++oz) {
                if (!in(ox, oy, oz)) continue;
                const double dist = sqrt((double)ox * ox + (double)oy * oy + (double)oz * oz);
                int lo = (int)ceil((dist - lag_width / 2 - first) / lag_sep);
                int hi = (int)floor((dist + lag_width / 2 - first) / lag_sep);
                if (lo >= lag_count || hi < 0) continue;
                if (hi >= lag_count) hi = lag_count - 1;
                if (lo < 0) lo = 0;
                for (int z = 0; z < nz; ++z)
                    for (int y = 0; y < ny; ++y)
                        for (int x = 0; x < nx; ++x) {
                            const long long i1 = x + (long long)nx * (y + (long long)ny * z);
                            if (!mask[i1]) continue;
                            const int x2 = x + ox, y2 = y + oy, z2 = z + oz;
                            if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= nx || y2 >= ny || z2 >= nz) continue;
                            const long long i2 = x2 + (long long)nx * (y2 + (long long)ny * z2);
                            if (!mask[i2]) continue;
                            double d = (double)data[i1] - (double)data[i2];
                            d = d * d;
                            for (int l = lo; l <= hi; ++l) {
                                sum[l] += d;
                                cnt[l] += 1;
                            }
                        }
            }
    for (int l = 0; l < lag_count; ++l) result[l] = cnt[l] > 0 ? (float)(sum[l] / cnt[l] / 2) : 0.0f;
    return 0;
}

// Semivariogram of a point set over every ordered pair (self pairs too):
// coordinate differences in float, lag l = [l * sep + first -/+ width / 2).
int orc_vario_points(const double* dirs9, double R2, double R3, double lag_width, double lag_sep, double tol,
                     double first, const float* xs, const float* ys, const float* zs, const float* vs, int n,
                     int lag_count, float* result) {
    Tunnel in{dirs9, R2, R3, tol};
    std::vector<double> st(lag_count), en(lag_count), sum(lag_count, 0.0);
    std::vector<long long> cnt(lag_count, 0);
    for (int l = 0; l < lag_count; ++l) {
        const double c = l * lag_sep + first;
        st[l] = c - lag_width / 2;
        en[l] = c + lag_width / 2;
    }
    for (int a = 0; a < n; ++a)
        for (int b = 0; b < n; ++b) {
            const double v0 = (float)(xs[a] - xs[b]), v1 = (float)(ys[a] - ys[b]), v2 = (float)(zs[a] - zs[b]);
            if (!in(v0, v1, v2)) continue;
            double q = 0.0;
            q += v0 * v0;
            q += v1 * v1;
            q += v2 * v2;
            const double dist = sqrt(q);
            const double dv = (float)(vs[a] - vs[b]);
            const double var = dv * dv;
            for (int l = 0; l < lag_count; ++l)
                if (st[l] <= dist && dist < en[l]) {
                    sum[l] += var;
                    cnt[l] += 1;
                }
        }
    for (int l = 0; l < lag_count; ++l) result[l] = cnt[l] > 0 ? (float)(sum[l] / cnt[l] / 2) : 0.0f;
    return 0;
}

// libstdc++ std::sort permutation for a key array under "greater key first"
// ordering: perm[i] = original position of the element sorted to slot i.
void orc_std_sort_perm(const double* keys, int n, int* perm) {
    std::vector<Entry> e(n);
    for (int i = 0; i < n; ++i) e[i] = Entry{i, keys[i]};
    std::sort(e.begin(), e.end(), [](const Entry& a, const Entry& b) { return a.cov > b.cov; });
    for (int i = 0; i < n; ++i) perm[i] = e[i].index;
}


// One uniform draw as the reference takes it (bs_random_generator.cpp:19-23):
// std::uniform_real_distribution<double>(0, 1) over std::mt19937 seeded with
// (uint32) seed.  Writes the first `n` draws.
void orc_mt_uniform(long seed, int n, double* out) {
    std::mt19937 gen;
    gen.seed(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    for (int i = 0; i < n; ++i) out[i] = dist(gen);
}

// The first `n` nodes of the reference's random path over `size` nodes.
int orc_sgs_path(long long size, int seed, int n, long long* out) {
    PathLcg st;
    if (!path_init(st, size, seed)) return -1;
    for (int i = 0; i < n; ++i) out[i] = path_next(st);
    return 0;
}

void orc_gauss(const double* x, int n, double mean, double var, int inverse, double* out) {
    for (int i = 0; i < n; ++i) out[i] = inverse ? gauss_inverse(x[i], mean, var) : gauss_prob(x[i], mean, var);
}

// Sequential Gaussian simulation (sequential_simulation.h:51-119 with
// sequential_gaussian_simulation.cpp:18-78).  kind 0 = ordinary, 1 = simple
// (mean 0); covariances from the precalculated table of the radius box (0
// outside) for both kinds; neighbours by the template walk only.  values and
// mask (N each) are updated in place; simmask (may be null) selects the cells
// to simulate (== 1); cdf_size > 0 applies the normal-score transform before
// and the back transform after.  out_order (may be null, N ints) receives the
// simulated nodes in path order, -1 padded.  Returns the number simulated, or
// -1 on invalid parameters.  lvm_mean (simple kriging only, may be null):
// one mean per cell (sequential_gaussian_simulation.cpp:80-131), taken
// through the normal-score transform with the values (cdf_utils.h:44-50);
// the centre and every neighbour then use their own mean.
long long orc_sgs(int kind, const orc_cov_t* cp, const int radiuses[3], int kmax, int nx, int ny, int nz,
                  float* values, unsigned char* mask, const unsigned char* simmask, long seed,
                  const float* cdf_vals, const float* cdf_probs, long long cdf_size, int* out_order,
                  const float* lvm_mean) {
    if (kmax < 1 || (kind != OK && kind != SK) || (lvm_mean && kind != SK)) return -1;
    Setup s;
    if (!setup(s, SK, (const CovParams*)cp, radiuses, kmax, nx, ny, nz, mask, false)) return -1;
    s.kind = (Kind)kind;  // algebra of the kind, covariances clipped to the box
    const long long N = s.g.n();
    std::vector<float> lm;
    if (lvm_mean) {
        lm.assign(lvm_mean, lvm_mean + N);
        if (cdf_size > 0)
            for (long long i = 0; i < N; ++i)
                lm[(size_t)i] = (float)gauss_inverse(np_prob(cdf_vals, cdf_probs, cdf_size, lm[(size_t)i]), 0, 1);
    }
    if (cdf_size > 0)
        for (long long i = 0; i < N; ++i)
            if (mask[i]) {
                const double P = np_prob(cdf_vals, cdf_probs, cdf_size, values[i]);
                values[i] = (float)gauss_inverse(P, 0, 1);
            }
    PathLcg path;
    if (!path_init(path, N, (int)seed)) return -1;
    std::mt19937 gen;
    gen.seed(static_cast<std::mt19937::result_type>(seed));
    std::uniform_real_distribution<double> dist(0.0, 1.0);
    long long nsim = 0;
    if (out_order)
        for (long long i = 0; i < N; ++i) out_order[i] = -1;
    std::vector<int> nb;
    for (long long counter = 0; counter < N; ++counter) {
        const int node = (int)path_next(path);
        if (mask[node]) continue;
        if (simmask && simmask[node] != 1) continue;
        int cx, cy, cz;
        s.g.loc(node, cx, cy, cz);
        nb.clear();
        for (size_t t = 0; t < s.tmpl.size() && (int)nb.size() < s.kmax; ++t) {
            const Off& o = s.tmpl[t];
            const int idx = s.g.index(cx + o.x, cy + o.y, cz + o.z);
            if (s.informed(idx)) nb.push_back(idx);
        }
        CellResult cr;
        if (nb.empty()) {
            cr.status = ST_NO_NEIGHBOURS;
        } else {
            cr = weights(s, node, nb);
        }
        const double u = dist(gen);
        double value;
        if (cr.status == ST_KRIGED) {
            // combine (combiner.h:10-25): mean 0 (single_mean_t(0) / no_mean_t)
            // or the cells' own means
            float est = lvm_mean ? lm[(size_t)node] : 0.0f;
            for (size_t i = 0; i < nb.size(); ++i)
                est += (float)(cr.w[i] * (values[nb[i]] - (lvm_mean ? lm[(size_t)nb[i]] : 0.0f)));
            value = gauss_inverse(u, (double)est, cr.var);
        } else {
            value = gauss_inverse(u, 0, 1);
        }
        values[node] = (float)value;
        mask[node] = 1;
        if (out_order) out_order[nsim] = node;
        ++nsim;
    }
    if (cdf_size > 0)
        for (long long i = 0; i < N; ++i)
            if (mask[i]) {
                const double P = gauss_prob(values[i], 0, 1);
                values[i] = np_inverse(cdf_vals, cdf_probs, cdf_size, (float)P);
            }
    return nsim;
}

}  // extern "C"

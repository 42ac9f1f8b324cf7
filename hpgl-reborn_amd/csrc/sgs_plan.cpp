// sgs_plan.cpp -- see sgs_plan.h.  Reference behaviour followed:
//   path       src/geo_bsd/hpgl/path_random_generator.cpp:20-94
//   draws      src/geo_bsd/hpgl/bs_random_generator.cpp:8-24, sample.h:14-25
//   deviates   src/geo_bsd/hpgl/gaussian_distribution.h:56-89
//   skips      src/geo_bsd/hpgl/sequential_simulation.h:83-90
//   transforms src/geo_bsd/hpgl/cdf_utils.h:14-25, non_parametric_cdf.h
//              (non_parametric_cdf_2_t), gaussian_distribution.h:24-41
#include "sgs_plan.h"

#include <algorithm>
#include <cmath>
#include <limits>
#include <random>
#include <thread>

namespace hpglx {
namespace {

// The reference walks its nodes with R <- (A R + C) mod size, R0 = seed mod
// size, where A - 1 is the product of the distinct primes dividing size
// (doubled when 4 divides size) and C is the largest prime not above size / 3
// that does not divide size; its prime table (primes.txt, not shipped with
// it) is the ascending list of primes, regenerated here by a sieve.
struct Walk {
    long long a = 1, c = 0, m = 1, r = 0;
    long long next() {
        const long long out = r;
        r = (a * r + c) % m;
        return out;
    }
};

bool make_walk(long long size, int seed, Walk& w, std::string& err) {
    if (size < 1 || size > 10000000) {
        err = "sequential simulation supports grids of 1 to 10,000,000 cells (the reference's path limit)";
        return false;
    }
    if (seed < 0) {
        err = "the simulation seed must be non-negative";
        return false;
    }
    // odd numbers only: bit i stands for 2 i + 1
    const long long top = 2 * size + 16;
    std::vector<bool> odd_composite((size_t)(top / 2 + 1), false);
    for (long long i = 3; i * i <= top; i += 2)
        if (!odd_composite[(size_t)(i / 2)])
            for (long long j = i * i; j <= top; j += 2 * i) odd_composite[(size_t)(j / 2)] = true;
    auto is_prime = [&](long long v) { return v == 2 || (v > 2 && (v & 1) && !odd_composite[(size_t)(v / 2)]); };
    w.m = size;
    w.a = 1;
    for (long long p = 2; p <= size; ++p)
        if (is_prime(p) && size % p == 0) w.a *= p;
    if (size % 4 == 0) w.a *= 2;
    w.a += 1;
    // the largest prime <= size / 3 (2 when there is none), then downwards
    // to the first that does not divide size (2 if every one does)
    long long start = 2;
    for (long long p = size / 3; p >= 3; --p)
        if (is_prime(p)) {
            start = p;
            break;
        }
    w.c = 2;
    for (long long p = start; p >= 2; --p) {
        if (!is_prime(p)) continue;
        w.c = p;
        if (std::floor(double(size) / double(p)) != double(size) / double(p)) break;
    }
    w.r = seed % size;
    return true;
}

// gaussian_cdf_t(0, 1).inverse(u): the quantile by rational approximation;
// NaN stands for the reference's "return the mean" outside (0, 1).
double standard_deviate(double u) {
    if (u <= 0.0 || u >= 1.0) return std::numeric_limits<double>::quiet_NaN();
    static const double ps[] = {-0.322232431088, -1.0, -0.342242088547, -0.0204231210245, -0.453642210148e-4};
    static const double qs[] = {0.0993484626060, 0.588581570495, 0.531103462366, 0.103537752850,
                                0.38560700634e-2};
    double q = u > 0.5 ? 1 - u : u;
    if (q < 1e-150) q = 1e-150;
    const double y = std::sqrt(std::log(1 / (q * q)));
    double x = y + ((((ps[4] * y + ps[3]) * y + ps[2]) * y + ps[1]) * y + ps[0]) /
                       ((((qs[4] * y + qs[3]) * y + qs[2]) * y + qs[1]) * y + qs[0]);
    if (u < 0.5) x = -x;
    return x * std::sqrt(1.0) + 0.0;
}

// gaussian_cdf_t(0, 1)(x): Abramowitz-Stegun 26.2.17.
double standard_cdf(double x) {
    static const double b[] = {0.319381530, -0.356563782, 1.781477937, -1.821255978, 1.330274429};
    double s = (x - 0) / std::sqrt(1.0);
    const bool neg = s < 0;
    if (neg) s = -s;
    const double t = 1 / (1 + 0.2316419 * s);
    const double zx = 1 / std::sqrt(2 * 3.14159265358) * std::exp(-s * s / 2);
    const double r = 1 - zx * ((((b[4] * t + b[3]) * t + b[2]) * t + b[1]) * t + b[0]) * t;
    return neg ? 1 - r : r;
}

// non_parametric_cdf_2_t: linear between the table's points, float
// arithmetic, the end values beyond them.
float table_prob(const float* xv, const float* pv, long long m, float x) {
    const long long i = std::lower_bound(xv, xv + m, x) - xv;
    if (i == m) return pv[m - 1];
    if (i == 0) return pv[0];
    const float x1 = xv[i - 1], x2 = xv[i], y1 = pv[i - 1], y2 = pv[i];
    return y1 + (y2 - y1) / (x2 - x1) * (x - x1);
}

float table_value(const float* xv, const float* pv, long long m, float p) {
    if (m <= 0) return 0;
    const long long i = std::lower_bound(pv, pv + m, p) - pv;
    if (i == 0) return xv[0];
    if (i == m) return xv[m - 1];
    const float x1 = xv[i - 1], x2 = xv[i], y1 = pv[i - 1], y2 = pv[i];
    return x1 + (x2 - x1) / (y2 - y1) * (p - y1);
}

}  // namespace

bool sgs_schedule(const int dims[3], const int r[3], const unsigned char* informed, const unsigned char* simmask,
                  long seed, SgsSchedule& s, std::string& err, bool draw_thread) {
    const long long nx = dims[0], ny = dims[1], nz = dims[2];
    const long long n = nx * ny * nz;
    Walk walk;
    if (!make_walk(n, (int)seed, walk, err)) return false;
    // The uniform draws, one per simulated cell in path order, and their
    // deviates: independent of the walk, so optionally on a thread of their
    // own.  Every cell the walk can simulate is counted; a walk that repeats
    // nodes simulates fewer and leaves the last deviates unused.
    long long eligible = 0;
    for (long long i = 0; i < n; ++i) eligible += !informed[i] && (!simmask || simmask[i] == 1);
    std::vector<double> dev((size_t)eligible);
    auto draws = [&dev, seed]() {
        std::mt19937 gen;
        gen.seed(static_cast<std::mt19937::result_type>(seed));
        std::uniform_real_distribution<double> uni(0.0, 1.0);
        for (double& d : dev) d = standard_deviate(uni(gen));
    };
    std::thread drawer;
    struct JoinOnExit {  // also when an allocation below throws
        std::thread& t;
        ~JoinOnExit() {
            if (t.joinable()) t.join();
        }
    } join_on_exit{drawer};
    if (draw_thread && eligible > 0)
        drawer = std::thread(draws);
    else
        draws();

    // Levels: a cell's level is one more than the highest level among the
    // cells simulated before it inside its +-r box.  The maximum is taken
    // over blocks of about r/5 cells in x and y and r/4 in z, which can only
    // over-separate; finer blocks give fewer levels for more host time
    // (100x100x50, r = (20, 20, 4): 22,552 levels in 0.09 s; blocks of r/2,
    // 27,965 levels; exact levels 20,235 in 4 s).
    const int bx = std::max(1, (r[0] + 1) / 5), by = std::max(1, (r[1] + 1) / 5), bz = std::max(1, (r[2] + 1) / 4);
    const long long mx = (nx + bx - 1) / bx, my = (ny + by - 1) / by, mz = (nz + bz - 1) / bz;
    // win[c][b][a] = the highest level placed in blocks a - hx .. a + hx of
    // row b of block layer c.  The +-r cells in x around a cell of block a
    // lie within those blocks (hx = ceil(r / bx)), so a query reads one entry
    // per block row of its y and z range, and placing a level raises the
    // 2 hx + 1 entries whose window holds its block (levels only ever rise).
    const int hx = (r[0] + bx - 1) / bx;
    std::vector<int> win((size_t)(mx * my * mz), 0);
    std::vector<int> lv;
    std::vector<int> order;
    lv.reserve((size_t)eligible);
    order.reserve((size_t)eligible);
    int levels = 0;
    // Where every prime up to size / 3 divides the size (2, 4, 6, 8, 12, ...)
    // the reference's walk repeats nodes and misses others; a repeat finds
    // the node already simulated and is skipped, as in the reference.
    std::vector<unsigned char> taken((size_t)n, 0);
    for (long long k = 0; k < n; ++k) {
        const long long node = walk.next();
        if (informed[node] || taken[(size_t)node]) continue;
        if (simmask && simmask[node] != 1) continue;
        taken[(size_t)node] = 1;
        const long long x = node % nx, y = (node / nx) % ny, z = node / (nx * ny);
        const long long a0 = x / bx, b0 = y / by;
        const long long z0 = std::max(0LL, z - r[2]) / bz, z1 = std::min(nz - 1, z + r[2]) / bz;
        const long long y0 = std::max(0LL, y - r[1]) / by, y1 = std::min(ny - 1, y + r[1]) / by;
        int m = 0;
        for (long long c = z0; c <= z1; ++c)
            for (long long b = y0; b <= y1; ++b) m = std::max(m, win[(size_t)((c * my + b) * mx + a0)]);
        const int v = m + 1;
        int* row = &win[(size_t)(((z / bz) * my + b0) * mx)];
        for (long long a = std::max(0LL, a0 - hx); a <= std::min(mx - 1, a0 + hx); ++a) row[a] = std::max(row[a], v);
        lv.push_back(m);
        order.push_back((int)node);
        levels = std::max(levels, m + 1);
    }
    if (drawer.joinable()) drawer.join();
    // Group by level, path order kept within each (a counting sort).
    s.level_off.assign((size_t)levels + 1, 0);
    for (int l : lv) s.level_off[(size_t)l + 1]++;
    s.max_level = 0;
    for (int l = 0; l < levels; ++l) {
        s.max_level = std::max(s.max_level, s.level_off[(size_t)l + 1]);
        s.level_off[(size_t)l + 1] += s.level_off[(size_t)l];
    }
    s.cells.assign(order.size(), 0);
    s.z.assign(order.size(), 0.0);
    std::vector<int> fill(s.level_off.begin(), s.level_off.end() - 1);
    for (size_t i = 0; i < order.size(); ++i) {
        const int pos = fill[(size_t)lv[i]]++;
        s.cells[(size_t)pos] = order[i];
        s.z[(size_t)pos] = dev[i];
    }
    return true;
}

void sgs_to_normal_scores(float* v, const unsigned char* informed, long long n, const float* cdf_values,
                          const float* cdf_probs, long long cdf_size) {
    for (long long i = 0; i < n; ++i)
        if (informed[i]) {
            const double p = table_prob(cdf_values, cdf_probs, cdf_size, v[i]);
            const double x = standard_deviate(p);
            v[i] = (float)(std::isnan(x) ? 0.0 : x);
        }
}

void sgs_means_to_normal_scores(float* m, long long n, const float* cdf_values, const float* cdf_probs,
                                long long cdf_size) {
    for (long long i = 0; i < n; ++i) {
        const double x = standard_deviate(table_prob(cdf_values, cdf_probs, cdf_size, m[i]));
        m[i] = (float)(std::isnan(x) ? 0.0 : x);
    }
}

void sgs_from_normal_scores(float* v, const unsigned char* informed, long long n, const float* cdf_values,
                            const float* cdf_probs, long long cdf_size) {
    for (long long i = 0; i < n; ++i)
        if (informed[i]) v[i] = table_value(cdf_values, cdf_probs, cdf_size, (float)standard_cdf(v[i]));
}

}  // namespace hpglx

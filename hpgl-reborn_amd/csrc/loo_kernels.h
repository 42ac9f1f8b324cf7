// loo_kernels.h -- leave-one-out ordinary kriging of scattered points, for
// many variogram parameter sets at once, with its adjoint (kernels; the
// host side is loo.hip).
//
// The reference has no such path; it is built on the reference's covariance
// model (src/geo_bsd/hpgl/cov_model.h:146-183: C = sill at h < 1e-4, else
// (sill - nugget) * rho(h / r0) with h the anisotropic distance) and on its
// OK system (src/geo_bsd/hpgl/my_kriging_weights.h:198-361).
//
// For point i with neighbours n_1..n_k (its k nearest other points, fixed for
// every parameter set) and parameters theta = (sill, nugget, rx, ry, rz):
//   A_ab = C(x_a - x_b), b_a = C(x_i - x_a)
//   [A 1; 1' 0] [w; mu] = [b; 1],  z_i = sum_a w_a v_a      (forward)
// With y = [w; mu] and K the bordered matrix, dz = lambda_w' (db - dA w) where
// K lambda = [v; 0] (K symmetric): the adjoint is one more solve with the
// forward Cholesky factor of A (Schur complement as in the forward), and the
// derivatives of C are recomputed pair by pair rather than stored.
//   dC/dsill = rho, dC/dnugget = -rho, dC/dr_k = (sill - nugget) rho'(x) dx/dr_k,
//   x = sqrt(sum_k (d_k / r_k)^2), dx/dr_k = -(d_k^2 / r_k^3) / x;
//   on the diagonal C = sill: dC/dsill = 1, all others 0.
// Layout: a group of G lanes per system, rows spread over lanes as in the
// grid solver (solve.h); all arithmetic in FP64.
#pragma once

#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "loo.h"
#include "solve.h"

namespace hpglx {

// rho and rho' of the unit-range models (cov_model.h:146-183).
template <int MODEL>
__device__ __forceinline__ void model_rho(double x, double& r, double& dr) {
    if (MODEL == 0) {  // spherical
        if (x < 1.0) {
            r = 1.0 - 1.5 * x + 0.5 * x * x * x;
            dr = -1.5 + 1.5 * x * x;
        } else {
            r = 0.0;
            dr = 0.0;
        }
    } else if (MODEL == 1) {  // exponential
        r = exp(-3.0 * x);
        dr = -3.0 * r;
    } else {  // gaussian
        r = exp(-3.0 * x * x);
        dr = -6.0 * x * r;
    }
}

struct LooParams {
    double sill, nugget;
    double ir[3];  // 1 / range
    double r0;     // first range
};

__device__ __forceinline__ LooParams load_params(const double* prm, int p) {
    LooParams q;
    q.sill = prm[5 * p + 0];
    q.nugget = prm[5 * p + 1];
    q.ir[0] = 1.0 / prm[5 * p + 2];
    q.ir[1] = 1.0 / prm[5 * p + 3];
    q.ir[2] = 1.0 / prm[5 * p + 4];
    q.r0 = prm[5 * p + 2];
    return q;
}

// Covariance of a pair of points: sill when the anisotropic distance
// h = r0 * x is below 1e-4 (the model's nugget convention), else
// (sill - nugget) * rho(x).
template <int MODEL>
__device__ __forceinline__ double pair_cov(const LooParams& q, double dx, double dy, double dz) {
    const double ux = dx * q.ir[0], uy = dy * q.ir[1], uz = dz * q.ir[2];
    const double x = sqrt(ux * ux + uy * uy + uz * uz);
    if (x * q.r0 < 1e-4) return q.sill;
    double r, dr;
    model_rho<MODEL>(x, r, dr);
    return (q.sill - q.nugget) * r;
}

// Derivatives of an off-diagonal covariance w.r.t. (sill, nugget, rx, ry, rz),
// scaled by `s` and added to g[5].
template <int MODEL>
__device__ __forceinline__ void pair_grad(const LooParams& q, double dx, double dy, double dz, double s,
                                          double (&g)[5]) {
    const double ux = dx * q.ir[0], uy = dy * q.ir[1], uz = dz * q.ir[2];
    const double x2 = ux * ux + uy * uy + uz * uz;
    const double x = sqrt(x2);
    if (x * q.r0 < 1e-4) {
        g[0] += s;
        return;
    }
    double r, dr;
    model_rho<MODEL>(x, r, dr);
    g[0] += s * r;
    g[1] -= s * r;
    if (x > 0.0) {
        const double f = s * (q.sill - q.nugget) * dr / x;
        // dx/dr_k = -(d_k^2 / r_k^3) / x = -(u_k^2 / r_k) / x
        g[2] -= f * ux * ux * q.ir[0];
        g[3] -= f * uy * uy * q.ir[1];
        g[4] -= f * uz * uz * q.ir[2];
    }
}

// Per-system state kept between forward and backward, element e of lane t of
// system sys at state[(e * nsys + sys) * G + t]: the factor (TOT), 1/diag (S),
// w (S), o = A^-1 1 (S), then sum(o) (lane 0's copy).
template <int NMAX, int G>
struct LooState {
    static constexpr int S = NMAX / G;
    static constexpr int TOT = slot_off<G>(S);
    static constexpr int N = TOT + 3 * S + 1;
};

template <int NMAX, int G>
__device__ __forceinline__ void load_system(const LooArgs& a, int i, int t, double (&px)[NMAX / G],
                                            double (&py)[NMAX / G], double (&pz)[NMAX / G],
                                            double (&pv)[NMAX / G], double& xi, double& yi, double& zi) {
    constexpr int S = NMAX / G;
    int nb[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int r = G * s + t;
        nb[s] = a.nbr[(long long)i * a.k + (r < a.k ? r : 0)];
    }
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int r = G * s + t;
        const bool live = r < a.k;
        px[s] = live ? (double)a.xyz[3LL * nb[s] + 0] : 0.0;
        py[s] = live ? (double)a.xyz[3LL * nb[s] + 1] : 0.0;
        pz[s] = live ? (double)a.xyz[3LL * nb[s] + 2] : 0.0;
        pv[s] = live ? (double)a.val[nb[s]] : 0.0;
    }
    xi = (double)a.xyz[3LL * i + 0];
    yi = (double)a.xyz[3LL * i + 1];
    zi = (double)a.xyz[3LL * i + 2];
}

template <int NMAX, int G, int MODEL>
__global__ __launch_bounds__(256) void k_loo_forward(LooArgs a) {
    constexpr int S = NMAX / G;
    typedef Chol<NMAX, G, double> CH;
    typedef LooState<NMAX, G> ST;
    const int t = threadIdx.x & (G - 1);
    const int i = (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) / G);
    const int p = blockIdx.y;
    if (i >= a.n) return;  // whole groups leave together
    const long long nsys = (long long)a.np * a.n, sys = (long long)p * a.n + i;
    const LooParams q = load_params(a.prm, p);
    double px[S], py[S], pz[S], pv[S], xi, yi, zi;
    load_system<NMAX, G>(a, i, t, px, py, pz, pv, xi, yi, zi);
    double L[CH::TOT], D[S];
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
        const double cx = gbc<G>(px[c / G], c % G), cy = gbc<G>(py[c / G], c % G), cz = gbc<G>(pz[c / G], c % G);
#pragma unroll
        for (int s = c / G; s < S; ++s) {
            const int r = G * s + t;
            const double v = (r == c) ? q.sill : pair_cov<MODEL>(q, px[s] - cx, py[s] - cy, pz[s] - cz);
            L[slot_off<G>(s) + c] = r < a.k ? -v : ((r == c) ? -1.0 : 0.0);  // Chol takes -A
        }
    }
    double bq[S];
#pragma unroll
    for (int s = 0; s < S; ++s)
        bq[s] = (G * s + t < a.k) ? pair_cov<MODEL>(q, px[s] - xi, py[s] - yi, pz[s] - zi) : 0.0;
    double minpiv;
    const bool ok = CH::factor(L, D, minpiv, t);
    double sv[S], ov[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        sv[s] = bq[s];
        ov[s] = (G * s + t < a.k) ? 1.0 : 0.0;
    }
    CH::forward(L, D, sv, t);
    CH::backward(L, D, sv, t);
    CH::forward(L, D, ov, t);
    CH::backward(L, D, ov, t);
    double ps = 0, po = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        ps += sv[s];
        po += ov[s];
    }
    const double sum_s = gsum<G>(ps), sum_o = gsum<G>(po);
    // numerically singular systems (coincident neighbours) fail too
    const bool good = ok && minpiv > 1e-12 * q.sill && fabs(sum_o) >= 1e-15;
    const double mu = good ? (sum_s - 1.0) / sum_o : 0.0;
    double w[S], pe = 0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        // failed systems: the neighbour mean, with a zero derivative
        w[s] = good ? sv[s] - mu * ov[s] : ((G * s + t < a.k) ? 1.0 / a.k : 0.0);
        pe += w[s] * pv[s];
    }
    const double z = gsum<G>(pe);
    if (t == 0) {
        a.est[sys] = z;
        if (a.status) a.status[sys] = good ? 0 : 2;
    }
    if (a.state) {
        double* st = a.state + sys * G + t;
        const long long es = nsys * G;  // stride between elements
#pragma unroll
        for (int e = 0; e < CH::TOT; ++e) st[e * es] = L[e];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            st[(CH::TOT + s) * es] = D[s];
            st[(CH::TOT + S + s) * es] = w[s];
            st[(CH::TOT + 2 * S + s) * es] = good ? ov[s] : 0.0;
        }
        st[(ST::N - 1) * es] = good ? sum_o : 0.0;
    }
}

// Backward: per system the Jacobian row dz/dtheta (optionally written), and
// grad_est * J reduced per block into part[(p * gridDim.x + bx) * 5 + c].
template <int NMAX, int G, int MODEL>
__global__ __launch_bounds__(256) void k_loo_backward(LooArgs a) {
    constexpr int S = NMAX / G;
    typedef Chol<NMAX, G, double> CH;
    typedef LooState<NMAX, G> ST;
    __shared__ double red[256 / G][5];
    const int t = threadIdx.x & (G - 1);
    const int gi = threadIdx.x / G;
    const int i = (int)((blockIdx.x * (long long)blockDim.x + threadIdx.x) / G);
    const int p = blockIdx.y;
    double jac[5] = {0, 0, 0, 0, 0};
    bool sys_ok = false;
    long long sys = 0;
    if (i < a.n) {
        const long long nsys = (long long)a.np * a.n;
        sys = (long long)p * a.n + i;
        const LooParams q = load_params(a.prm, p);
        double px[S], py[S], pz[S], pv[S], xi, yi, zi;
        load_system<NMAX, G>(a, i, t, px, py, pz, pv, xi, yi, zi);
        const double* st = a.state + sys * G + t;
        const long long es = nsys * G;  // stride between elements
        double L[CH::TOT], D[S], w[S], ov[S];
#pragma unroll
        for (int e = 0; e < CH::TOT; ++e) L[e] = st[e * es];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            D[s] = st[(CH::TOT + s) * es];
            w[s] = st[(CH::TOT + S + s) * es];
            ov[s] = st[(CH::TOT + 2 * S + s) * es];
        }
        const double sum_o = st[(ST::N - 1) * es - t];  // lane 0's copy
        sys_ok = sum_o != 0.0;
        if (sys_ok) {
            // adjoint: A s' = v, lambda_mu = sum(s') / sum(o), lambda = s' - lambda_mu o
            double lam[S];
#pragma unroll
            for (int s = 0; s < S; ++s) lam[s] = pv[s];
            CH::forward(L, D, lam, t);
            CH::backward(L, D, lam, t);
            double pl = 0;
#pragma unroll
            for (int s = 0; s < S; ++s) pl += lam[s];
            const double lmu = gsum<G>(pl) / sum_o;
#pragma unroll
            for (int s = 0; s < S; ++s) lam[s] -= lmu * ov[s];
            // + sum_a lambda_a db_a
#pragma unroll
            for (int s = 0; s < S; ++s)
                if (G * s + t < a.k) pair_grad<MODEL>(q, px[s] - xi, py[s] - yi, pz[s] - zi, lam[s], jac);
            // - sum_{r,c} lambda_r w_c dA_rc (full matrix; diagonal dA = dsill)
#pragma unroll
            for (int c = 0; c < NMAX; ++c) {
                const double cx = gbc<G>(px[c / G], c % G), cy = gbc<G>(py[c / G], c % G),
                             cz = gbc<G>(pz[c / G], c % G), wc = gbc<G>(w[c / G], c % G);
                if (c >= a.k) continue;  // uniform
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const int r = G * s + t;
                    if (r >= a.k) continue;
                    const double f = -lam[s] * wc;
                    if (r == c)
                        jac[0] += f;
                    else
                        pair_grad<MODEL>(q, px[s] - cx, py[s] - cy, pz[s] - cz, f, jac);
                }
            }
        }
#pragma unroll
        for (int c = 0; c < 5; ++c) jac[c] = gsum<G>(jac[c]);
        if (t == 0 && a.jac) {
#pragma unroll
            for (int c = 0; c < 5; ++c) a.jac[sys * 5 + c] = jac[c];
        }
    }
    // grad_est-weighted block sums, in a fixed order
    if (t == 0) {
        const double g = (i < a.n && a.gest) ? a.gest[sys] : 0.0;
#pragma unroll
        for (int c = 0; c < 5; ++c) red[gi][c] = g * jac[c];
    }
    __syncthreads();
    if (threadIdx.x < 5) {
        double acc = 0;
        for (int k = 0; k < 256 / G; ++k) acc += red[k][threadIdx.x];
        a.part[((long long)p * gridDim.x + blockIdx.x) * 5 + threadIdx.x] = acc;
    }
}

// Neighbour-count bucket and lanes per system (8 for 32 neighbours keeps the
// unrolled bodies small).
inline int nmax_for(int k) { return k <= 8 ? 8 : (k <= 16 ? 16 : (k <= 24 ? 24 : 32)); }
constexpr int loo_g(int nmax) { return nmax <= 24 ? 4 : 8; }

template <int NMAX, int MODEL>
void launch_loo_m(const LooArgs& a, bool backward, hipStream_t st) {
    constexpr int G = loo_g(NMAX);
    const dim3 grid((unsigned)((a.n * (long long)G + 255) / 256), (unsigned)a.np);
    if (!backward)
        hipLaunchKernelGGL((k_loo_forward<NMAX, G, MODEL>), grid, dim3(256), 0, st, a);
    else
        hipLaunchKernelGGL((k_loo_backward<NMAX, G, MODEL>), grid, dim3(256), 0, st, a);
}

// One per covariance model: loo_m<MODEL>.hip.
void launch_loo_model_0(const LooArgs& a, bool backward, hipStream_t st);
void launch_loo_model_1(const LooArgs& a, bool backward, hipStream_t st);
void launch_loo_model_2(const LooArgs& a, bool backward, hipStream_t st);

}  // namespace hpglx

// loo.hip -- host side of the leave-one-out kriging (kernels in
// loo_kernels.h, compiled per covariance model in loo_m0/1/2.hip) and the
// k-nearest-neighbour kernel.
#include <hip/hip_runtime.h>

#include <cstdint>
#include <string>

#include "loo_kernels.h"

namespace hpglx {

__global__ void k_loo_reduce(int np, int nblk, const double* part, double* grad) {
    const int p = blockIdx.x, c = threadIdx.x;
    if (p >= np || c >= 5) return;
    double acc = 0;
    for (int b = 0; b < nblk; ++b) acc += part[((long long)p * nblk + b) * 5 + c];
    grad[p * 5 + c] = acc;
}

// k nearest other points by Euclidean distance (ties: lower index first),
// one lane per query, candidates staged through LDS.
template <int KMAX>
__global__ __launch_bounds__(256) void k_knn(const float* xyz, int n, int k, int* nbr) {
    __shared__ float tile[256 * 3];
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    const bool q = i < n;
    const double xi = q ? (double)xyz[3LL * i] : 0.0, yi = q ? (double)xyz[3LL * i + 1] : 0.0,
                 zi = q ? (double)xyz[3LL * i + 2] : 0.0;
    double bd[KMAX];
    int bi[KMAX];
#pragma unroll
    for (int m = 0; m < KMAX; ++m) {
        bd[m] = 1e300;
        bi[m] = 0x7fffffff;
    }
    double worst = 1e300;  // entry k - 1 of the list
    int wi = 0x7fffffff;
    for (int j0 = 0; j0 < n; j0 += 256) {
        const int jl = j0 + threadIdx.x;
        __syncthreads();
        if (jl < n) {
            tile[3 * threadIdx.x + 0] = xyz[3LL * jl + 0];
            tile[3 * threadIdx.x + 1] = xyz[3LL * jl + 1];
            tile[3 * threadIdx.x + 2] = xyz[3LL * jl + 2];
        }
        __syncthreads();
        const int jn = min(256, n - j0);
        for (int u = 0; u < jn; ++u) {
            const int j = j0 + u;
            const double dx = (double)tile[3 * u] - xi, dy = (double)tile[3 * u + 1] - yi,
                         dz = (double)tile[3 * u + 2] - zi;
            const double d = dx * dx + dy * dy + dz * dz;
            if (!q || j == i || !(d < worst || (d == worst && j < wi))) continue;
            // insert (d, j) into the first k entries, ordered by (distance, index)
            bool placed = false;
#pragma unroll
            for (int m = KMAX - 1; m >= 1; --m) {
                if (m > k - 1 || placed) continue;
                if (d < bd[m - 1] || (d == bd[m - 1] && j < bi[m - 1])) {
                    bd[m] = bd[m - 1];
                    bi[m] = bi[m - 1];
                } else {
                    bd[m] = d;
                    bi[m] = j;
                    placed = true;
                }
            }
            if (!placed) {
                bd[0] = d;
                bi[0] = j;
            }
#pragma unroll
            for (int m = 0; m < KMAX; ++m)
                if (m == k - 1) {
                    worst = bd[m];
                    wi = bi[m];
                }
        }
    }
    if (q) {
#pragma unroll
        for (int m = 0; m < KMAX; ++m)
            if (m < k) nbr[(long long)i * k + m] = bi[m];
    }
}


namespace {

void launch_loo(const LooArgs& a, bool backward, hipStream_t st) {
    if (a.model == 0)
        launch_loo_model_0(a, backward, st);
    else if (a.model == 1)
        launch_loo_model_1(a, backward, st);
    else
        launch_loo_model_2(a, backward, st);
}

}  // namespace

size_t loo_state_doubles(int n, int k, int np) {
    size_t per = 0;
    switch (nmax_for(k)) {
        case 8: per = (size_t)LooState<8, loo_g(8)>::N * loo_g(8); break;
        case 16: per = (size_t)LooState<16, loo_g(16)>::N * loo_g(16); break;
        case 24: per = (size_t)LooState<24, loo_g(24)>::N * loo_g(24); break;
        default: per = (size_t)LooState<32, loo_g(32)>::N * loo_g(32); break;
    }
    return per * (size_t)n * (size_t)np;
}

size_t loo_part_doubles(int n, int k, int np) {
    const size_t g = (size_t)loo_g(nmax_for(k));
    const size_t nblk = ((size_t)n * g + 255) / 256;
    return nblk * (size_t)np * 5;
}

int loo_forward(const LooArgs& a, hipStream_t st, std::string& err) {
    if (a.n < 2 || a.k < 1 || a.k > 32 || a.k > a.n - 1 || a.np < 1 || a.np > 65535 || a.model < 0 || a.model > 2) {
        err = "loo: need n >= 2, 1 <= k <= min(32, n - 1), 1 <= parameter sets <= 65535, model 0..2";
        return -1;
    }
    launch_loo(a, false, st);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("loo forward launch: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

int loo_backward(const LooArgs& a, double* grad, hipStream_t st, std::string& err) {
    if (!a.state || !a.part) {
        err = "loo backward: the forward state and a partial-sum buffer are required";
        return -1;
    }
    launch_loo(a, true, st);
    const int nblk = (int)((a.n * (long long)loo_g(nmax_for(a.k)) + 255) / 256);
    if (grad) hipLaunchKernelGGL(k_loo_reduce, dim3(a.np), dim3(64), 0, st, a.np, nblk, (const double*)a.part, grad);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("loo backward launch: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

int knn(const float* xyz, int n, int k, int* nbr, hipStream_t st, std::string& err) {
    if (n < 2 || k < 1 || k > 32 || k > n - 1) {
        err = "knn: need n >= 2 and 1 <= k <= min(32, n - 1)";
        return -1;
    }
    const dim3 grid((unsigned)((n + 255) / 256));
    if (k <= 8)
        hipLaunchKernelGGL(k_knn<8>, grid, dim3(256), 0, st, xyz, n, k, nbr);
    else if (k <= 16)
        hipLaunchKernelGGL(k_knn<16>, grid, dim3(256), 0, st, xyz, n, k, nbr);
    else if (k <= 24)
        hipLaunchKernelGGL(k_knn<24>, grid, dim3(256), 0, st, xyz, n, k, nbr);
    else
        hipLaunchKernelGGL(k_knn<32>, grid, dim3(256), 0, st, xyz, n, k, nbr);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("knn launch: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

}  // namespace hpglx

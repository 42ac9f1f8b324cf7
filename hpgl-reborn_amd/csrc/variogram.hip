// variogram.hip -- the experimental variogram on the GPU
// (_cvariogram/variograms.cpp:353-506).
//
// Grid form (calc_variograms): for every offset of the search window that
// lies in the tunnel, the squared differences of all informed cell pairs at
// that offset go to the lags the offset's length falls in.  The informed
// cells are compacted in grid order; one block per (offset, 2048 informed
// cells) sums its pairs in a fixed order, the chunks of an offset are summed
// in order, and one lane per lag adds the offsets in the reference's window
// order.  Point-set form (calc_variograms_from_point_set): every ordered
// pair of points; a block holds 128 first points, streams a 2048-point range
// of second points through LDS and keeps per-lane per-lag sums in LDS;
// blocks are summed in order.  Products and sums use the _rn intrinsics so that nothing is fused:
// each squared difference and each distance rounds as the reference's does.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <cstdint>
#include <string>
#include <vector>

#include "krige_launch.h"
#include "variogram_plan.h"

namespace hpglx {

namespace {

constexpr int VB = 256;          // threads per block
constexpr int VCELLS = 16;       // cells per thread in the compaction
constexpr int VCHUNK = 2048;     // informed cells per block in the pair kernel
constexpr int PB = 128;          // first points per block in the point-set kernel
constexpr int PLAGS = 64;        // lags per pass of the point-set kernel
constexpr int PJ = 2048;         // second points per block of the point-set kernel

__global__ __launch_bounds__(VB) void k_vario_count(const unsigned char* mask, long long n, int* block_cnt) {
    __shared__ int part[VB];
    const long long base = (long long)blockIdx.x * VB * VCELLS + (long long)threadIdx.x * VCELLS;
    int c = 0;
    for (int k = 0; k < VCELLS; ++k) c += (base + k < n && mask[base + k]) ? 1 : 0;
    part[threadIdx.x] = c;
    __syncthreads();
    for (int d = VB / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) part[threadIdx.x] += part[threadIdx.x + d];
        __syncthreads();
    }
    if (threadIdx.x == 0) block_cnt[blockIdx.x] = part[0];
}

// Exclusive prefix of the block counts (one block); total at block_cnt[nb].
__global__ __launch_bounds__(1024) void k_vario_scan(int* block_cnt, int nb) {
    __shared__ long long part[1024];
    const int t = threadIdx.x, per = (nb + 1023) / 1024;
    const int a = min(nb, t * per), b = min(nb, a + per);
    long long acc = 0;
    for (int c = a; c < b; ++c) acc += block_cnt[c];
    part[t] = acc;
    __syncthreads();
    if (t == 0) {
        long long run = 0;
        for (int i = 0; i < 1024; ++i) {
            const long long x = part[i];
            part[i] = run;
            run += x;
        }
        block_cnt[nb] = (int)run;
    }
    __syncthreads();
    long long run = part[t];
    for (int c = a; c < b; ++c) {
        const int x = block_cnt[c];
        block_cnt[c] = (int)run;
        run += x;
    }
}

__global__ __launch_bounds__(VB) void k_vario_scatter(const unsigned char* mask, long long n, const int* block_off,
                                                      int* list) {
    __shared__ int part[VB];
    const long long base = (long long)blockIdx.x * VB * VCELLS + (long long)threadIdx.x * VCELLS;
    int c = 0;
    for (int k = 0; k < VCELLS; ++k) c += (base + k < n && mask[base + k]) ? 1 : 0;
    part[threadIdx.x] = c;
    __syncthreads();
    for (int d = 1; d < VB; d <<= 1) {
        const int o = (int)threadIdx.x >= d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += o;
        __syncthreads();
    }
    int pos = block_off[blockIdx.x] + part[threadIdx.x] - c;
    for (int k = 0; k < VCELLS; ++k)
        if (base + k < n && mask[base + k]) list[pos++] = (int)(base + k);
}

struct PairArgs {
    const float* data;
    const unsigned char* mask;
    const int* list;
    int ninf;
    int nx, ny, nz;
    const int4* offs;  // dx, dy, dz, flat delta
    int percent;
    int o_base;        // index of this launch's first offset (for the sampling hash)
    double* psum;      // [offset][chunk]
    int* pcnt;
    int nchunks;
};

__device__ inline uint32_t mix32(uint32_t a, uint32_t b) {
    uint32_t h = a * 0x9E3779B1u ^ (b + 0x7F4A7C15u + (a << 6) + (a >> 2));
    h ^= h >> 16;
    h *= 0x85EBCA6Bu;
    h ^= h >> 13;
    h *= 0xC2B2AE35u;
    h ^= h >> 16;
    return h;
}

__global__ __launch_bounds__(VB) void k_vario_pairs(PairArgs a) {
    __shared__ double ssum[VB];
    __shared__ int scnt[VB];
    const int o = blockIdx.y;
    const int4 off = a.offs[o];
    const int first = blockIdx.x * VCHUNK;
    const int last = min(a.ninf, first + VCHUNK);
    const int nxy = a.nx * a.ny;
    double sum = 0.0;
    int cnt = 0;
    for (int i = first + threadIdx.x; i < last; i += VB) {
        const int idx = a.list[i];
        const int z = idx / nxy, r = idx - z * nxy, y = r / a.nx, x = r - y * a.nx;
        const int x2 = x + off.x, y2 = y + off.y, z2 = z + off.z;
        if (x2 < 0 || y2 < 0 || z2 < 0 || x2 >= a.nx || y2 >= a.ny || z2 >= a.nz) continue;
        const int idx2 = idx + off.w;
        if (!a.mask[idx2]) continue;
        if (a.percent < 100 && (int)(mix32((uint32_t)(a.o_base + o), (uint32_t)idx) % 100u) >= a.percent) continue;
        const double d = __dsub_rn((double)a.data[idx], (double)a.data[idx2]);
        sum = __dadd_rn(sum, __dmul_rn(d, d));
        ++cnt;
    }
    ssum[threadIdx.x] = sum;
    scnt[threadIdx.x] = cnt;
    __syncthreads();
    for (int d = VB / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            ssum[threadIdx.x] = __dadd_rn(ssum[threadIdx.x], ssum[threadIdx.x + d]);
            scnt[threadIdx.x] += scnt[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        a.psum[(long long)o * a.nchunks + blockIdx.x] = ssum[0];
        a.pcnt[(long long)o * a.nchunks + blockIdx.x] = scnt[0];
    }
}

// One lane per offset: its chunk sums in chunk order.
__global__ void k_vario_offsets(const double* psum, const int* pcnt, int nchunks, int noff, double* osum,
                                long long* ocnt) {
    const int o = blockIdx.x * blockDim.x + threadIdx.x;
    if (o >= noff) return;
    double s = 0.0;
    long long c = 0;
    for (int k = 0; k < nchunks; ++k) {
        s = __dadd_rn(s, psum[(long long)o * nchunks + k]);
        c += pcnt[(long long)o * nchunks + k];
    }
    osum[o] = s;
    ocnt[o] = c;
}

// One block per lag: the offsets feeding it (lag_min <= l <= lag_max), each
// lane a strided share in window order, then a fixed tree over the lanes.
__global__ __launch_bounds__(VB) void k_vario_lags(const double* osum, const long long* ocnt, const int2* lag_range,
                                                   int noff, float* result) {
    __shared__ double ss[VB];
    __shared__ long long sc[VB];
    const int l = blockIdx.x;
    double s = 0.0;
    long long c = 0;
    for (int o = threadIdx.x; o < noff; o += VB) {
        const int2 lr = lag_range[o];
        if (l < lr.x || l > lr.y) continue;
        s = __dadd_rn(s, osum[o]);
        c += ocnt[o];
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int d = VB / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            ss[threadIdx.x] = __dadd_rn(ss[threadIdx.x], ss[threadIdx.x + d]);
            sc[threadIdx.x] += sc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) result[l] = sc[0] > 0 ? (float)(ss[0] / (double)sc[0] / 2) : 0.0f;
}

struct DevTemplate {
    double dir[3][3];
    double R2, R3, tol;
};

__device__ inline double dot3d(double v0, double v1, double v2, const double d[3]) {
    double s = 0.0;
    s = __dadd_rn(s, __dmul_rn(v0, d[0]));
    s = __dadd_rn(s, __dmul_rn(v1, d[1]));
    s = __dadd_rn(s, __dmul_rn(v2, d[2]));
    return s;
}

__device__ inline bool in_tunnel_dev(const DevTemplate& t, double v0, double v1, double v2) {
    const double ss1 = dot3d(v0, v1, v2, t.dir[0]);
    const double ss2 = dot3d(v0, v1, v2, t.dir[1]);
    const double ss3 = dot3d(v0, v1, v2, t.dir[2]);
    const double s2 = __ddiv_rn(ss2, t.R2), s3 = __ddiv_rn(ss3, t.R3);
    const double dist = __dsqrt_rn(__dadd_rn(__dmul_rn(s2, s2), __dmul_rn(s3, s3)));
    return dist <= 1.0 && __dmul_rn(t.tol, dist) <= ss1;
}

struct PointArgs {
    const float4* pts;  // x, y, z, value
    int n;
    DevTemplate t;
    const double* lag_start;
    const double* lag_end;
    int lag0, nl;       // lags [lag0, lag0 + nl) in this pass
    double far2;        // pairs with squared distance above this reach no lag of the pass
    double* psum;       // [block][nl]
    long long* pcnt;
};

__global__ __launch_bounds__(PB) void k_vario_points(PointArgs a) {
    extern __shared__ unsigned char smem[];
    float4* tile = reinterpret_cast<float4*>(smem);
    double* asum = reinterpret_cast<double*>(smem + PB * sizeof(float4));
    int* acnt = reinterpret_cast<int*>(asum + (size_t)a.nl * PB);
    double* lstart = reinterpret_cast<double*>(acnt + (size_t)a.nl * PB);
    double* lend = lstart + a.nl;
    const int t = threadIdx.x;
    for (int l = 0; l < a.nl; ++l) {
        asum[l * PB + t] = 0.0;
        acnt[l * PB + t] = 0;
    }
    for (int l = t; l < a.nl; l += PB) {
        lstart[l] = a.lag_start[a.lag0 + l];
        lend[l] = a.lag_end[a.lag0 + l];
    }
    const int i1 = blockIdx.x * PB + t;
    const float4 p1 = i1 < a.n ? a.pts[i1] : make_float4(0.f, 0.f, 0.f, 0.f);
    const int jend = min(a.n, (int)(blockIdx.y + 1) * PJ);
    for (int j0 = blockIdx.y * PJ; j0 < jend; j0 += PB) {
        __syncthreads();
        if (j0 + t < jend) tile[t] = a.pts[j0 + t];
        __syncthreads();
        if (i1 >= a.n) continue;
        const int jn = min(PB, jend - j0);
        for (int j = 0; j < jn; ++j) {
            const float4 p2 = tile[j];
            const double v0 = (double)(p1.x - p2.x), v1 = (double)(p1.y - p2.y), v2 = (double)(p1.z - p2.z);
            double q = 0.0;
            q = __dadd_rn(q, __dmul_rn(v0, v0));
            q = __dadd_rn(q, __dmul_rn(v1, v1));
            q = __dadd_rn(q, __dmul_rn(v2, v2));
            // exact skips: too far for any lag of this pass (sqrt is monotone,
            // far2 sits above the last lag end), or behind the tunnel's axis
            // (tol * dist >= 0 > ss1 when tol >= 0)
            if (q > a.far2) continue;
            if (a.t.tol >= 0.0 && dot3d(v0, v1, v2, a.t.dir[0]) < 0.0) continue;
            if (!in_tunnel_dev(a.t, v0, v1, v2)) continue;
            const double dist = __dsqrt_rn(q);
            const double dv = (double)(p1.w - p2.w);
            const double var = __dmul_rn(dv, dv);
            for (int l = 0; l < a.nl; ++l)
                if (lstart[l] <= dist && dist < lend[l]) {
                    asum[l * PB + t] = __dadd_rn(asum[l * PB + t], var);
                    acnt[l * PB + t] += 1;
                }
        }
    }
    __syncthreads();
    // per lag, the block's lanes in order
    for (int l = t; l < a.nl; l += PB) {
        double s = 0.0;
        long long c = 0;
        for (int k = 0; k < PB; ++k) {
            s = __dadd_rn(s, asum[l * PB + k]);
            c += acnt[l * PB + k];
        }
        const long long b = (long long)blockIdx.y * gridDim.x + blockIdx.x;
        a.psum[b * a.nl + l] = s;
        a.pcnt[b * a.nl + l] = c;
    }
}

// One block per lag: the pair blocks' partial sums, each lane a strided
// share in block order, then a fixed tree over the lanes.
__global__ __launch_bounds__(VB) void k_vario_points_lags(const double* psum, const long long* pcnt, int nblocks,
                                                          int nl, float* result) {
    __shared__ double ss[VB];
    __shared__ long long sc[VB];
    const int l = blockIdx.x;
    double s = 0.0;
    long long c = 0;
    for (int b = threadIdx.x; b < nblocks; b += VB) {
        s = __dadd_rn(s, psum[(long long)b * nl + l]);
        c += pcnt[(long long)b * nl + l];
    }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int d = VB / 2; d > 0; d >>= 1) {
        if ((int)threadIdx.x < d) {
            ss[threadIdx.x] = __dadd_rn(ss[threadIdx.x], ss[threadIdx.x + d]);
            sc[threadIdx.x] += sc[threadIdx.x + d];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) result[l] = sc[0] > 0 ? (float)(ss[0] / (double)sc[0] / 2) : 0.0f;
}

DevTemplate dev_template(const VarioTemplate& t) {
    DevTemplate d;
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) d.dir[i][j] = t.dir[i][j];
    d.R2 = t.R[1];
    d.R3 = t.R[2];
    d.tol = t.tol_distance;
    return d;
}

}  // namespace

int vario_grid(const VarioTemplate& t, const float* d_data, const unsigned char* d_mask, const int shape[3],
               int lag_count, int percent, float* d_result, Workspace& ws, hipStream_t st, std::string& err) {
    const long long n = (long long)shape[0] * shape[1] * shape[2];
    if (n <= 0 || n > 2147483647LL) {
        err = "calc_variograms: grid volume must be 1..INT_MAX";
        return -1;
    }
    const std::vector<VarioOffset> offs = vario_grid_offsets(t, lag_count);
    const int nb = (int)((n + VB * VCELLS - 1) / (VB * VCELLS));
    auto* bc = (int*)ws.get("vario_bcnt", (size_t)(nb + 1) * 4, err);
    auto* list = (int*)ws.get("vario_list", (size_t)n * 4, err);
    if (!err.empty()) return -1;
    hipLaunchKernelGGL(k_vario_count, dim3(nb), dim3(VB), 0, st, d_mask, n, bc);
    hipLaunchKernelGGL(k_vario_scan, dim3(1), dim3(1024), 0, st, bc, nb);
    hipLaunchKernelGGL(k_vario_scatter, dim3(nb), dim3(VB), 0, st, d_mask, n, (const int*)bc, list);
    int ninf = 0;
    if (hipMemcpyAsync(&ninf, bc + nb, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "variogram compaction failed";
        return -1;
    }
    const int noff = (int)offs.size();
    const int nchunks = std::max(1, (ninf + VCHUNK - 1) / VCHUNK);
    std::vector<int4> h_off(std::max(1, noff));
    std::vector<int2> h_lr(std::max(1, noff));
    for (int o = 0; o < noff; ++o) {
        const VarioOffset& v = offs[o];
        h_off[o] = make_int4(v.d[0], v.d[1], v.d[2], v.d[0] + shape[0] * (v.d[1] + shape[1] * v.d[2]));
        h_lr[o] = make_int2(v.lag_min, v.lag_max);
    }
    auto* d_off = (int4*)ws.get("vario_off", h_off.size() * sizeof(int4), err);
    auto* d_lr = (int2*)ws.get("vario_lr", h_lr.size() * sizeof(int2), err);
    auto* psum = (double*)ws.get("vario_psum", (size_t)std::max(1, noff) * nchunks * 8, err);
    auto* pcnt = (int*)ws.get("vario_pcnt", (size_t)std::max(1, noff) * nchunks * 4, err);
    if (!err.empty()) return -1;
    if (hipMemcpyAsync(d_off, h_off.data(), h_off.size() * sizeof(int4), hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_lr, h_lr.data(), h_lr.size() * sizeof(int2), hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "host to device copy failed";
        return -1;
    }
    if (noff > 0 && ninf > 0) {
        PairArgs a;
        a.data = d_data;
        a.mask = d_mask;
        a.list = list;
        a.ninf = ninf;
        a.nx = shape[0];
        a.ny = shape[1];
        a.nz = shape[2];
        a.offs = d_off;
        a.percent = percent;
        a.psum = psum;
        a.pcnt = pcnt;
        a.nchunks = nchunks;
        for (int o0 = 0; o0 < noff; o0 += 65535) {
            PairArgs b = a;
            b.o_base = o0;
            b.offs = d_off + o0;
            b.psum = psum + (long long)o0 * nchunks;
            b.pcnt = pcnt + (long long)o0 * nchunks;
            hipLaunchKernelGGL(k_vario_pairs, dim3(nchunks, std::min(65535, noff - o0)), dim3(VB), 0, st, b);
        }
    }
    auto* osum = (double*)ws.get("vario_osum", (size_t)std::max(1, noff) * 8, err);
    auto* ocnt = (long long*)ws.get("vario_ocnt", (size_t)std::max(1, noff) * 8, err);
    if (!err.empty()) return -1;
    const int nused = ninf > 0 ? noff : 0;
    if (nused > 0)
        hipLaunchKernelGGL(k_vario_offsets, dim3((nused + VB - 1) / VB), dim3(VB), 0, st, (const double*)psum,
                           (const int*)pcnt, nchunks, nused, osum, ocnt);
    hipLaunchKernelGGL(k_vario_lags, dim3(lag_count), dim3(VB), 0, st, (const double*)osum, (const long long*)ocnt,
                       (const int2*)d_lr, nused, d_result);
    if (hipGetLastError() != hipSuccess) {
        err = "variogram kernels failed to launch";
        return -1;
    }
    return 0;
}

int vario_points(const VarioTemplate& t, const float* d_pts4, int n, int lag_count, float* d_result, Workspace& ws,
                 hipStream_t st, std::string& err) {
    if (n < 0) {
        err = "calc_variograms_from_point_set: negative point count";
        return -1;
    }
    std::vector<double> ls, le;
    vario_lag_bounds(t, lag_count, ls, le);
    const int nbx = std::max(1, (n + PB - 1) / PB), nby = std::max(1, (n + PJ - 1) / PJ);
    const int nblocks = nbx * nby;
    auto* d_ls = (double*)ws.get("vario_ls", (size_t)std::max(1, lag_count) * 8, err);
    auto* d_le = (double*)ws.get("vario_le", (size_t)std::max(1, lag_count) * 8, err);
    auto* psum = (double*)ws.get("vario_ppsum", (size_t)nblocks * PLAGS * 8, err);
    auto* pcnt = (long long*)ws.get("vario_ppcnt", (size_t)nblocks * PLAGS * 8, err);
    if (!err.empty()) return -1;
    if (lag_count > 0 &&
        (hipMemcpyAsync(d_ls, ls.data(), (size_t)lag_count * 8, hipMemcpyHostToDevice, st) != hipSuccess ||
         hipMemcpyAsync(d_le, le.data(), (size_t)lag_count * 8, hipMemcpyHostToDevice, st) != hipSuccess)) {
        err = "host to device copy failed";
        return -1;
    }
    for (int l0 = 0; l0 < lag_count; l0 += PLAGS) {
        const int nl = std::min(PLAGS, lag_count - l0);
        if (n > 0) {
            PointArgs a;
            a.pts = reinterpret_cast<const float4*>(d_pts4);
            a.n = n;
            a.t = dev_template(t);
            a.lag_start = d_ls;
            a.lag_end = d_le;
            a.lag0 = l0;
            a.nl = nl;
            double top = 0.0;
            for (int l = l0; l < l0 + nl; ++l) top = std::max(top, le[l]);
            a.far2 = top > 0.0 ? (top * (1.0 + 1e-9)) * (top * (1.0 + 1e-9)) : -1.0;
            a.psum = psum;
            a.pcnt = pcnt;
            const size_t lds = PB * sizeof(float4) + (size_t)nl * PB * (8 + 4) + (size_t)nl * 16;
            if (lds > 65536 &&
                hipFuncSetAttribute((const void*)&k_vario_points, hipFuncAttributeMaxDynamicSharedMemorySize,
                                    (int)lds) != hipSuccess) {
                err = "hipFuncSetAttribute failed";
                return -1;
            }
            hipLaunchKernelGGL(k_vario_points, dim3(nbx, nby), dim3(PB), lds, st, a);
        }
        hipLaunchKernelGGL(k_vario_points_lags, dim3(nl), dim3(VB), 0, st, (const double*)psum,
                           (const long long*)pcnt, n > 0 ? nblocks : 0, nl, d_result + l0);
    }
    if (hipGetLastError() != hipSuccess) {
        err = "variogram kernels failed to launch";
        return -1;
    }
    return 0;
}

}  // namespace hpglx

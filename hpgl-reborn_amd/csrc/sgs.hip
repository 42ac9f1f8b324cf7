// sgs.hip -- sequential Gaussian simulation, level by level (see sgs.h).
//
// Reference: src/geo_bsd/hpgl/sequential_simulation.h:51-119 (the walk),
// kriging_interpolation.h:269-313 (search, weights, combine with variance),
// sequential_gaussian_simulation.cpp:18-78 (SK with mean 0 or OK, both over
// the precalculated covariance table of the radius box), sample.h:14-25 and
// gaussian_distribution.h:56-89 (the draw).
#include "sgs.h"

#include <memory>
#include <thread>

#include <algorithm>
#include <cmath>
#include <cstring>
#include <vector>

#include "krige_kernels.h"
#include "solve_dispatch.h"

namespace hpglx {
namespace {

// Padded 0/1 mask (halo of h cells: the template's reach; x pitch px) from
// the cell mask.
__global__ void k_sgs_pmask(GridDims g, const unsigned char* mask, unsigned char* pmask) {
    const long long P = (long long)g.px * g.py * g.pz;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < P;
         i += (long long)gridDim.x * blockDim.x) {
        const int x = (int)(i % g.px) - g.hx, y = (int)((i / g.px) % g.py) - g.hy, z = (int)(i / ((long long)g.px * g.py)) - g.hz;
        const bool in = x >= 0 && x < g.nx && y >= 0 && y < g.ny && z >= 0 && z < g.nz;
        pmask[i] = in ? (mask[((long long)z * g.ny + y) * g.nx + x] != 0) : 0;
    }
}

// Template entries as deltas in the padded mask and as packed offsets.
__global__ void k_sgs_tdelta(GridDims g, int T, const int* tdx, const int* tdy, const int* tdz, int* tdelta,
                             int* tpack) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= T) return;
    tdelta[t] = (tdz[t] * g.py + tdy[t]) * g.px + tdx[t];
    tpack[t] = pack_off(tdx[t], tdy[t], tdz[t]);
}

unsigned blocks_of(long long work, int per) {
    return (unsigned)std::max(1LL, (work + per - 1) / per);
}

}  // namespace

int sgs_run(DevPlan& plan, Workspace& ws, int kind, const int dims[3], float* d_values, unsigned char* d_mask,
            const float* d_lvm, const std::vector<SgsSchedule>& sched, hipStream_t st, std::string& err) {
    const PlanHost& h = plan.h;
    if (kind != KIND_OK && kind != KIND_SK) {
        err = "sequential simulation takes simple or ordinary kriging";
        return -1;
    }
    if (d_lvm && kind != KIND_SK) {
        err = "a locally varying mean applies to simple kriging only";
        return -1;
    }
    if (h.kmax > 64) {
        err = "max_neighbours above 64 is not supported by the GPU simulation";
        return -1;
    }
    if (!plan.upload(err)) return -1;
    GridDims g;
    g.nx = dims[0];
    g.ny = dims[1];
    g.nz = dims[2];
    g.n = (long long)g.nx * g.ny * g.nz;
    g.rx = h.r[0];
    g.ry = h.r[1];
    g.rz = h.r[2];
    g.hx = h.H[0];
    g.hy = h.H[1];
    g.hz = h.H[2];
    g.px = g.nx + 2 * g.hx;
    g.py = g.ny + 2 * g.hy;
    g.pz = g.nz + 2 * g.hz;
    g.cx = g.cy = g.cz = 0;
    const long long n = g.n, P = (long long)g.px * g.py * g.pz;
    const int K = h.kmax, T = plan.tsize;
    const int R = (int)sched.size();
    // Step l runs level l of every realization.
    int steps = 0;
    long long m = 0;
    for (const SgsSchedule& s : sched) {
        steps = std::max(steps, (int)s.level_off.size() - 1);
        m += (long long)s.cells.size();
    }
    if (m == 0) return 0;
    // Entries of step l: level l of realization 0, then of 1, ...  Sizes and
    // offsets first, then the copies on up to 16 threads over disjoint step
    // ranges (the buffers are left uninitialised: each thread touches its own).
    std::vector<int> step_off((size_t)steps + 1, 0);
    for (int l = 0; l < steps; ++l) {
        long long c = 0;
        for (const SgsSchedule& s : sched)
            if (l + 1 < (int)s.level_off.size()) c += s.level_off[(size_t)l + 1] - s.level_off[(size_t)l];
        step_off[(size_t)l + 1] = step_off[(size_t)l] + (int)c;
    }
    std::unique_ptr<int[]> ecell(new int[(size_t)m]), ereal(new int[(size_t)m]);
    std::unique_ptr<double[]> ez(new double[(size_t)m]);
    auto fill = [&](int l0, int l1) {
        for (int l = l0; l < l1; ++l) {
            size_t e = (size_t)step_off[(size_t)l];
            for (int r = 0; r < R; ++r) {
                const SgsSchedule& s = sched[(size_t)r];
                if (l + 1 >= (int)s.level_off.size()) continue;
                for (int i = s.level_off[(size_t)l]; i < s.level_off[(size_t)l + 1]; ++i, ++e) {
                    ecell[e] = s.cells[(size_t)i];
                    ereal[e] = r;
                    ez[e] = s.z[(size_t)i];
                }
            }
        }
    };
    const int nthr = (int)std::min<long long>(16, std::max<long long>(1, m / 200000));
    if (nthr == 1) {
        fill(0, steps);
    } else {
        // step ranges of about equal entry counts
        std::vector<std::thread> pool;
        int l0 = 0;
        for (int t = 0; t < nthr; ++t) {
            const long long target = m * (t + 1) / nthr;
            int l1 = l0;
            while (l1 < steps && (t == nthr - 1 || step_off[(size_t)l1] < target)) ++l1;
            pool.emplace_back(fill, l0, l1);
            l0 = l1;
        }
        for (auto& th : pool) th.join();
    }
    auto* pmask = (unsigned char*)ws.get("sgs_pmask", (size_t)(R * P), err);
    auto* nboff = (int*)ws.get("sgs_nboff", (size_t)R * n * K * 4, err);
    auto* est = (float*)ws.get("sgs_est", (size_t)R * n * 4, err);
    auto* var = (double*)ws.get("sgs_var", (size_t)R * n * 8, err);
    auto* status = (unsigned char*)ws.get("sgs_status", (size_t)R * n, err);
    auto* tdelta = (int*)ws.get("sgs_tdelta", (size_t)std::max(T, 1) * 4, err);
    auto* tpack = (int*)ws.get("sgs_tpack", (size_t)std::max(T, 1) * 4, err);
    auto* cells = (int*)ws.get("sgs_cells", (size_t)m * 4, err);
    auto* reals = (int*)ws.get("sgs_reals", (size_t)m * 4, err);
    auto* zdev = (double*)ws.get("sgs_z", (size_t)m * 8, err);
    if (!err.empty()) return -1;
    if (hipMemcpyAsync(cells, ecell.get(), (size_t)m * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(reals, ereal.get(), (size_t)m * 4, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(zdev, ez.get(), (size_t)m * 8, hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "host to device copy failed";
        (void)hipStreamSynchronize(st);
        return -1;
    }
    for (int r = 0; r < R; ++r)
        hipLaunchKernelGGL(k_sgs_pmask, dim3((unsigned)std::min<long long>(4096, blocks_of(P, 256))), dim3(256), 0, st,
                           g, (const unsigned char*)(d_mask + (long long)r * n), pmask + (long long)r * P);
    if (T > 0)
        hipLaunchKernelGGL(k_sgs_tdelta, dim3(blocks_of(T, 256)), dim3(256), 0, st, g, T, (const int*)plan.tdx,
                           (const int*)plan.tdy, (const int*)plan.tdz, tdelta, tpack);

    SolveArgs so;
    std::memset(&so, 0, sizeof so);
    so.g = g;
    so.kind = kind;
    so.kmax = K;
    so.values = d_values;
    so.mask = d_mask;
    so.mean = 0.0;  // SK: single_mean_t(0) (sequential_gaussian_simulation.cpp:37-43)
    so.lvm_mean = d_lvm;
    so.count = nullptr;  // the fused level kernel passes each cell's count directly
    so.nboff = nboff;
    so.ctab = plan.ctab;
    so.ctab32 = plan.ctab32;
    so.wx = h.W[0];
    so.wy = h.W[1];
    so.wz = h.W[2];
    so.otab = nullptr;
    so.otab_w = plan.otab_w;
    so.ex = h.E[0];
    so.ey = h.E[1];
    so.ez = h.E[2];
    so.c0 = h.c0();
    so.out_val = est;
    so.out_var = var;
    so.out_status = status;
    SgsLevelArgs la;
    la.T = T;
    la.tdelta = tdelta;
    la.tpack = tpack;
    la.pmask = pmask;
    la.n = n;
    la.pn = P;
    la.values = d_values;
    la.mask = d_mask;
    for (int l = 0; l < steps; ++l) {
        const int off = step_off[(size_t)l];
        la.cells = cells + off;
        la.real = reals + off;
        la.ncells = step_off[(size_t)l + 1] - off;
        la.z = zdev + off;
        if (K <= 8)
            kind == KIND_OK ? sgs_level_8_0(so, la, st) : sgs_level_8_1(so, la, st);
        else if (K <= 16)
            kind == KIND_OK ? sgs_level_16_0(so, la, st) : sgs_level_16_1(so, la, st);
        else if (K <= 24)
            kind == KIND_OK ? sgs_level_24_0(so, la, st) : sgs_level_24_1(so, la, st);
        else if (K <= 32)
            kind == KIND_OK ? sgs_level_32_0(so, la, st) : sgs_level_32_1(so, la, st);
        else if (K <= 48)
            kind == KIND_OK ? sgs_level_48_0(so, la, st) : sgs_level_48_1(so, la, st);
        else
            kind == KIND_OK ? sgs_level_64_0(so, la, st) : sgs_level_64_1(so, la, st);
    }
    if (hipGetLastError() != hipSuccess) {
        err = "sequential simulation kernel launch failed";
        (void)hipStreamSynchronize(st);  // the copies above read host arrays freed on return
        return -1;
    }
    // the host vectors above are read by the copies
    if (hipStreamSynchronize(st) != hipSuccess) {
        err = "sequential simulation failed on the device";
        return -1;
    }
    return 0;
}

}  // namespace hpglx

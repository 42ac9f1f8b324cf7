// solve_dispatch_impl.h -- launch logic of the solve kernels (solve.h),
// included by one translation unit per neighbour-count bucket
// (solve_n8.hip ... solve_n32.hip) so the buckets compile in parallel.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <algorithm>
#include <map>
#include <mutex>
#include <tuple>
#include <utility>

#include "solve.h"
#include "solve_dispatch.h"
#include "solve_wave.h"

namespace hpglx {
namespace {

// Persistent grid of a kernel: the blocks resident at once on every CU
// (cached per kernel, device and LDS bytes), raising the kernel's dynamic-LDS
// limit when needed.  0 when the kernel cannot run with that LDS.
inline int persistent_grid(const void* fn, int block, int lds) {
    static std::mutex mu;
    static std::map<std::tuple<const void*, int, int>, int> grids;
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return 0;
    std::lock_guard<std::mutex> lk(mu);
    auto key = std::make_tuple(fn, dev, lds);
    auto it = grids.find(key);
    if (it != grids.end()) return it->second;
    int cus = 0, nb = 0;
    if ((lds > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess) ||
        hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
        hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, block, lds) != hipSuccess)
        nb = 0;
    (void)hipGetLastError();
    grids[key] = nb * cus;
    return nb * cus;
}

// One launch of the bucketed solve (solve.h), 4 lanes per cell: the sizes up
// to TOP (TOP <= 24) or the 25-32 bucket (its registers allow one wave per
// SIMD), over one path's range of the work list.  Path 0 with nelem > 0 stages the octant table in LDS;
// otherwise the global tables serve.  Returns false when nothing could be
// launched (the LDS table does not fit).
template <int TOP, int KIND, bool MIXED>
bool launch_sorted_p(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    const hipEvent_t e0 = l.ev ? l.ev[0] : nullptr, e1 = l.ev ? l.ev[1] : nullptr;
    const int blocks_per_cu_cap = l.blocks_per_cu > 0 ? l.blocks_per_cu : 1 << 20;
    int cus = 0, dev = 0;
    (void)hipGetDevice(&dev);
    (void)hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev);
    if (l.nelem > 0) {
        constexpr int BLOCK = (TOP <= 24 && MIXED) ? 512 : 256;
        const void* fn = (const void*)&k_solve_sorted_lds<TOP, MIXED, KIND, BLOCK>;
        const int lds = l.nelem * (int)sizeof(double);
        int grid = persistent_grid(fn, BLOCK, lds);
        if (grid <= 0) return false;
        grid = std::min(grid, blocks_per_cu_cap * std::max(cus, 1));
        HPGLX_LAUNCH_EV_ANY(l.any_order, (k_solve_sorted_lds<TOP, MIXED, KIND, BLOCK>), dim3((unsigned)grid),
                            dim3(BLOCK), (uint32_t)lds, st, e0, e1, sa, l.nelem, l.path);
        return true;
    }
    if (sa.otab_w) {
        const void* fn = (const void*)&k_solve_sorted<TOP, MIXED, KIND, true>;
        const int grid = std::min(persistent_grid(fn, 256, 0), blocks_per_cu_cap * std::max(cus, 1));
        HPGLX_LAUNCH_EV_ANY(l.any_order, (k_solve_sorted<TOP, MIXED, KIND, true>), dim3((unsigned)std::max(grid, 1)),
                            dim3(256), 0, st, e0, e1, sa, l.path);
    } else {
        const void* fn = (const void*)&k_solve_sorted<TOP, MIXED, KIND, false>;
        const int grid = std::min(persistent_grid(fn, 256, 0), blocks_per_cu_cap * std::max(cus, 1));
        HPGLX_LAUNCH_EV_ANY(l.any_order, (k_solve_sorted<TOP, MIXED, KIND, false>),
                            dim3((unsigned)std::max(grid, 1)), dim3(256), 0, st, e0, e1, sa, l.path);
    }
    return true;
}

template <int TOP, int KIND>
bool launch_sorted(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    return l.mixed ? launch_sorted_p<TOP, KIND, true>(sa, l, st) : launch_sorted_p<TOP, KIND, false>(sa, l, st);
}

// FP64 re-solve of the mixed form's fallback list (sa.cell_list / n_list),
// sized for up to max_cells cells.
template <bool B32, int KIND>
void launch_fallback(const SolveArgs& sa, long long max_cells, hipStream_t st) {
    constexpr int G = 4;
    long long blocks = (max_cells * G + 255) / 256;
    // the list is usually empty or short: 16k groups stride over it
    blocks = std::max(1LL, std::min(blocks, 256LL));
    if (sa.otab_w)
        hipLaunchKernelGGL((k_solve_fallback<B32, KIND, true>), dim3((unsigned)blocks), dim3(256), 0, st, sa);
    else
        hipLaunchKernelGGL((k_solve_fallback<B32, KIND, false>), dim3((unsigned)blocks), dim3(256), 0, st, sa);
}

// One level of sequential Gaussian simulation (see SgsLevelArgs): one wave
// (block of 64) per cell.  The search ranks the template's hits with a
// ballot, 64 entries at a time, and keeps the first kmax in template order
// (sugarbox_neighbour_lookup.h:37-62); the first G lanes then solve the cell
// as the list solve does; lane 0 draws (sample.h:14-25,
// gaussian_distribution.h:56-89, products and sums rounded singly).  Each
// cell works in its realization's slices of the arrays.
template <int NMAX, int G, int KIND, bool OCT>
__global__ __launch_bounds__(64) void k_sgs_level(SolveArgs a, SgsLevelArgs la) {
    const int lane = threadIdx.x;
    const int cell = la.cells[blockIdx.x];
    const long long r = la.real[blockIdx.x];
    a.values += r * la.n;
    a.nboff += r * la.n * a.kmax;
    a.out_val += r * la.n;
    a.out_var += r * la.n;
    a.out_status += r * la.n;
    float* values = la.values + r * la.n;
    unsigned char* mask = la.mask + r * la.n;
    unsigned char* pmask = la.pmask + r * la.pn;
    const GridDims& g = a.g;
    const int x = cell % g.nx, y = (cell / g.nx) % g.ny, zc = cell / (g.nx * g.ny);
    const long long pcell = ((long long)(zc + g.hz) * g.py + (y + g.hy)) * g.px + (x + g.hx);
    int* nboff = const_cast<int*>(a.nboff) + (long long)cell * a.kmax;
    int cnt = 0;
    for (int t0 = 0; t0 < la.T && cnt < a.kmax; t0 += 64) {
        const int t = t0 + lane;
        const bool hit = t < la.T && pmask[pcell + la.tdelta[t]] != 0;
        const unsigned long long b = __ballot(hit);
        const int rank = cnt + __popcll(b & ((1ull << lane) - 1ull));
        if (hit && rank < a.kmax) nboff[rank] = la.tpack[t];
        cnt += __popcll(b);
    }
    if (cnt > a.kmax) cnt = a.kmax;
    __threadfence_block();
    __syncthreads();
    const unsigned ox = (unsigned)(a.wx + 1), oxy = ox * (unsigned)(a.wy + 1);
    if constexpr (NMAX > 32) {
        // above the register solvers: the whole wave, a row per lane
        __shared__ double col[64];
        wave_solve_cell<NMAX, KIND>(a, cell, cnt, lane, col);
    } else {
        if (lane < G) solve_cell<NMAX, G, false, KIND, OCT>(a, cell, cnt, lane, a.otab_w, ox, oxy);
    }
    __threadfence_block();
    __syncthreads();
    if (lane == 0) {
        const double d = la.z[blockIdx.x];
        float v;
        if (a.out_status[cell] == ST_KRIGED) {
            const double m = (double)a.out_val[cell];
            v = isnan(d) ? (float)m : (float)__dadd_rn(__dmul_rn(d, __dsqrt_rn(a.out_var[cell])), m);
        } else {
            v = isnan(d) ? 0.0f : (float)d;
        }
        values[cell] = v;
        mask[cell] = 1;
        pmask[pcell] = 1;
    }
}

template <int NMAX, int G64, int KIND>
void launch_sgs_level(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    if constexpr (KIND == 0 || KIND == 1) {  // ordinary or simple kriging (median IK is not simulated)
        // a level holds few cells, so each cell's latency counts: 8 lanes per
        // cell from 16 neighbours up, the whole wave above 32
        constexpr int GS = NMAX >= 16 ? 8 : G64;
        if (la.ncells <= 0) return;
        if (sa.otab_w)
            hipLaunchKernelGGL((k_sgs_level<NMAX, GS, KIND, true>), dim3((unsigned)la.ncells), dim3(64), 0, st, sa,
                               la);
        else
            hipLaunchKernelGGL((k_sgs_level<NMAX, GS, KIND, false>), dim3((unsigned)la.ncells), dim3(64), 0, st, sa,
                               la);
    } else {
        (void)sa;
        (void)la;
        (void)st;
    }
}

}  // namespace
}  // namespace hpglx

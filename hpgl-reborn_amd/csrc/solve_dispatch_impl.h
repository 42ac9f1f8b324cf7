// solve_dispatch_impl.h -- launch logic of the solve kernels (solve.h),
// included by one translation unit per neighbour-count bucket
// (solve_n8.hip ... solve_n32.hip) so the buckets compile in parallel.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>

#include <map>
#include <mutex>
#include <utility>

#include "solve.h"
#include "solve_dispatch.h"

namespace hpglx {
namespace {

template <int NMAX, int G, bool MIXED, int KIND>
void launch_solve(const SolveArgs& sa, long long ncells, const hipEvent_t* ev, hipStream_t st) {
    const int block = 256;
    const unsigned grid = (unsigned)((ncells * G + block - 1) / block);
    if (sa.otab_w)
        hipExtLaunchKernelGGL((k_solve<NMAX, G, MIXED, KIND, true>), dim3(grid), dim3(block), 0, st, ev[0], ev[1], 0,
                              sa);
    else
        hipExtLaunchKernelGGL((k_solve<NMAX, G, MIXED, KIND, false>), dim3(grid), dim3(block), 0, st, ev[0], ev[1],
                              0, sa);
}

// Global-table form over a device-sized cell list (cluster-path cells, and
// the FP64 re-solve of the mixed form's fallbacks).
template <int NMAX, int G, bool MIXED, int KIND>
void launch_solve_list(const SolveArgs& sa, long long max_cells, hipStream_t st) {
    const int block = 256;
    long long blocks = (max_cells * G + block - 1) / block;
    if (blocks > 2048) blocks = 2048;
    if (blocks < 1) blocks = 1;
    if (sa.otab_w)
        hipLaunchKernelGGL((k_solve<NMAX, G, MIXED, KIND, true>), dim3((unsigned)blocks), dim3(block), 0, st, sa);
    else
        hipLaunchKernelGGL((k_solve<NMAX, G, MIXED, KIND, false>), dim3((unsigned)blocks), dim3(block), 0, st, sa);
}

// Persistent LDS-table form: as many blocks as are resident at once.  Returns
// false (nothing launched) when the table does not fit a block's LDS.
template <int NMAX, int G, bool MIXED, int KIND, int BLOCK, bool PACK>
bool launch_solve_lds(const SolveArgs& sa, int nelem, const hipEvent_t* ev, hipStream_t st) {
    static std::mutex mu;
    static std::map<std::pair<int, int>, int> grids;  // (device, LDS bytes) -> grid
    const void* fn = (const void*)&k_solve_lds<NMAX, G, MIXED, KIND, BLOCK, PACK>;
    const int lds = nelem * (int)sizeof(double);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess) return false;
    int grid = 0;
    {
        std::lock_guard<std::mutex> lk(mu);
        auto it = grids.find({dev, lds});
        if (it != grids.end()) {
            grid = it->second;
        } else {
            int cus = 0, nb = 0;
            if (hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, lds) != hipSuccess ||
                hipDeviceGetAttribute(&cus, hipDeviceAttributeMultiprocessorCount, dev) != hipSuccess ||
                hipOccupancyMaxActiveBlocksPerMultiprocessor(&nb, fn, BLOCK, lds) != hipSuccess)
                nb = 0;
            (void)hipGetLastError();
            grid = nb * cus;
            grids[{dev, lds}] = grid;
        }
    }
    if (grid <= 0) return false;
    hipExtLaunchKernelGGL((k_solve_lds<NMAX, G, MIXED, KIND, BLOCK, PACK>), dim3((unsigned)grid), dim3(BLOCK),
                          (uint32_t)lds, st, ev[0], ev[1], 0, sa, nelem);
    return true;
}

// Where the plan has an octant table: the LDS form over template-path cells,
// then the global form over the cluster-path list; otherwise the global form
// over every cell.  The mixed form then re-solves its fallbacks in FP64.
// G64 / G32: lanes per cell of the FP64 and of the mixed form.
template <int NMAX, int G64, int G32, int KIND>
void dispatch_kind(const SolveArgs& sa, bool mixed, long long n, int nelem, const hipEvent_t* ev, hipStream_t st) {
    constexpr int BLOCK32 = NMAX <= 24 ? 512 : 256;
    // A caller's own cell list (n = its largest length): the global-table
    // form over the list only.
    if (sa.cell_list) {
        if (mixed)
            launch_solve_list<NMAX, G32, true, KIND>(sa, n, st);
        else
            launch_solve_list<NMAX, G64, false, KIND>(sa, n, st);
        return;
    }
    // Packed x/y table addresses when both 16-bit fields stay below 32768:
    // x up to E_x * 8, y up to E_y * (E_x + 1) * 8.
    const bool pack = (long long)sa.ex * 8 < 32768 && (long long)sa.ey * (sa.ex + 1) * 8 < 32768;
    bool lds = false;
    if (nelem > 0) {
        if (mixed)
            lds = pack ? launch_solve_lds<NMAX, G32, true, KIND, BLOCK32, true>(sa, nelem, ev, st)
                       : launch_solve_lds<NMAX, G32, true, KIND, BLOCK32, false>(sa, nelem, ev, st);
        else
            lds = pack ? launch_solve_lds<NMAX, G64, false, KIND, 256, true>(sa, nelem, ev, st)
                       : launch_solve_lds<NMAX, G64, false, KIND, 256, false>(sa, nelem, ev, st);
    }
    if (!lds) {
        if (mixed)
            launch_solve<NMAX, G32, true, KIND>(sa, n, ev, st);
        else
            launch_solve<NMAX, G64, false, KIND>(sa, n, ev, st);
    }
    if (lds) {
        SolveArgs cl = sa;
        cl.cell_list = sa.cl_cells;
        cl.n_list = sa.n_cl_cells;
        if (mixed)
            launch_solve_list<NMAX, G32, true, KIND>(cl, n, st);
        else
            launch_solve_list<NMAX, G64, false, KIND>(cl, n, st);
    }
    if (mixed) {
        SolveArgs fb = sa;
        fb.cell_list = sa.fallback_cells;
        fb.n_list = sa.n_fallback;
        fb.fallback_cells = nullptr;
        fb.n_fallback = nullptr;
        launch_solve_list<NMAX, G64, false, KIND>(fb, n, st);
    }
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
    const long long pcell = ((long long)(zc + g.rz) * g.py + (y + g.ry)) * g.px + (x + g.rx);
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
    if (lane < G) solve_cell<NMAX, G, false, KIND, OCT>(a, cell, cnt, lane, a.otab_w, ox, oxy);
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
        // cell from 16 neighbours up
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

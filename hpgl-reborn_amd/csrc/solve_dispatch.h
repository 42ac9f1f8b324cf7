// solve_dispatch.h -- entry points of the solve kernels, one per
// neighbour-count bucket (max_neighbours <= 8, 16, 24, 32).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "krige_kernels.h"
#include "krige_launch.h"

namespace hpglx {

// One launch of the count-bucketed solve (solve.h, k_solve_sorted*).
struct SortedLaunch {
    bool mixed = false;
    int path = 0;                  // 0 template walk, 1 cluster search (and every cell without an LDS table)
    int nelem = 0;                 // octant-table entries staged in LDS (path 0), 0 = global tables
    int blocks_per_cu = 0;         // cap on the persistent grid's blocks per CU (0 = as many as fit)
    const hipEvent_t* ev = nullptr;  // start / stop events of this launch (entries may be null)
    // The launch covers a range of the work list no launch before it since the
    // list was built touches, so it may start before they finish
    // (hipExtAnyOrderLaunch: no barrier bit on its dispatch packet).
    bool any_order = false;
};

// The sizes up to top = 8, 16 or 24 (4 lanes per cell), or top = 32 for the
// 25-32 bucket, 4 lanes per cell, of `kind` over one path of the work list.
// Returns false when nothing was launched (the LDS table does not fit).
bool solve_sorted(int top, int kind, const SolveArgs& sa, const SortedLaunch& l, hipStream_t st);
// FP64 re-solve of the mixed form's fallbacks (sa.cell_list / sa.n_list):
// buckets 0-2 (b32 false) or bucket 3 (b32 true).
void solve_fallback(bool b32, int kind, const SolveArgs& sa, long long max_cells, hipStream_t st);

// Systems above 32 neighbours (sa.kmax > 32): one workgroup per cell, FP64,
// every cell of the grid; scratch from ws.  Returns 0, or -1 with err.
int dispatch_solve_wide(const SolveArgs& sa, int kind, Workspace& ws, const hipEvent_t* ev, hipStream_t st,
                        std::string& err);

// One level of sequential Gaussian simulation in one launch: a wave per
// listed cell runs the template search over the padded mask, the FP64 solve
// with variance (SolveArgs: kind, tables, values, outputs est/var/status) and
// the draw z * sqrt(variance) + estimate, written back as data.
struct SgsLevelArgs {
    int T;                    // template entries
    const int* tdelta;        // their deltas in the padded mask
    const int* tpack;         // their packed offsets
    unsigned char* pmask;     // padded 0/1 mask (halo h, pitch px)
    const int* cells;         // the level's cells
    int ncells;
    const double* z;          // their standard deviates (NaN: the draw gives the mean)
    const int* real;          // each cell's realization (several run in one launch)
    long long n;              // cells per realization: values, mask, results, neighbours
    long long pn;             // padded cells per realization
    float* values;            // property values (realizations x N), updated
    unsigned char* mask;      // informed masks (realizations x N), updated
};

// One unit per (bucket size, kind): solve_n<NMAX>_k<KIND>.hip.
#define HPGLX_DECL_SOLVE(NMAX, KIND)                                                                      \
    bool solve_sorted_##NMAX##_##KIND(const SolveArgs& sa, const SortedLaunch& l, hipStream_t st);      \
    void sgs_level_##NMAX##_##KIND(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st);
HPGLX_DECL_SOLVE(8, 0)
HPGLX_DECL_SOLVE(8, 1)
HPGLX_DECL_SOLVE(8, 2)
HPGLX_DECL_SOLVE(16, 0)
HPGLX_DECL_SOLVE(16, 1)
HPGLX_DECL_SOLVE(16, 2)
HPGLX_DECL_SOLVE(24, 0)
HPGLX_DECL_SOLVE(24, 1)
HPGLX_DECL_SOLVE(24, 2)
HPGLX_DECL_SOLVE(32, 0)
HPGLX_DECL_SOLVE(32, 1)
HPGLX_DECL_SOLVE(32, 2)
#undef HPGLX_DECL_SOLVE
// simulation levels above 32 neighbours (a wavefront per cell; solve_wide.hip)
void sgs_level_48_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st);
void sgs_level_48_1(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st);
void sgs_level_64_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st);
void sgs_level_64_1(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st);
void solve_fallback_b24_0(const SolveArgs& sa, long long max_cells, hipStream_t st);
void solve_fallback_b24_1(const SolveArgs& sa, long long max_cells, hipStream_t st);
void solve_fallback_b24_2(const SolveArgs& sa, long long max_cells, hipStream_t st);
void solve_fallback_b32_0(const SolveArgs& sa, long long max_cells, hipStream_t st);
void solve_fallback_b32_1(const SolveArgs& sa, long long max_cells, hipStream_t st);
void solve_fallback_b32_2(const SolveArgs& sa, long long max_cells, hipStream_t st);

}  // namespace hpglx

// solve_dispatch.h -- entry points of the solve kernels, one per
// neighbour-count bucket (max_neighbours <= 8, 16, 24, 32).
#pragma once

#include <hip/hip_runtime.h>

#include <string>

#include "krige_kernels.h"
#include "krige_launch.h"

namespace hpglx {

// Launches the solve of one call: the LDS-table kernel over template-path
// cells when nelem > 0 (octant table entries), the global-table kernel over
// the cluster-path list (or over every cell), and for the mixed form the
// FP64 re-solve of its fallbacks.  ev = start/stop events for the main
// kernel (entries may be null).
void dispatch_solve_8(const SolveArgs& sa, int kind, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                      hipStream_t st);
void dispatch_solve_16(const SolveArgs& sa, int kind, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                       hipStream_t st);
void dispatch_solve_24(const SolveArgs& sa, int kind, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                       hipStream_t st);
void dispatch_solve_32(const SolveArgs& sa, int kind, bool mixed, long long n, int nelem, const hipEvent_t* ev,
                       hipStream_t st);

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
    unsigned char* pmask;     // padded 0/1 mask (halo r, pitch px)
    const int* cells;         // the level's cells
    int ncells;
    const double* z;          // their standard deviates (NaN: the draw gives the mean)
    const int* real;          // each cell's realization (several run in one launch)
    long long n;              // cells per realization: values, mask, results, neighbours
    long long pn;             // padded cells per realization
    float* values;            // property values (realizations x N), updated
    unsigned char* mask;      // informed masks (realizations x N), updated
};

// One per (bucket, kind): solve_n<NMAX>_k<KIND>.hip.
#define HPGLX_DECL_SOLVE(NMAX, KIND)                                                                      \
    void dispatch_solve_##NMAX##_##KIND(const SolveArgs& sa, bool mixed, long long n, int nelem,          \
                                        const hipEvent_t* ev, hipStream_t st);                            \
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

}  // namespace hpglx

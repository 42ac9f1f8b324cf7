// krige_launch.h -- host launcher of the grid-kriging pipeline.
#pragma once

#include <hip/hip_runtime.h>

#include <map>
#include <string>
#include <vector>

#include "plan.h"

namespace hpglx {

// A plan mirrored in device memory.
struct DevPlan {
    PlanHost h;
    double* ctab = nullptr;
    float* ctab32 = nullptr;
    int* tdx = nullptr;
    int* tdy = nullptr;
    int* tdz = nullptr;
    double* otab = nullptr;
    double* otab_w = nullptr;
    unsigned* rank_w = nullptr;
    int tsize = 0;
    // Template deltas for one grid geometry (padded pitch and tile box);
    // tkey = (px, py, lx, ly) they were computed for.
    int* tdelta = nullptr;
    int* tdelta_lds = nullptr;
    int* tpack = nullptr;
    int tpad = 0;
    int tkey[4] = {-1, -1, -1, -1};
    unsigned long long serial = 0;   // distinguishes plans that reuse an address
    std::vector<float> ctab32_host;  // host copy of ctab32 while place()'s copies may be pending
    bool upload(std::string& err);
    // Lays the device tables out in the region [base, base + cap) instead of
    // allocating them and enqueues their copies from the host tables on st;
    // `used` receives the bytes taken.  The host tables must outlive the
    // copies.  release() must not be called on a placed plan.
    bool place(char* base, size_t cap, size_t& used, hipStream_t st, std::string& err);
    void release();
};

// Named device buffers reused across calls.  Heap mode (arena == null):
// grow-only allocations.  Arena mode: every buffer of a call is carved, in
// request order, from the fixed region [arena, arena + arena_cap), nothing
// is allocated or freed; reset_arena() starts the next call's layout.  A
// request beyond the region fails with a message naming the bytes needed.
struct Workspace {
    std::map<std::string, std::pair<void*, size_t>> bufs;
    std::map<std::string, std::string> tags;  // state kept in a buffer between calls
    char* arena = nullptr;
    size_t arena_cap = 0, arena_used = 0;
    unsigned long long gen = 0;  // bumped by every allocation and release (heap mode)
    void* get(const std::string& name, size_t bytes, std::string& err);
    void reset_arena() {
        bufs.clear();
        arena_used = 0;
    }
    void drop(const std::string& name);  // frees one heap-mode buffer
    void release();
};

enum Precision { PREC_FP64 = 0, PREC_MIXED = 1 };

struct GridCall {
    int kind = 0;                 // 0 OK, 1 SK, 2 median IK
    int precision = PREC_FP64;
    int nx = 0, ny = 0, nz = 0;
    const float* values = nullptr;          // OK/SK data (device, N)
    const unsigned char* mask = nullptr;    // informed mask (device, N)
    const unsigned char* ind = nullptr;     // median IK indicator data (device, nind x N)
    const float* zfield = nullptr;          // median IK CDF mode: field (device, N)
    const float* thresholds = nullptr;      // median IK CDF mode: thresholds (device, nind)
    int nind = 0;
    const double* marg_host = nullptr;      // median IK marginal P(I=1) per threshold (host, nind)
    const int* cats_host = nullptr;         // IK categories (host, nind): indicator k = (ind == cats[k]),
                                            // float accumulation, probability row cats[k] of out_prob
    const float* lvm_mean = nullptr;        // SK with a locally varying mean (device, N)
    int auto_mean = 0;                      // SK
    double mean = 0.0;                      // SK explicit mean
    float* out_val = nullptr;
    unsigned char* out_mask = nullptr;
    unsigned char* out_class = nullptr;
    float* out_prob = nullptr;
    double* out_var = nullptr;
    double* out_est64 = nullptr;
    unsigned char* out_status = nullptr;
    int order_relations = 0;
    int* out_count = nullptr;               // debug: neighbour counts (device, N)
    int* out_nboff = nullptr;               // debug: packed offsets (device, N x kmax)
    double pivot_tol = 1e-3;
    int refine_steps = 1;
    int search_variant = 0;                 // 0 = LDS tiles when the box fits (hits staged in LDS when
                                            // max_neighbours <= 32), 1 = global walk, 2 = LDS tiles, per-hit stores
    int solve_variant = 0;                  // 0 = LDS / octant tables, 1 = full table only, 2 = no LDS table
    int search_head = 0;                    // packed offsets staged in LDS by the search (0 = default)
    int search_slabs = 0;                   // z slices per LDS search block (0 = the most resident waves)
    int cluster_cap = 0;                    // kept candidates per lane of the LDS cluster search (0 = default)
    int cluster_cand = 0;                   // candidates a block of the LDS cluster search stages (0 = default)
    int cluster_lanes = 0;                  // lanes of the full-table cluster search (multiple of 256; 0 = default)
    int blocks_per_cu = 0;                  // cap on the persistent solve grids' blocks per CU (0 = all that fit)
    int dyn_sched = 0;                      // solve chunks dealt from a counter instead of a fixed stride
    int overlap = 1;                        // independent launches may run concurrently (hipExtAnyOrderLaunch)
    bool profile = false;
};

// Events of one profiled call: start and stop of the search kernel and of
// the main solve kernel, recorded by the kernel dispatches.
struct CallTiming {
    static constexpr int N = 4;
    hipEvent_t ev[N] = {nullptr, nullptr, nullptr, nullptr};
    bool valid = false;
};

int launch_grid(DevPlan& plan, Workspace& ws, const GridCall& call, hipStream_t stream, CallTiming* timing,
                std::string& err);

// Order relations and most probable category over `ncat` probability rows of
// n cells (indicator kriging over categories); writes category and mask = 1.
int ik_finalize(long long n, int ncat, float* prob, unsigned char* out_cat, unsigned char* out_mask,
                hipStream_t stream, std::string& err);

}  // namespace hpglx

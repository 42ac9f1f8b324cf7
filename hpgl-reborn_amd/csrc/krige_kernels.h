// krige_kernels.h -- argument blocks shared by the grid-kriging kernels and
// their host launcher.
#pragma once

#include <hip/hip_ext.h>
#include <hip/hip_runtime.h>
#include <cstdint>

// A launch that records start/stop events on its dispatch when given them
// (profiled calls); without events a plain launch, which stream capture
// takes into a graph.
#define HPGLX_LAUNCH_EV(kernel, grid, block, lds, st, e0, e1, ...)                        \
    do {                                                                                  \
        if ((e0) != nullptr || (e1) != nullptr)                                           \
            hipExtLaunchKernelGGL(kernel, grid, block, lds, st, e0, e1, 0, __VA_ARGS__); \
        else                                                                              \
            hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__);                \
    } while (0)
// The same, and with `any` the dispatch carries no barrier bit: the kernel
// may start while earlier kernels of the stream still run (it must not read
// what they write); a later ordinary launch still waits for all of them.
#define HPGLX_LAUNCH_EV_ANY(any, kernel, grid, block, lds, st, e0, e1, ...)                              \
    do {                                                                                                 \
        if ((any) || (e0) != nullptr || (e1) != nullptr)                                                 \
            hipExtLaunchKernelGGL(kernel, grid, block, lds, st, e0, e1, (any) ? hipExtAnyOrderLaunch : 0, \
                                  __VA_ARGS__);                                                          \
        else                                                                                             \
            hipLaunchKernelGGL(kernel, grid, block, lds, st, __VA_ARGS__);                               \
    } while (0)

namespace hpglx {

// Per-cell neighbour offsets are packed into one int: dx and dy in 11 bits,
// dz in 10 bits, each biased (dx + 1024, dy + 1024, dz + 512).
__host__ __device__ inline int pack_off(int dx, int dy, int dz) {
    return (dx + 1024) | ((dy + 1024) << 11) | ((dz + 512) << 22);
}
__host__ __device__ inline void unpack_off(int p, int& dx, int& dy, int& dz) {
    dx = (p & 2047) - 1024;
    dy = ((p >> 11) & 2047) - 1024;
    dz = ((unsigned)p >> 22) - 512;
}

// Cell status written by the solve kernel (same codes as the oracle).
enum CellStatus { ST_KRIGED = 0, ST_NO_NEIGHBOURS = 1, ST_SINGULAR = 2, ST_INFORMED = 3 };

// count[] markers set by the search kernels.  A non-negative count holds the
// number of neighbours in bits 0-15; CNT_CLUSTER marks a cell whose neighbours
// came from the cluster path (offsets up to 4r - 2 instead of r).
enum { CNT_INFORMED = -1, CNT_PENDING_CLUSTER = -2, CNT_CLUSTER = 1 << 16, CNT_MASK = 0xFFFF };

struct GridDims {
    int nx, ny, nz;
    long long n;
    int rx, ry, rz;          // search radiuses (= cluster sizes)
    int hx, hy, hz;          // padded mask halo: the template's reach (<= r)
    int px, py, pz;          // padded mask dims (n + 2h)
    int cx, cy, cz;          // cluster box dims (n / r + 2)
};

struct SearchArgs {
    GridDims g;
    int kind;                      // 0 OK, 1 SK, 2 MIK
    int kmax;
    const unsigned char* mask;     // input mask (N)
    const unsigned char* pmask;    // padded 0/1 mask
    const int* tdelta;             // template: padded-linear delta
    const int* tdelta_lds;         // template: delta in the tile's LDS box
    const int* tpack;              // template: packed offset
    int tsize;
    int lx, ly, lz;                // LDS box of a tile (tile + 2r halo; lx multiple of 4)
    const unsigned char* cl_nbhd_exceeded;  // per cluster: any of its 27 neighbours exceeded
    int* count;                    // per cell neighbour count / marker
    int* nboff;                    // per cell kmax packed offsets
    int* cl_cells;                 // list of cluster-path cells
    int* n_cl_cells;               // its length (device counter)
    int cluster_list;              // 1: classify cluster-path cells and list them; 0: leave the cells of
                                   // within-limit clusters to the octant cluster search
};

struct ClusterSearchArgs {
    GridDims g;
    int kmax;
    const int* cl_cells;
    const int* n_cl_cells;
    const int* cl_count;           // per cluster: informed count
    const int* cl_nodes;           // per cluster: ascending node list at ci * node_cap
    int node_cap;                  // list capacity per cluster (the limit, at least 1)
    long long limit;
    const double* ctab;            // covariance table over +-W
    int wx, wy, wz;                // table half-widths
    double threshold;              // C(0) / 100
    double* skey;                  // scratch keys   [nlanes][capacity]
    int* soff;                     // scratch offsets [nlanes][capacity]
    int capacity;                  // scratch entries per lane
    int nlanes;                    // total lanes of the launch
    int* count;
    int* nboff;
    int* overflow;                 // set when a candidate list exceeds capacity
};

struct SolveArgs {
    GridDims g;
    int kind;
    int kmax;
    const float* values;           // OK/SK input values (N)
    const unsigned char* mask;     // input mask (N)
    const unsigned char* ind;      // MIK: indicator data, nind arrays of N
    const float* zfield;           // MIK CDF mode: continuous field (N), else null
    const float* thresholds;       // MIK CDF mode: nind thresholds
    int nind;
    const double* mean_dev;        // SK automatic mean (device), or null
    double mean;                   // SK explicit mean
    double marg[16];               // MIK: nind marginal probabilities
    int cats[16];                  // IK categories: indicator k is (ind == cats[k]) when has_cats
    int has_cats;
    const float* lvm_mean;         // SK with a locally varying mean (N), else null
    const int* count;
    const int* nboff;
    const double* ctab;
    const float* ctab32;
    int wx, wy, wz;
    const double* otab;            // octant table (or null), staged in LDS by k_solve_lds
    int ex, ey, ez;                // its extents (|d| <= e)
    const double* otab_w;          // octant table over |d| <= (wx, wy, wz) (or null)
    double c0;                     // C(0)
    double pivot_tol;              // mixed precision: FP64 fallback threshold
    int refine_steps;
    // outputs (any may be null)
    float* out_val;                // OK/SK estimate (N)
    unsigned char* out_mask;       // OK/SK/MIK mask (N)
    unsigned char* out_class;      // MIK: indicator class per threshold (nind x N)
    float* out_prob;               // MIK: probability per threshold (nind x N)
    double* out_var;               // kriging variance (N), -1 when not kriged
    double* out_est64;             // estimate in FP64 (nind x N for MIK)
    unsigned char* out_status;     // CellStatus (N)
    int order_relations;           // MIK CDF: apply the order-relation correction
    const int* cell_list;          // when set: krige only these cells
    const int* n_list;             // its length (device)
    const int* cl_cells;           // cluster-path cells (solved from the global table)
    const int* n_cl_cells;         // their number (device)
    int* fallback_cells;           // mixed precision: cells queued for the FP64 form
    int* n_fallback;               // its length (device counter)
    const int* sorted;             // work list ordered by path and count bucket (k_bucket_scatter), -1 = gap
    const int* seg;                // its segments: [begin, end) of (path, bucket) at seg[2 * (4 path + b)]
    int* sched;                    // dynamic chunk dealing: the launch's chunk counter (null: fixed stride)
};

// Count buckets of the solve: the register solver sizes 8, 16, 24 and 32
// (bucket 0 also takes cells without neighbours).
__host__ __device__ inline int count_bucket(int n) { return n <= 8 ? 0 : (n <= 16 ? 1 : (n <= 24 ? 2 : 3)); }
// Neighbour count of a count[] entry (0 for the negative markers).
__host__ __device__ inline int count_n(int cnt) { return cnt < 0 ? 0 : (cnt & CNT_MASK); }

// Work-list chunk: the cells one wave solves together (4 lanes per cell);
// segments start at multiples of it.
constexpr int SOLVE_CHUNK = 16;

// Layout of the per-call device counters (ints; cleared by k_cluster_build).
enum {
    CTR_N_CL_CELLS = 0,
    CTR_OVERFLOW = 2,
    CTR_N_FALLBACK = 3,
    CTR_N_OVF = 4,
    CTR_N_OVF2 = 5,
    CTR_SEG = 16,        // (path, bucket) segments of the work list: 2 x 4 x [begin, end)
    CTR_BIN_TOTAL = 32,  // cells per work-list bin (NBINS)
    CTR_SCHED = 100,     // chunk counters of the solve launches (dynamic dealing): 2 paths x 2 sizes
    CTR_COUNT = 128      // ints cleared per call
};
constexpr int NBINS = 66;  // work-list bins: 2 paths x 33 counts

}  // namespace hpglx

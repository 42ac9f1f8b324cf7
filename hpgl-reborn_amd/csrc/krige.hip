// krige.hip -- gfx950 kernels of the grid-kriging hot path.
//
// Pipeline for one call (all on one stream, no host synchronisation):
//   k_cluster_build    per cluster: padded 0/1 mask, informed count, ascending
//                      node list; clears the call's counters
//   k_cluster_flags    per cluster: does any of its 27 neighbours (itself
//                      included) exceed the limit (then its cells take the
//                      template walk)
//   k_search_tile      cell paths, then the template walk in LDS tiles, one
//                      lane per cell (k_search_tmpl: the same from global memory)
//   k_search_cluster_oct, _list_oct, k_search_cluster
//                      cluster path: gather, rank, libstdc++-identical
//                      introsort, first max_neighbours
//   k_bucket_count, _scan, _scatter
//                      the work list: cells by path and neighbour count
//   k_solve_sorted_lds, k_solve_sorted (solve.h)
//                      covariance assembly + Cholesky + SK/OK/MIK solve +
//                      combine, G lanes per cell, persistent over the list;
//                      k_solve_fallback re-solves mixed-precision fallbacks
//
// Reference semantics reproduced (paths relative to the reference repository):
//   template walk   src/geo_bsd/hpgl/sugarbox_neighbour_lookup.h:37-62
//   path selection  src/geo_bsd/hpgl/sugarbox_indexed_neighbour_lookup.h:133-140,
//                   src/geo_bsd/hpgl/clusterizer.cpp:169-189
//   cluster path    src/geo_bsd/hpgl/sugarbox_indexed_neighbour_lookup.h:141-190,
//                   src/geo_bsd/hpgl/clusterizer.cpp:144-167
//   weights         src/geo_bsd/hpgl/my_kriging_weights.h:60-196 (SK), :198-361 (OK)
//   combine         src/geo_bsd/hpgl/combiner.h:10-25 (float accumulation in
//                   neighbour order for OK/SK, double for median IK)
//   failure policy  src/geo_bsd/hpgl/cont_kriging.h:73-105, median_ik.cpp:60-78
//   order relations src/geo_bsd/hpgl/indicator_kriging.h:34-59
#include <hip/hip_ext.h>

#include "introsort.h"
#include "krige_kernels.h"
#include "krige_launch.h"
#include "solve_dispatch.h"

#include <algorithm>
#include <map>
#include <mutex>
#include <utility>
#include <vector>

namespace hpglx {

// ------------------------------------------------------------------ helpers
__global__ void k_fill_u8(unsigned char* p, long long n, unsigned char v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

__global__ void k_fill_i32(int* p, long long n, int v) {
    long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    long long stride = (long long)gridDim.x * blockDim.x;
    for (; i < n; i += stride) p[i] = v;
}

// Indicator kriging over categories, last step per cell
// (indicator_kriging.h:34-59, 141-147; cdf_utils.cpp most_probable_category):
// clamp the category probabilities to [0,1], one forward pass replacing each
// decreasing adjacent pair by its average (float sum halved in double), then
// the first category of largest probability.  Probabilities are rewritten
// corrected; the rows are categories, N cells each.
__global__ void k_ik_finalize(long long n, int ncat, float* prob, unsigned char* out_cat, unsigned char* out_mask) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    auto clamp01 = [](float v) { return v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v); };
    float prev = clamp01(prob[i]);
    double best = 0.0;
    int best_k = 0;
    for (int k = 1; k <= ncat; ++k) {
        float cur = 0.0f;
        if (k < ncat) {
            cur = clamp01(prob[(long long)k * n + i]);
            if (prev > cur) {
                const float avg = (float)((double)(prev + cur) / 2.0);
                prev = avg;
                cur = avg;
            }
        }
        // entry k - 1 is final once the pair (k - 1, k) has been seen
        prob[(long long)(k - 1) * n + i] = prev;
        if (k == 1 || (double)prev > best) {
            best = prev;
            best_k = k - 1;
        }
        prev = cur;
    }
    if (out_cat) out_cat[i] = (unsigned char)best_k;
    if (out_mask) out_mask[i] = 1;
}

int ik_finalize(long long n, int ncat, float* prob, unsigned char* out_cat, unsigned char* out_mask, hipStream_t st,
                std::string& err) {
    const long long blocks = (n + 255) / 256;
    hipLaunchKernelGGL(k_ik_finalize, dim3((unsigned)blocks), dim3(256), 0, st, n, ncat, prob, out_cat, out_mask);
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        err = std::string("indicator kriging finalize launch: ") + hipGetErrorString(e);
        return -1;
    }
    return 0;
}

__device__ __forceinline__ void cell_xyz(const GridDims& g, long long cell, int& x, int& y, int& z) {
    const long long nxy = (long long)g.nx * g.ny;
    z = (int)(cell / nxy);
    const long long rem = cell - (long long)z * nxy;
    y = (int)(rem / g.nx);
    x = (int)(rem - (long long)y * g.nx);
}

__device__ __forceinline__ int cluster_of(const GridDims& g, int x, int y, int z) {
    return ((z / g.rz) * g.cy + (y / g.ry)) * g.cx + (x / g.rx);
}

// Appends the cells of the lanes with `take` to list[*n], in lane order, with
// one atomic per wave (per-lane atomics on one counter serialise).  Every
// lane of the wave calls it; blocks are one-dimensional.
__device__ __forceinline__ void wave_append(int* list, int* n, bool take, int cell) {
    const unsigned long long b = __ballot(take);
    if (b == 0) return;
    const int lane = threadIdx.x & 63;
    const int leader = __ffsll(b) - 1;
    int base = 0;
    if (lane == leader) base = atomicAdd(n, __popcll(b));
    base = __shfl(base, leader);
    if (take) list[base + __popcll(b & ((1ull << lane) - 1ull))] = cell;
}

// ------------------------------------------------------------ cluster build
// One block (256 lanes) per cluster (clusterizer.cpp:14-189: cluster = cell
// coordinate / radius, the grid split into boxes of r cells).  The block
// writes its cells into the padded mask, counts its informed cells and lists
// them, in ascending cell order, at cl_nodes[ci * cap] (at most cap = limit
// entries, which holds every cluster within the limit).  Lists of clusters
// within the limit are exactly what the cluster path reads; those of exceeded
// clusters are never read.  Block 0 also clears the call's counters.
__global__ __launch_bounds__(256) void k_cluster_build(GridDims g, const unsigned char* mask, unsigned char* pmask,
                                                       int* cl_count, int* cl_nodes, int cap, int* counters,
                                                       int ncounters) {
    __shared__ int wsum[4];
    const int tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    const int ci = blockIdx.x;
    if (ci == 0 && tid < ncounters) counters[tid] = 0;
    const int kx = ci % g.cx, ky = (ci / g.cx) % g.cy, kz = ci / (g.cx * g.cy);
    const int xa = kx * g.rx, ya = ky * g.ry, za = kz * g.rz;
    const int wx = max(0, min(g.nx, xa + g.rx) - xa), wy = max(0, min(g.ny, ya + g.ry) - ya),
              wz = max(0, min(g.nz, za + g.rz) - za);
    const int ncell = wx * wy * wz;
    int total = 0;
    for (int i0 = 0; i0 < ncell; i0 += 256) {
        const int i = i0 + tid;
        bool inf = false;
        long long cell = 0;
        if (i < ncell) {
            const int lx = i % wx, r = i / wx, ly = r % wy, lz = r / wy;
            const int x = xa + lx, y = ya + ly, z = za + lz;
            cell = ((long long)z * g.ny + y) * g.nx + x;
            inf = mask[cell] != 0;
            pmask[((long long)(z + g.hz) * g.py + (y + g.hy)) * g.px + (x + g.hx)] = inf ? 1 : 0;
        }
        // Box order (z, y, x) is ascending cell order, so ranks by a
        // block-wide prefix count give the ascending list.
        const unsigned long long b = __ballot(inf);
        if (lane == 0) wsum[w] = __popcll(b);
        __syncthreads();
        int before = 0, chunk = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            before += k < w ? wsum[k] : 0;
            chunk += wsum[k];
        }
        const int pos = total + before + __popcll(b & ((1ull << lane) - 1ull));
        if (inf && pos < cap) cl_nodes[(long long)ci * cap + pos] = (int)cell;
        total += chunk;
        __syncthreads();
    }
    if (tid == 0) cl_count[ci] = total;
}

// Mean of informed values in double (calc_mean.cpp:6-30); per-block partial
// sums in a fixed order, so the result is reproducible run to run.
__global__ void k_mean_partial(long long n, const float* values, const unsigned char* mask, double* psum,
                               int* pcnt) {
    __shared__ double ss[256];
    __shared__ int sc[256];
    double s = 0;
    int c = 0;
    for (long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x; i < n; i += (long long)gridDim.x * blockDim.x)
        if (mask[i]) {
            s += values[i];
            c += 1;
        }
    ss[threadIdx.x] = s;
    sc[threadIdx.x] = c;
    __syncthreads();
    for (int o = 128; o > 0; o >>= 1) {
        if ((int)threadIdx.x < o) {
            ss[threadIdx.x] += ss[threadIdx.x + o];
            sc[threadIdx.x] += sc[threadIdx.x + o];
        }
        __syncthreads();
    }
    if (threadIdx.x == 0) {
        psum[blockIdx.x] = ss[0];
        pcnt[blockIdx.x] = sc[0];
    }
}

__global__ void k_mean_final(int nblocks, const double* psum, const int* pcnt, double* mean_out) {
    if (threadIdx.x != 0 || blockIdx.x != 0) return;
    double s = 0;
    long long c = 0;
    for (int i = 0; i < nblocks; ++i) {
        s += psum[i];
        c += pcnt[i];
    }
    mean_out[0] = c == 0 ? 0.0 : s / (double)c;
}

__global__ void k_cluster_flags(GridDims g, const int* cl_count, long long limit, unsigned char* nbhd_exceeded) {
    int ci = blockIdx.x * blockDim.x + threadIdx.x;
    const int ncl = g.cx * g.cy * g.cz;
    if (ci >= ncl) return;
    const int kx = ci % g.cx, ky = (ci / g.cx) % g.cy, kz = ci / (g.cx * g.cy);
    unsigned char ex = 0;
    for (int k = kz - 1; k <= kz + 1; ++k)
        for (int j = ky - 1; j <= ky + 1; ++j)
            for (int i = kx - 1; i <= kx + 1; ++i) {
                if (i < 0 || j < 0 || k < 0 || i >= g.cx || j >= g.cy || k >= g.cz) continue;
                if ((long long)cl_count[(k * g.cy + j) * g.cx + i] > limit) ex = 1;
            }
    nbhd_exceeded[ci] = ex;
}

// Template offsets as deltas in the padded mask and in a tile's LDS box;
// entries from tsize up to the padded length are zero (never counted).
__global__ void k_tdelta(GridDims g, int tsize, int tpad, const int* tdx, const int* tdy, const int* tdz,
                         int* tdelta, int* tdelta_lds, int lx, int ly, int* tpack) {
    int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= tpad) return;
    if (t >= tsize) {
        tdelta[t] = 0;
        tdelta_lds[t] = 0;
        tpack[t] = 0;
        return;
    }
    tdelta[t] = (tdz[t] * g.py + tdy[t]) * g.px + tdx[t];
    tdelta_lds[t] = (tdz[t] * ly + tdy[t]) * lx + tdx[t];
    tpack[t] = pack_off(tdx[t], tdy[t], tdz[t]);
}

// --------------------------------------------------------- tiled template walk
// One block per TILE_X x TILE_Y tile of `nslab` consecutive z slices (256
// lanes per slice, one lane per cell).  The block copies the padded-mask box
// its cells can reach (tile + 2h halo, nslab + 2 hz slices deep) and the
// packed offsets of the first SEARCH_HEAD template entries into LDS.  Every
// lane then walks the template in reference order, 32 entries per step: the
// step's box deltas are wave-uniform and come through scalar loads (the next
// step's are loaded while this step's probes are in flight), each probe is
// one LDS byte read, two probes share a register (low and high half), and the
// 32 results form a bit mask whose set bits (ctz order) are the hits until
// max_neighbours.  Slice groups are dispatched from both grid faces inwards,
// because cells near a z face need the most probes and would otherwise finish
// last.  Cells whose 27 clusters are all below the limit are handed to the
// cluster path instead.  Several slices per block share one box: its halo in
// z is paid once per block, so more cells fit in LDS at a time.
constexpr int TILE_X = 16, TILE_Y = 16;
constexpr int TILE_CELLS = TILE_X * TILE_Y;
constexpr int TMPL_U = 32;       // probes per step; template arrays are padded to a multiple (plus one more)
constexpr int LOAD_BATCH = 8;    // loads in flight per thread while filling LDS
constexpr int SEARCH_HEAD = 1024;  // packed offsets staged in LDS
constexpr int SEARCH_MAX_SLABS = 4;  // z slices per block (1024 lanes)
constexpr size_t kSearchLdsMax = 160 * 1024;  // one workgroup's LDS

typedef const int __attribute__((address_space(4)))* const_int_ptr;
typedef unsigned short probe_pair __attribute__((ext_vector_type(2)));

// LDS bytes of a search block: the box, the head of packed offsets and, when
// staged, 16-bit hit indices (max_neighbours per lane) and per-lane counts.
static inline size_t search_tile_lds(int lx, int ly, int hz, int nslab, int K, bool staged) {
    size_t b = ((size_t)lx * ly * (nslab + 2 * hz) + 15) / 16 * 16 + (size_t)SEARCH_HEAD * 4;
    if (staged) b += ((size_t)TILE_CELLS * nslab * K * 2 + 15) / 16 * 16 + (size_t)TILE_CELLS * nslab * 4;
    return b;
}

// head = number of staged packed offsets (<= SEARCH_HEAD); blockDim.x =
// 256 * nslab.
//
// STAGE (max_neighbours <= 32, template shorter than 65,536 entries): a lane
// keeps its hits in LDS as 16-bit template indices, and at the end each wave
// writes its four 16-cell rows' neighbour lists -- one contiguous run of
// 16 max_neighbours ints per row -- with consecutive lanes on consecutive
// ints.  Without it every hit is a 4-byte store of its own (after an LDS read
// of its packed offset), and a wave's store touches 64 cache lines (the
// lanes' lists are max_neighbours ints apart).
template <bool STAGE>
__global__ __launch_bounds__(TILE_CELLS * SEARCH_MAX_SLABS) void k_search_tile(SearchArgs a, int head) {
    extern __shared__ unsigned int lds_dw[];
    const GridDims& g = a.g;
    const int qx = a.lx >> 2;  // dwords per box row
    const int nrows = a.ly * a.lz;
    const int nslab = (int)blockDim.x / TILE_CELLS;
    const int nlanes = (int)blockDim.x;
    const unsigned char* box = (const unsigned char*)lds_dw;
    int* tpk = (int*)(lds_dw + ((qx * nrows + 3) & ~3));
    unsigned short* hits = (unsigned short*)(tpk + ((SEARCH_HEAD + 3) & ~3));
    // per-lane counts after the hits (16-byte aligned, as the launcher sizes them)
    int* hitn = (int*)((unsigned char*)hits + ((nlanes * a.kmax * 2 + 15) & ~15));
    const int tid = threadIdx.x;
    const int nzb = (int)gridDim.z;
    const int bz = (blockIdx.z & 1) ? nzb - 1 - (int)(blockIdx.z >> 1) : (int)(blockIdx.z >> 1);
    const int x0 = blockIdx.x * TILE_X, y0 = blockIdx.y * TILE_Y, z0 = bz * nslab;

    // The cells' paths first: informed cells are done, cells whose clusters
    // are all within the limit take the cluster path (listed here, or left
    // to the octant cluster search, which classifies them itself), and a
    // block left with no cell to walk loads no box.
    const int slab = tid / TILE_CELLS, tc = tid % TILE_CELLS;
    const int tx = tc % TILE_X, ty = tc / TILE_X;
    const int x = x0 + tx, y = y0 + ty, z = z0 + slab;
    const bool inside = x < g.nx && y < g.ny && z < g.nz;
    const long long cell = inside ? ((long long)z * g.ny + y) * g.nx + x : 0;
    bool active = false, pending = false;
    if (inside) {
        // both loads issue before either is used
        const bool ex = a.cl_nbhd_exceeded[cluster_of(g, x, y, z)] != 0;
        const bool needs = a.kind == 2 || !a.mask[cell];
        if (ex || a.cluster_list) {
            if (!needs) {
                a.count[cell] = CNT_INFORMED;
            } else if (ex) {
                active = true;
            } else {
                a.count[cell] = CNT_PENDING_CLUSTER;
                pending = true;
            }
        }
    }
    wave_append(a.cl_cells, a.n_cl_cells, pending, (int)cell);
    if (!__syncthreads_or(active)) return;

    // Box origin is padded (x0, y0, z0) = cell (x0 - hx, y0 - hy, z0 - hz).
    // Thread -> (dword column xq, row); rows advance by rstep per pass.
    {
        const int rstep = nlanes / qx;  // qx <= 256 by the launcher
        const int xq = tid % qx;
        int row = tid / qx;
        const bool lane_ok = row < rstep;
        int yy = row % a.ly, zz = row / a.ly;
        const int px = x0 + 4 * xq;
        while (lane_ok && row < nrows) {
            unsigned v[LOAD_BATCH];
            int rr[LOAD_BATCH];
#pragma unroll
            for (int k = 0; k < LOAD_BATCH; ++k) {
                const int py = y0 + yy, pz = z0 + zz;
                const bool ok = row < nrows && px < g.px && py < g.py && pz < g.pz;
                // unconditional load (from offset 0 when out of range), so the
                // batch issues without branches
                const long long off = ok ? ((long long)pz * g.py + py) * g.px + px : 0;
                const unsigned w = *(const unsigned int*)(a.pmask + off);
                v[k] = ok ? w : 0u;
                rr[k] = row < nrows ? row : -1;
                row += rstep;
                yy += rstep;
                while (yy >= a.ly) {
                    yy -= a.ly;
                    ++zz;
                }
            }
#pragma unroll
            for (int k = 0; k < LOAD_BATCH; ++k)
                if (rr[k] >= 0) lds_dw[rr[k] * qx + xq] = v[k];
        }
    }
    head = min(head, a.tsize);
    for (int i = tid; i < head; i += nlanes) tpk[i] = a.tpack[i];
    __syncthreads();

    const int lcell = ((g.hz + slab) * a.ly + ty + g.hy) * a.lx + tx + g.hx;  // the cell's own box byte
    // Inactive lanes probe around the box position of cell (0, 0) of the
    // first slice, which keeps every template delta inside the box; their
    // bits are discarded.
    const int lbase = active ? lcell : (g.hz * a.ly + g.hy) * a.lx + g.hx;
    const const_int_ptr tdel = (const_int_ptr)a.tdelta_lds;
    int cnt = 0;
    const int K = a.kmax;
    int* out = a.nboff + (active ? cell * K : 0);
    // This step's deltas; the template arrays hold one padding step past the
    // last, so the load of the step after the last stays in bounds.
    int dl[TMPL_U];
#pragma unroll
    for (int u = 0; u < TMPL_U; ++u) dl[u] = tdel[u];
    for (int t0 = 0; t0 < a.tsize; t0 += TMPL_U) {
        const bool want = active && cnt < K;
        if (!__any(want)) break;
        // 32 probes, two per register, issued before the first is used, so
        // the step waits for the LDS once; the next step's deltas load
        // meanwhile.
        probe_pair pv[TMPL_U / 2];
#pragma unroll
        for (int u = 0; u < TMPL_U / 2; ++u) {
            probe_pair p;
            p.x = box[lbase + dl[u]];
            p.y = box[lbase + dl[u + TMPL_U / 2]];
            pv[u] = p;
        }
        __builtin_amdgcn_sched_barrier(0);
        const int tn = __builtin_amdgcn_readfirstlane(t0 + TMPL_U);
#pragma unroll
        for (int u = 0; u < TMPL_U; ++u) dl[u] = tdel[tn + u];
        __builtin_amdgcn_sched_barrier(0);
        // mask bytes are 0 or 1: pair u sets bits u and u + 16
        unsigned bits = 0;
#pragma unroll
        for (int u = 0; u < TMPL_U / 2; ++u) bits |= __builtin_bit_cast(unsigned, pv[u]) << u;
        const int left = a.tsize - t0;
        if (left < TMPL_U) bits &= (1u << left) - 1u;
        if (!want) bits = 0;
        // Hits in template order until max_neighbours.
        if (STAGE) {
            while (bits != 0 && cnt < K) {
                const int u = __builtin_ctz(bits);
                hits[tid * K + cnt++] = (unsigned short)(t0 + u);
                bits &= bits - 1u;
            }
        } else if (t0 + TMPL_U <= head) {
            while (bits != 0 && cnt < K) {
                const int u = __builtin_ctz(bits);
                out[cnt++] = tpk[t0 + u];
                bits &= bits - 1u;
            }
        } else {
            while (bits != 0 && cnt < K) {
                const int u = __builtin_ctz(bits);
                out[cnt++] = a.tpack[t0 + u];
                bits &= bits - 1u;
            }
        }
    }
    if (active) a.count[cell] = cnt;
    if (STAGE) {
        hitn[tid] = active ? cnt : 0;
        // The wave's cells are rows 4 (w mod 4) .. + 3 of slice w / 4 (16
        // cells each); a row's lists are one run of 16 K ints in nboff.
        // Slots past a cell's count, and the lists of cells the walk skipped
        // (informed cells; cluster-path cells, whose lists the cluster search
        // writes later), are written as 0.
        __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        const int lane = tid & 63, w = tid >> 6;
        const int wz = z0 + (w >> 2), wrow = 4 * (w & 3);
        const int nvalid = min(TILE_X, g.nx - x0);
        // i / K for i < 16 K <= 512 and K <= 32: (i m) >> 16 with m = 65536 / K + 1
        const unsigned m = 65536u / (unsigned)K + 1u;
        for (int r = 0; r < 4 && wz < g.nz; ++r) {
            const int yy = y0 + wrow + r;
            if (yy >= g.ny) break;
            int* dst = a.nboff + (((long long)wz * g.ny + yy) * g.nx + x0) * K;
            const int src0 = (w >> 2) * TILE_CELLS + (wrow + r) * TILE_X;
            for (int i = lane; i < nvalid * K; i += 64) {
                const int cx = (int)(((unsigned)i * m) >> 16);
                const int slot = i - cx * K;
                // slots past the count are written as 0; every index read was staged
                int v = 0;
                if (slot < hitn[src0 + cx]) {
                    const int idx = hits[(src0 + cx) * K + slot];
                    v = idx < head ? tpk[idx] : a.tpack[idx];
                }
                dst[i] = v;
            }
        }
    }
}

// ------------------------------------------------------------ template walk
// Global-memory form for boxes too large for LDS: one lane per cell, the
// template read with uniform vector loads, the mask from the padded array.
__global__ __launch_bounds__(256) void k_search_tmpl(SearchArgs a) {
    const long long cell = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    const GridDims& g = a.g;
    bool active = false;
    // Inactive lanes probe around grid cell (0, 0, 0): every template delta
    // stays inside the padded mask; their bits are discarded.
    long long pbase = ((long long)g.hz * g.py + g.hy) * g.px + g.hx;
    bool pending = false;
    if (cell < g.n) {
        int x, y, z;
        cell_xyz(g, cell, x, y, z);
        const bool ex = a.cl_nbhd_exceeded[cluster_of(g, x, y, z)] != 0;
        const bool needs = a.kind == 2 || !a.mask[cell];
        if (ex || a.cluster_list) {
            if (!needs) {
                a.count[cell] = CNT_INFORMED;
            } else if (ex) {
                active = true;
                pbase = ((long long)(z + g.hz) * g.py + (y + g.hy)) * g.px + (x + g.hx);
            } else {
                a.count[cell] = CNT_PENDING_CLUSTER;
                pending = true;
            }
        }
    }
    wave_append(a.cl_cells, a.n_cl_cells, pending, (int)cell);
    int cnt = 0;
    const int K = a.kmax;
    int* out = a.nboff + (active ? cell * K : 0);
    for (int t0 = 0; t0 < a.tsize; t0 += TMPL_U) {
        const bool want = active && cnt < K;
        if (!__any(want)) break;
        // The template arrays are padded to a multiple of TMPL_U, so these
        // loads are unconditional and issue together.
        unsigned bits = 0;
#pragma unroll
        for (int u = 0; u < TMPL_U; ++u) {
            const unsigned m = a.pmask[pbase + a.tdelta[t0 + u]];
            bits |= m << u;
        }
        const int left = a.tsize - t0;
        if (left < TMPL_U) bits &= (1u << left) - 1u;
        if (!want) bits = 0;
        while (bits != 0 && cnt < K) {
            const int u = __builtin_ctz(bits);
            out[cnt++] = a.tpack[t0 + u];
            bits &= bits - 1u;
        }
    }
    if (active) a.count[cell] = cnt;
}

// Scratch store for the introsort: lane l's elements are contiguous
// (element i at l * capacity + i), so one lane's sort stays in a few cache
// lines.
struct LaneStore {
    struct Item {
        double key;
        int off;
    };
    double* keys;
    int* offs;
    int stride;
    __device__ double key(int i) const { return keys[(long long)i * stride]; }
    __device__ Item get(int i) const {
        Item it;
        it.key = keys[(long long)i * stride];
        it.off = offs[(long long)i * stride];
        return it;
    }
    __device__ void put(int i, const Item& it) {
        keys[(long long)i * stride] = it.key;
        offs[(long long)i * stride] = it.off;
    }
    __device__ void swap(int i, int j) {
        Item a = get(i), b = get(j);
        put(i, b);
        put(j, a);
    }
};

// Cluster path with the full covariance table, one lane per pending cell of
// a cell list (all cluster-path cells when the plan has no octant tables,
// else the cells k_search_cluster_oct could not hold), persistent.
__global__ __launch_bounds__(256) void k_search_cluster(ClusterSearchArgs a) {
    const int lane = blockIdx.x * blockDim.x + threadIdx.x;
    const GridDims& g = a.g;
    const int ncells = *a.n_cl_cells;
    const int SX = 2 * a.wx + 1, SY = 2 * a.wy + 1;
    const long long center = ((long long)a.wz * SY + a.wy) * SX + a.wx;
    LaneStore st;
    st.keys = a.skey + (long long)lane * a.capacity;
    st.offs = a.soff + (long long)lane * a.capacity;
    st.stride = 1;
    for (int item = lane; item < ncells; item += a.nlanes) {
        const long long cell = a.cl_cells[item];
        if (a.count[cell] != CNT_PENDING_CLUSTER) continue;
        int x, y, z;
        cell_xyz(g, cell, x, y, z);
        const int kx = x / g.rx, ky = y / g.ry, kz = z / g.rz;
        int m = 0;
        bool overflow = false;
        for (int k = kz - 1; k <= kz + 1; ++k)
            for (int j = ky - 1; j <= ky + 1; ++j)
                for (int i = kx - 1; i <= kx + 1; ++i) {
                    if (i < 0 || j < 0 || k < 0 || i >= g.cx || j >= g.cy || k >= g.cz) continue;
                    const int ci = (k * g.cy + j) * g.cx + i;
                    const int c = a.cl_count[ci];
                    const int* nodes = a.cl_nodes + (long long)ci * a.node_cap;
                    for (int s = 0; s < c; ++s) {
                        const int nd = nodes[s];
                        int nx_, ny_, nz_;
                        cell_xyz(g, nd, nx_, ny_, nz_);
                        const int dx = nx_ - x, dy = ny_ - y, dz = nz_ - z;
                        // beyond the table the covariance is below C(0)/100
                        const bool in = abs(dx) <= a.wx && abs(dy) <= a.wy && abs(dz) <= a.wz;
                        const double key = in ? a.ctab[center + ((long long)dz * SY + dy) * SX + dx] : 0.0;
                        if (in && key > a.threshold) {
                            if (m < a.capacity) {
                                LaneStore::Item it;
                                it.key = key;
                                it.off = pack_off(dx, dy, dz);
                                st.put(m, it);
                                ++m;
                            } else {
                                overflow = true;
                            }
                        }
                    }
                }
        if (overflow) atomicAdd(a.overflow, 1);
        Introsort<LaneStore> srt(st);
        srt.sort(m, a.kmax);
        const int n = m < a.kmax ? m : a.kmax;
        for (int i = 0; i < n; ++i) a.nboff[cell * a.kmax + i] = st.get(i).off;
        a.count[cell] = n | CNT_CLUSTER;
    }
}

// ----------------------------------------------- cluster path, octant ranks
// Cells of one cluster share the candidate list of its 27 neighbouring
// clusters.  A block takes a cluster, stages that list in LDS (packed cell
// coordinates, in the reference's gather order: neighbour clusters k, j, i,
// nodes ascending), and each lane takes one of the cluster's pending cells:
// for every candidate the key is the dense rank of its covariance (octant
// rank table; 0 = not above C(0)/100), and the kept candidates go to a
// per-lane LDS buffer as one word, rank << qbits | candidate index, which the
// introsort orders exactly as std::sort orders the keys.  Cells with more than
// `cap1` kept candidates are appended to an overflow list for
// k_search_cluster_list_oct, and so are all pending cells of a cluster whose
// 27 neighbours hold more than `cand_cap` candidates.  The LDS holds the
// sort buffers (cap1 words per lane) and cand_cap coordinates: at the
// defaults 79 KB, two blocks (eight waves) per CU.
constexpr int CL_BLOCK = 256;
constexpr int kClusterSortCap = 64;    // kept candidates per lane in LDS
constexpr int kClusterCandCap = 3840;  // candidate coordinates per block in LDS
constexpr int CL_GATHER = 8;         // rank gathers in flight per lane

struct WordStore {
    struct Item {
        unsigned key;
        unsigned w;
    };
    unsigned* buf;  // element i at buf[i * stride]
    int qbits;
    int stride = CL_BLOCK;
    __device__ unsigned key(int i) const { return buf[i * stride] >> qbits; }
    __device__ Item get(int i) const {
        Item it;
        it.w = buf[i * stride];
        it.key = it.w >> qbits;
        return it;
    }
    __device__ void put(int i, const Item& it) { buf[i * stride] = it.w; }
    __device__ void swap(int i, int j) {
        const unsigned a = buf[i * stride], b = buf[j * stride];
        buf[i * stride] = b;
        buf[j * stride] = a;
    }
};

struct ClusterOctArgs {
    GridDims g;
    int kind;                      // 0 OK, 1 SK, 2 MIK (every cell kriged)
    const unsigned char* mask;     // input mask (N)
    int* n_cl_cells;               // cluster-path cells of the call (device counter)
    int kmax;
    const int* cl_count;
    const int* cl_nodes;
    int node_cap;
    const unsigned char* nbhd_exceeded;
    const unsigned* rank_w;        // octant rank table over |d| <= W
    int wx1, wy1, wz1;             // its extents (W + 1)
    int qbits;                     // bits of the candidate index in a sort word
    int cap1;                      // sort buffer entries per lane
    int cand_cap;                  // candidate coordinates the block's LDS holds
    int* count;
    int* nboff;
    int* ovf_cells;                // cells left for k_search_cluster_list_oct
    int* n_ovf;
};

__device__ __forceinline__ unsigned pack_xyz(int x, int y, int z) {
    return (unsigned)x | ((unsigned)y << 11) | ((unsigned)z << 22);
}

// BOUNDED: candidates can lie beyond the rank table (a radius whose 27
// clusters reach farther than the covariance does); those are skipped.
template <bool BOUNDED>
__global__ __launch_bounds__(CL_BLOCK) void k_search_cluster_oct(ClusterOctArgs a) {
    extern __shared__ unsigned cl_lds[];
    __shared__ int off27[28];
    __shared__ int npend;
    const GridDims& g = a.g;
    const int tid = threadIdx.x;
    unsigned* sbuf = cl_lds;                                  // cap1 x CL_BLOCK words
    unsigned* cand = cl_lds + (size_t)a.cap1 * CL_BLOCK;      // 27 x node_cap packed coordinates
    WordStore st;
    st.buf = sbuf + tid;
    st.qbits = a.qbits;
    const int ncl = g.cx * g.cy * g.cz;
    const int K = a.kmax;
    for (int ci = blockIdx.x; ci < ncl; ci += gridDim.x) {
        if (a.nbhd_exceeded[ci]) continue;  // its cells all took the template path
        const int kx = ci % g.cx, ky = (ci / g.cx) % g.cy, kz = ci / (g.cx * g.cy);
        const int xa = kx * g.rx, ya = ky * g.ry, za = kz * g.rz;
        const int wx = max(0, min(g.nx, xa + g.rx) - xa), wy = max(0, min(g.ny, ya + g.ry) - ya),
                  wz = max(0, min(g.nz, za + g.rz) - za);
        const int ncell = wx * wy * wz;
        if (ncell == 0) continue;
        if (tid == 0) {
            npend = 0;
            int acc = 0, q = 0;
            for (int k = kz - 1; k <= kz + 1; ++k)
                for (int j = ky - 1; j <= ky + 1; ++j)
                    for (int i = kx - 1; i <= kx + 1; ++i, ++q) {
                        off27[q] = acc;
                        if (i < 0 || j < 0 || k < 0 || i >= g.cx || j >= g.cy || k >= g.cz) continue;
                        acc += a.cl_count[(k * g.cy + j) * g.cx + i];
                    }
            off27[27] = acc;
        }
        __syncthreads();
        const int total = off27[27];
        // The cluster's cells: informed ones are done (but for median IK,
        // which krige every cell); the rest take the cluster path.
        if (total > a.cand_cap) {
            // too many candidates to stage: the cluster's cells take the list kernel
            for (int c = tid; c < ncell; c += CL_BLOCK) {
                const int lx = c % wx, r = c / wx, ly = r % wy, lz = r / wy;
                const long long cell = ((long long)(za + lz) * g.ny + (ya + ly)) * g.nx + (xa + lx);
                if (a.kind != 2 && a.mask[cell]) {
                    a.count[cell] = CNT_INFORMED;
                    continue;
                }
                a.count[cell] = CNT_PENDING_CLUSTER;
                atomicAdd(&npend, 1);
                const int pos = atomicAdd(a.n_ovf, 1);
                a.ovf_cells[pos] = (int)cell;
            }
            __syncthreads();  // off27 is rewritten for the next cluster
            if (tid == 0) atomicAdd(a.n_cl_cells, npend);
            continue;
        }
        for (int idx = tid; idx < total; idx += CL_BLOCK) {
            int q = 0;
            while (off27[q + 1] <= idx) ++q;
            const int k = kz - 1 + q / 9, j = ky - 1 + (q / 3) % 3, i = kx - 1 + q % 3;
            const int nd = a.cl_nodes[(long long)((k * g.cy + j) * g.cx + i) * a.node_cap + (idx - off27[q])];
            int x, y, z;
            cell_xyz(g, nd, x, y, z);
            cand[idx] = pack_xyz(x, y, z);
        }
        __syncthreads();
        for (int c = tid; c < ncell; c += CL_BLOCK) {
            const int lx = c % wx, r = c / wx, ly = r % wy, lz = r / wy;
            const int x = xa + lx, y = ya + ly, z = za + lz;
            const long long cell = ((long long)z * g.ny + y) * g.nx + x;
            if (a.kind != 2 && a.mask[cell]) {
                a.count[cell] = CNT_INFORMED;
                continue;
            }
            atomicAdd(&npend, 1);
            int m = 0;
            bool over = false;
            for (int q0 = 0; q0 < total; q0 += CL_GATHER) {
                unsigned rk[CL_GATHER];
#pragma unroll
                for (int u = 0; u < CL_GATHER; ++u) {
                    const int q = q0 + u < total ? q0 + u : total - 1;
                    const unsigned p = cand[q];
                    const int dx = (int)(p & 2047u) - x, dy = (int)((p >> 11) & 2047u) - y, dz = (int)(p >> 22) - z;
                    const int ax = dx < 0 ? -dx : dx, ay = dy < 0 ? -dy : dy, az = dz < 0 ? -dz : dz;
                    // beyond the table the covariance is below C(0)/100 (rank 0)
                    const bool in = !BOUNDED || (ax < a.wx1 && ay < a.wy1 && az < a.wz1);
                    rk[u] = in ? a.rank_w[((long long)az * a.wy1 + ay) * a.wx1 + ax] : 0u;
                }
#pragma unroll
                for (int u = 0; u < CL_GATHER; ++u) {
                    if (q0 + u < total && rk[u] != 0u) {
                        if (m < a.cap1)
                            sbuf[m * CL_BLOCK + tid] = (rk[u] << a.qbits) | (unsigned)(q0 + u);
                        else
                            over = true;
                        ++m;
                    }
                }
            }
            if (over) {
                a.count[cell] = CNT_PENDING_CLUSTER;
                const int pos = atomicAdd(a.n_ovf, 1);
                a.ovf_cells[pos] = (int)cell;
                continue;
            }
            Introsort<WordStore> srt(st);
            srt.sort(m, K);
            const int n = m < K ? m : K;
            const unsigned qmask = (1u << a.qbits) - 1u;
            for (int i = 0; i < n; ++i) {
                const unsigned p = cand[sbuf[i * CL_BLOCK + tid] & qmask];
                a.nboff[cell * K + i] =
                    pack_off((int)(p & 2047u) - x, (int)((p >> 11) & 2047u) - y, (int)(p >> 22) - z);
            }
            a.count[cell] = n | CNT_CLUSTER;
        }
        __syncthreads();
        if (tid == 0) atomicAdd(a.n_cl_cells, npend);
    }
}

// Cluster path for a cell list (the overflow of k_search_cluster_oct): one
// lane per cell, the candidates gathered from the 27 clusters' node lists in
// global memory, rank keys as above, and a deeper per-lane LDS sort buffer
// (cap2 words; CL_LIST_BLOCK lanes per block: 32 KB at the default, five
// blocks per CU).  Cells with more kept candidates stay pending for
// k_search_cluster.
constexpr int CL_LIST_BLOCK = 64;
constexpr int kClusterListCap = 128;

template <bool BOUNDED>
__global__ __launch_bounds__(CL_LIST_BLOCK) void k_search_cluster_list_oct(ClusterOctArgs a, const int* cells,
                                                                           const int* ncells_p, int cap2,
                                                                           int* ovf2, int* n_ovf2) {
    extern __shared__ unsigned cll_lds[];
    const GridDims& g = a.g;
    const int tid = threadIdx.x;
    unsigned* sbuf = cll_lds;  // cap2 x CL_LIST_BLOCK words, element i of lane t at i * CL_LIST_BLOCK + t
    const int ncells = *ncells_p;
    const int K = a.kmax;
    const unsigned qmask = (1u << a.qbits) - 1u;
    for (int item = blockIdx.x * CL_LIST_BLOCK + tid; item < ncells; item += gridDim.x * CL_LIST_BLOCK) {
        const long long cell = cells[item];
        if (a.count[cell] != CNT_PENDING_CLUSTER) continue;
        int x, y, z;
        cell_xyz(g, cell, x, y, z);
        const int kx = x / g.rx, ky = y / g.ry, kz = z / g.rz;
        int m = 0, q = 0;
        bool over = false;
        for (int k = kz - 1; k <= kz + 1; ++k)
            for (int j = ky - 1; j <= ky + 1; ++j)
                for (int i = kx - 1; i <= kx + 1; ++i) {
                    if (i < 0 || j < 0 || k < 0 || i >= g.cx || j >= g.cy || k >= g.cz) continue;
                    const int ci = (k * g.cy + j) * g.cx + i;
                    const int c = a.cl_count[ci];
                    const int* nodes = a.cl_nodes + (long long)ci * a.node_cap;
                    for (int s0 = 0; s0 < c; s0 += CL_GATHER) {
                        int nd[CL_GATHER];
#pragma unroll
                        for (int u = 0; u < CL_GATHER; ++u) nd[u] = nodes[s0 + u < c ? s0 + u : c - 1];
                        unsigned rk[CL_GATHER];
#pragma unroll
                        for (int u = 0; u < CL_GATHER; ++u) {
                            int nx_, ny_, nz_;
                            cell_xyz(g, nd[u], nx_, ny_, nz_);
                            const int ax = abs(nx_ - x), ay = abs(ny_ - y), az = abs(nz_ - z);
                            const bool in = !BOUNDED || (ax < a.wx1 && ay < a.wy1 && az < a.wz1);
                            rk[u] = in ? a.rank_w[((long long)az * a.wy1 + ay) * a.wx1 + ax] : 0u;
                        }
#pragma unroll
                        for (int u = 0; u < CL_GATHER; ++u) {
                            if (s0 + u < c && rk[u] != 0u) {
                                // the word keeps the candidate's node id position: q = running index
                                if (m < cap2)
                                    sbuf[m * CL_LIST_BLOCK + tid] = (rk[u] << a.qbits) | (unsigned)(q + s0 + u);
                                else
                                    over = true;
                                ++m;
                            }
                        }
                    }
                    q += c;
                }
        if (over) {
            const int pos = atomicAdd(n_ovf2, 1);
            ovf2[pos] = (int)cell;
            continue;
        }
        WordStore st;
        st.buf = sbuf + tid;
        st.qbits = a.qbits;
        st.stride = CL_LIST_BLOCK;
        Introsort<WordStore> srt(st);
        srt.sort(m, K);
        const int n = m < K ? m : K;
        for (int r = 0; r < n; ++r) {
            // recover the node from its running index over the 27 clusters
            int idx = (int)(sbuf[r * CL_LIST_BLOCK + tid] & qmask), nd = 0;
            for (int k = kz - 1; k <= kz + 1; ++k)
                for (int j = ky - 1; j <= ky + 1; ++j)
                    for (int i = kx - 1; i <= kx + 1; ++i) {
                        if (i < 0 || j < 0 || k < 0 || i >= g.cx || j >= g.cy || k >= g.cz) continue;
                        const int ci = (k * g.cy + j) * g.cx + i;
                        const int c = a.cl_count[ci];
                        if (idx >= 0 && idx < c) nd = a.cl_nodes[(long long)ci * a.node_cap + idx];
                        idx -= c;
                    }
            int nx_, ny_, nz_;
            cell_xyz(g, nd, nx_, ny_, nz_);
            a.nboff[cell * K + r] = pack_off(nx_ - x, ny_ - y, nz_ - z);
        }
        a.count[cell] = n | CNT_CLUSTER;
    }
}

// ------------------------------------------------------------------ launcher



#define HPGLX_CHECK(expr)                                              \
    do {                                                               \
        hipError_t e_ = (expr);                                        \
        if (e_ != hipSuccess) {                                        \
            err = std::string(#expr) + ": " + hipGetErrorString(e_);   \
            return -1;                                                 \
        }                                                              \
    } while (0)

bool DevPlan::upload(std::string& err) {
    if (ctab) return true;
    const size_t nt = h.ctab.size();
    std::vector<float> t32(nt);
    for (size_t i = 0; i < nt; ++i) t32[i] = (float)h.ctab[i];
    tsize = (int)h.tdx.size();
    const size_t tb = (size_t)(tsize > 0 ? tsize : 1) * sizeof(int);
    if (hipMalloc(&ctab, nt * sizeof(double)) != hipSuccess || hipMalloc(&ctab32, nt * sizeof(float)) != hipSuccess ||
        hipMalloc(&tdx, tb) != hipSuccess || hipMalloc(&tdy, tb) != hipSuccess || hipMalloc(&tdz, tb) != hipSuccess) {
        err = "hipMalloc failed for the plan tables";
        release();
        return false;
    }
    tpad = (tsize + TMPL_U - 1) / TMPL_U * TMPL_U + TMPL_U;
    if (hipMalloc(&tdelta, (size_t)tpad * 4) != hipSuccess || hipMalloc(&tdelta_lds, (size_t)tpad * 4) != hipSuccess ||
        hipMalloc(&tpack, (size_t)tpad * 4) != hipSuccess) {
        err = "hipMalloc failed for the template deltas";
        release();
        return false;
    }
    tkey[0] = tkey[1] = tkey[2] = tkey[3] = -1;
    if (h.octant) {
        const size_t nw = h.otab_w.size();
        if (hipMalloc(&otab_w, nw * sizeof(double)) != hipSuccess || hipMalloc(&rank_w, nw * sizeof(unsigned)) != hipSuccess ||
            hipMemcpy(otab_w, h.otab_w.data(), nw * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
            hipMemcpy(rank_w, h.rank_w.data(), nw * sizeof(unsigned), hipMemcpyHostToDevice) != hipSuccess) {
            err = "hipMalloc/hipMemcpy failed for the octant tables";
            release();
            return false;
        }
    }
    if (h.octant && !h.otab.empty()) {
        if (hipMalloc(&otab, h.otab.size() * sizeof(double)) != hipSuccess ||
            hipMemcpy(otab, h.otab.data(), h.otab.size() * sizeof(double), hipMemcpyHostToDevice) != hipSuccess) {
            err = "hipMalloc/hipMemcpy failed for the octant table";
            release();
            return false;
        }
    }
    if (hipMemcpy(ctab, h.ctab.data(), nt * sizeof(double), hipMemcpyHostToDevice) != hipSuccess ||
        hipMemcpy(ctab32, t32.data(), nt * sizeof(float), hipMemcpyHostToDevice) != hipSuccess ||
        (tsize > 0 && (hipMemcpy(tdx, h.tdx.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemcpy(tdy, h.tdy.data(), tb, hipMemcpyHostToDevice) != hipSuccess ||
                       hipMemcpy(tdz, h.tdz.data(), tb, hipMemcpyHostToDevice) != hipSuccess))) {
        err = "hipMemcpy failed for the plan tables";
        release();
        return false;
    }
    return true;
}

bool DevPlan::place(char* base, size_t cap, size_t& used, hipStream_t st, std::string& err) {
    const size_t nt = h.ctab.size();
    ctab32_host.resize(nt);
    for (size_t i = 0; i < nt; ++i) ctab32_host[i] = (float)h.ctab[i];
    tsize = (int)h.tdx.size();
    const size_t tb = (size_t)(tsize > 0 ? tsize : 1) * sizeof(int);
    tpad = (tsize + TMPL_U - 1) / TMPL_U * TMPL_U + TMPL_U;
    const size_t nw = h.octant ? h.otab_w.size() : 0;
    const size_t no = h.octant ? h.otab.size() : 0;
    size_t off = 0;
    auto take = [&](size_t bytes) -> char* {
        char* p = base + off;
        off = (off + bytes + 255) / 256 * 256;
        return p;
    };
    ctab = (double*)take(nt * sizeof(double));
    ctab32 = (float*)take(nt * sizeof(float));
    tdx = (int*)take(tb);
    tdy = (int*)take(tb);
    tdz = (int*)take(tb);
    tdelta = (int*)take((size_t)tpad * 4);
    tdelta_lds = (int*)take((size_t)tpad * 4);
    tpack = (int*)take((size_t)tpad * 4);
    otab_w = nw ? (double*)take(nw * sizeof(double)) : nullptr;
    rank_w = nw ? (unsigned*)take(nw * sizeof(unsigned)) : nullptr;
    otab = no ? (double*)take(no * sizeof(double)) : nullptr;
    tkey[0] = tkey[1] = tkey[2] = tkey[3] = -1;
    if (off > cap) {
        err = "the allocation-free scratch is too small for this parameter set's tables (" + std::to_string(off) +
              " bytes, scratch " + std::to_string(cap) + ")";
        ctab = nullptr;
        ctab32 = nullptr;
        tdx = tdy = tdz = tdelta = tdelta_lds = tpack = nullptr;
        otab_w = otab = nullptr;
        rank_w = nullptr;
        return false;
    }
    used = off;
    auto put = [&](void* dst, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpyAsync(dst, src, bytes, hipMemcpyHostToDevice, st) == hipSuccess;
    };
    if (!put(ctab, h.ctab.data(), nt * sizeof(double)) || !put(ctab32, ctab32_host.data(), nt * sizeof(float)) ||
        (tsize > 0 && (!put(tdx, h.tdx.data(), tb) || !put(tdy, h.tdy.data(), tb) || !put(tdz, h.tdz.data(), tb))) ||
        !put(otab_w, h.otab_w.data(), nw * sizeof(double)) || !put(rank_w, h.rank_w.data(), nw * sizeof(unsigned)) ||
        !put(otab, h.otab.data(), no * sizeof(double))) {
        err = "hipMemcpyAsync failed for the plan tables";
        return false;
    }
    return true;
}

void DevPlan::release() {
    if (ctab) (void)hipFree(ctab);
    if (ctab32) (void)hipFree(ctab32);
    if (tdx) (void)hipFree(tdx);
    if (tdy) (void)hipFree(tdy);
    if (tdz) (void)hipFree(tdz);
    if (otab) (void)hipFree(otab);
    otab = nullptr;
    if (otab_w) (void)hipFree(otab_w);
    if (rank_w) (void)hipFree(rank_w);
    otab_w = nullptr;
    rank_w = nullptr;
    if (tdelta) (void)hipFree(tdelta);
    if (tdelta_lds) (void)hipFree(tdelta_lds);
    if (tpack) (void)hipFree(tpack);
    tdelta = tdelta_lds = tpack = nullptr;
    ctab = nullptr;
    ctab32 = nullptr;
    tdx = tdy = tdz = nullptr;
}

// ----------------------------------------------------------- work list
// The solve's work list (solve.h): the cells to krige ordered by search path
// (0 template walk, 1 cluster search; every cell is path 1 when the solve
// has no LDS table), count bucket (25-32, 17-24, 9-16, 0-8 neighbours) and
// count, descending.  Bin = 33 path + 32 - n.  Informed OK/SK cells are not
// listed: the counting pass writes their outputs (the data value, mask 1,
// variance 0).
__device__ __forceinline__ int work_bin(int cnt, int all_global) {
    const bool cl = all_global || cnt < CNT_INFORMED || (cnt >= 0 && (cnt & CNT_CLUSTER));
    return (cl ? 33 : 0) + 32 - count_n(cnt);
}

// Wave-aggregated add of one to h[bin] (bin < 0: none): one atomic per
// distinct bin of the wave, which also returns each lane's rank among the
// wave's lanes of its bin added before it (base from the atomic + lanes of
// lower index).  Every lane of the wave must call it.
__device__ __forceinline__ int wave_bin_add(int* h, int bin) {
    const int lane = threadIdx.x & 63;
    unsigned long long todo = __ballot(bin >= 0);
    int rank = 0;
    while (todo) {
        const int leader = __ffsll((unsigned long long)todo) - 1;
        const int lb = __shfl(bin, leader);
        const unsigned long long same = __ballot(bin == lb) & todo;
        int wbase = 0;
        if (lane == leader) wbase = atomicAdd(&h[lb], __popcll(same));
        wbase = __shfl(wbase, leader);
        if ((same >> lane) & 1ull) rank = wbase + __popcll(same & ((1ull << lane) - 1ull));
        todo &= ~same;
    }
    return rank;
}

// The count and scatter passes run BUCKET_BLOCKS blocks, block b over the
// same contiguous range of cells, so the list is a counting sort without
// global atomics: the count pass writes each block's bin counts, the scan
// turns them into each block's place in each bin, the scatter writes.
constexpr int BUCKET_BLOCKS = 1024;
constexpr int BUCKET_UNROLL = 4;  // tiles of 256 cells whose loads issue together

__device__ __forceinline__ void bucket_range(long long n, long long& lo, long long& hi) {
    const long long per = ((n + BUCKET_BLOCKS - 1) / BUCKET_BLOCKS + 255) / 256 * 256;
    lo = (long long)blockIdx.x * per;
    hi = lo + per < n ? lo + per : n;
}

__global__ __launch_bounds__(256) void k_bucket_count(SolveArgs a, int all_global, int* hist) {
    __shared__ int h[NBINS];
    for (int i = threadIdx.x; i < NBINS; i += 256) h[i] = 0;
    __syncthreads();
    long long lo, hi;
    bucket_range(a.g.n, lo, hi);
    for (long long base = lo; base < hi; base += 256 * BUCKET_UNROLL) {
        int cnt[BUCKET_UNROLL];
#pragma unroll
        for (int k = 0; k < BUCKET_UNROLL; ++k) {
            const long long cell = base + 256 * k + threadIdx.x;
            cnt[k] = cell < hi ? a.count[cell] : CNT_INFORMED;
        }
#pragma unroll
        for (int k = 0; k < BUCKET_UNROLL; ++k) {
            const long long cell = base + 256 * k + threadIdx.x;
            int bin = -1;
            if (cell < hi) {
                if (a.kind != 2 && cnt[k] == CNT_INFORMED) {
                    const float v = a.values[cell];
                    if (a.out_val) a.out_val[cell] = v;
                    if (a.out_mask) a.out_mask[cell] = 1;
                    if (a.out_var) a.out_var[cell] = 0.0;
                    if (a.out_est64) a.out_est64[cell] = v;
                    if (a.out_status) a.out_status[cell] = ST_INFORMED;
                } else {
                    bin = work_bin(cnt[k], all_global);
                }
            }
            (void)wave_bin_add(h, bin);
        }
    }
    __syncthreads();
    for (int i = threadIdx.x; i < NBINS; i += 256) hist[blockIdx.x * NBINS + i] = h[i];
}

// One block per bin: exclusive prefix of the bin's column of block counts,
// written back in place, and the bin's total.
__global__ __launch_bounds__(256) void k_bucket_scan(int* hist, int* tot) {
    __shared__ int wsum[4];
    constexpr int PER = BUCKET_BLOCKS / 256;
    const int bin = blockIdx.x, tid = threadIdx.x, lane = tid & 63, w = tid >> 6;
    int v[PER];
    int local = 0;
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        v[k] = hist[(tid * PER + k) * NBINS + bin];
        local += v[k];
    }
    int x = local;  // inclusive scan over the block
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const int y = __shfl_up(x, d);
        if (lane >= d) x += y;
    }
    if (lane == 63) wsum[w] = x;
    __syncthreads();
    int before = 0;
#pragma unroll
    for (int k = 0; k < 4; ++k) before += k < w ? wsum[k] : 0;
    int run = before + x - local;  // exclusive
#pragma unroll
    for (int k = 0; k < PER; ++k) {
        hist[(tid * PER + k) * NBINS + bin] = run;
        run += v[k];
    }
    if (tid == 255) tot[bin] = run;
}

// Lays the bins out from their totals (each bucket starting at a multiple of
// SOLVE_CHUNK; block 0 also writes the segment table and the gaps, -1) and
// writes the block's cells at its places.
__global__ __launch_bounds__(256) void k_bucket_scatter(SolveArgs a, int all_global, const int* hist, const int* tot,
                                                        int* seg, int* sorted) {
    __shared__ int tot_s[NBINS], off[NBINS], loc[NBINS];
    const int tid = threadIdx.x;
    if (tid < NBINS) {
        tot_s[tid] = tot[tid];
        off[tid] = hist[blockIdx.x * NBINS + tid];
        loc[tid] = 0;
    }
    __syncthreads();
    if (tid == 0) {
        int pos = 0;
        for (int p = 0; p < 2; ++p)
            for (int b = 3; b >= 0; --b) {
                const int begin = (pos + SOLVE_CHUNK - 1) / SOLVE_CHUNK * SOLVE_CHUNK;
                if (blockIdx.x == 0)
                    for (int i = pos; i < begin; ++i) sorted[i] = -1;
                pos = begin;
                const int nhi = 8 * b + 8, nlo = b == 0 ? 0 : 8 * b + 1;
                for (int n = nhi; n >= nlo; --n) {
                    const int bin = 33 * p + 32 - n;
                    off[bin] += pos;
                    pos += tot_s[bin];
                }
                if (blockIdx.x == 0) {
                    seg[2 * (4 * p + b)] = begin;
                    seg[2 * (4 * p + b) + 1] = pos;
                }
            }
    }
    __syncthreads();
    long long lo, hi;
    bucket_range(a.g.n, lo, hi);
    for (long long base = lo; base < hi; base += 256 * BUCKET_UNROLL) {
        int cnt[BUCKET_UNROLL];
#pragma unroll
        for (int k = 0; k < BUCKET_UNROLL; ++k) {
            const long long cell = base + 256 * k + tid;
            cnt[k] = cell < hi ? a.count[cell] : CNT_INFORMED;
        }
#pragma unroll
        for (int k = 0; k < BUCKET_UNROLL; ++k) {
            const long long cell = base + 256 * k + tid;
            int bin = -1;
            if (cell < hi && !(a.kind != 2 && cnt[k] == CNT_INFORMED)) bin = work_bin(cnt[k], all_global);
            const int rank = wave_bin_add(loc, bin);
            if (bin >= 0) sorted[off[bin] + rank] = (int)cell;
        }
    }
}

bool solve_sorted(int top, int kind, const SolveArgs& sa, const SortedLaunch& l, hipStream_t st) {
    switch (top * 4 + kind) {
        case 8 * 4 + 0: return solve_sorted_8_0(sa, l, st);
        case 8 * 4 + 1: return solve_sorted_8_1(sa, l, st);
        case 8 * 4 + 2: return solve_sorted_8_2(sa, l, st);
        case 16 * 4 + 0: return solve_sorted_16_0(sa, l, st);
        case 16 * 4 + 1: return solve_sorted_16_1(sa, l, st);
        case 16 * 4 + 2: return solve_sorted_16_2(sa, l, st);
        case 24 * 4 + 0: return solve_sorted_24_0(sa, l, st);
        case 24 * 4 + 1: return solve_sorted_24_1(sa, l, st);
        case 24 * 4 + 2: return solve_sorted_24_2(sa, l, st);
        case 32 * 4 + 0: return solve_sorted_32_0(sa, l, st);
        case 32 * 4 + 1: return solve_sorted_32_1(sa, l, st);
        default: return solve_sorted_32_2(sa, l, st);
    }
}

void solve_fallback(bool b32, int kind, const SolveArgs& sa, long long max_cells, hipStream_t st) {
    if (b32) {
        if (kind == 0) solve_fallback_b32_0(sa, max_cells, st);
        else if (kind == 1) solve_fallback_b32_1(sa, max_cells, st);
        else solve_fallback_b32_2(sa, max_cells, st);
    } else {
        if (kind == 0) solve_fallback_b24_0(sa, max_cells, st);
        else if (kind == 1) solve_fallback_b24_1(sa, max_cells, st);
        else solve_fallback_b24_2(sa, max_cells, st);
    }
}

void* Workspace::get(const std::string& name, size_t bytes, std::string& err) {
    if (bytes == 0) bytes = 256;
    auto it = bufs.find(name);
    if (it != bufs.end() && it->second.second >= bytes) return it->second.first;
    if (arena) {
        // Buffers keep their place between calls with the same layout, so
        // state recorded in tags (keyed by address and geometry) stays valid.
        const size_t off = (arena_used + 255) / 256 * 256;
        if (off + bytes > arena_cap) {
            if (err.empty())
                err = "the allocation-free scratch is too small for this call: buffer '" + name + "' needs " +
                      std::to_string(bytes) + " bytes at offset " + std::to_string(off) + " of " +
                      std::to_string(arena_cap);
            return nullptr;
        }
        arena_used = off + bytes;
        bufs[name] = std::make_pair((void*)(arena + off), bytes);
        return arena + off;
    }
    if (it != bufs.end()) {
        // The previous contents are not needed; callers wait for prior work
        // before the launcher runs, so freeing here is safe.
        (void)hipDeviceSynchronize();
        (void)hipFree(it->second.first);
        bufs.erase(it);
    }
    void* p = nullptr;
    if (hipMalloc(&p, bytes) != hipSuccess) {
        err = "hipMalloc failed for workspace buffer '" + name + "' (" + std::to_string(bytes) + " bytes)";
        return nullptr;
    }
    ++gen;
    bufs[name] = std::make_pair(p, bytes);
    tags.erase(name);
    return p;
}

void Workspace::drop(const std::string& name) {
    auto it = bufs.find(name);
    if (it == bufs.end() || arena) return;
    (void)hipFree(it->second.first);
    bufs.erase(it);
    tags.erase(name);
    ++gen;
}

void Workspace::release() {
    ++gen;
    for (auto& kv : bufs) (void)hipFree(kv.second.first);
    bufs.clear();
    tags.clear();
}

namespace {

unsigned blocks_for(long long n, int block) { return (unsigned)((n + block - 1) / block); }

// Raises a kernel's dynamic-LDS limit (once per device and size).
hipError_t allow_lds(const void* fn, int bytes) {
    static std::mutex mu;
    static std::map<std::pair<const void*, int>, int> done;  // (kernel, device) -> bytes allowed
    int dev = 0;
    hipError_t e = hipGetDevice(&dev);
    if (e != hipSuccess) return e;
    std::lock_guard<std::mutex> lk(mu);
    int& have = done[{fn, dev}];
    if (have >= bytes) return hipSuccess;
    e = hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
    if (e == hipSuccess) have = bytes;
    return e;
}

}  // namespace

int launch_grid(DevPlan& plan, Workspace& ws, const GridCall& c, hipStream_t st, CallTiming* timing,
                std::string& err) {
    const PlanHost& h = plan.h;
    if (!plan.upload(err)) return -1;
    if (c.nx < 1 || c.ny < 1 || c.nz < 1) {
        err = "grid dimensions must be positive";
        return -1;
    }
    const long long n = (long long)c.nx * c.ny * c.nz;
    if (n > 2147483647LL) {
        err = "grid volume exceeds INT_MAX";
        return -1;
    }
    if (h.r[0] < 1 || h.r[1] < 1 || h.r[2] < 1) {
        // the reference's clusterizer divides the grid by them
        // (clusterizer.cpp:87-89)
        err = "search radiuses must be >= 1 for kriging";
        return -1;
    }
    if (c.kind == KIND_MIK && (c.nind < 1 || c.nind > 16)) {
        err = "median IK needs between 1 and 16 thresholds";
        return -1;
    }
    if (c.cats_host && (c.kind != KIND_MIK || !c.ind || c.zfield)) {
        err = "indicator categories need indicator data";
        return -1;
    }
    if (c.lvm_mean && c.kind != KIND_SK) {
        err = "a locally varying mean applies to simple kriging only";
        return -1;
    }
    GridDims g;
    g.nx = c.nx;
    g.ny = c.ny;
    g.nz = c.nz;
    g.n = n;
    g.rx = h.r[0];
    g.ry = h.r[1];
    g.rz = h.r[2];
    g.hx = h.H[0];
    g.hy = h.H[1];
    g.hz = h.H[2];
    g.px = (c.nx + 2 * g.hx + 15) / 16 * 16;  // row pitch: 16-byte aligned rows
    g.py = c.ny + 2 * g.hy;
    g.pz = c.nz + 2 * g.hz;
    // clusterizer.cpp:87-89: g / r + 2, in double, truncated.
    g.cx = (int)(c.nx / (double)g.rx + 2);
    g.cy = (int)(c.ny / (double)g.ry + 2);
    g.cz = (int)(c.nz / (double)g.rz + 2);
    const long long P = (long long)g.px * g.py * g.pz;
    const int ncl = g.cx * g.cy * g.cz;
    // A cell has at most n neighbours (a caller's neighbour buffer keeps the
    // stride it was given); above 32 the wide solver runs.
    const int K = (int)(c.out_nboff ? (long long)h.kmax : std::min<long long>(h.kmax, n));
    if (K > 65535) {
        err = "max_neighbours above 65535 is not supported";
        return -1;
    }
    const int T = plan.tsize;

    auto* pmask = (unsigned char*)ws.get("pmask", (size_t)P, err);
    auto* cl_count = (int*)ws.get("cl_count", (size_t)ncl * 4, err);
    auto* nbhd = (unsigned char*)ws.get("cl_nbhd", (size_t)ncl, err);
    // Node lists: a fixed slot of `limit` entries per cluster.
    const int node_cap = (int)std::max<long long>(1, std::min<long long>(h.limit, n));
    auto* cl_nodes = (int*)ws.get("cl_nodes", (size_t)ncl * node_cap * 4, err);
    // Tiled search: LDS = mask box (tile + 2r halo) + the template's first
    // SEARCH_HEAD packed offsets (+ the staged hits).  The slices per block
    // are the count that keeps the most waves resident on a CU (160 KiB of
    // LDS, allocated in 512-byte units; at most 32 waves), the smaller count
    // on ties.
    const int lx = (TILE_X + 2 * g.hx + 3) / 4 * 4, ly = TILE_Y + 2 * g.hy;
    const bool staged_ok = K <= 32 && T < 65536 && c.search_variant != 2;
    auto pick_slabs = [&](bool stg) {
        int best_waves = 0, ns = 0;
        for (int s = 1; s <= SEARCH_MAX_SLABS; ++s) {
            if (c.search_slabs > 0 && s != c.search_slabs) continue;
            const size_t b = search_tile_lds(lx, ly, g.hz, s, K, stg);
            if (b > kSearchLdsMax) continue;
            const int blocks = std::min<int>((int)(kSearchLdsMax / ((b + 511) / 512 * 512)), 32 / (4 * s));
            if (blocks * 4 * s > best_waves) {
                best_waves = blocks * 4 * s;
                ns = s;
            }
        }
        return ns;
    };
    int nslab = 0;
    bool staged = false;
    if (lx / 4 <= 256 && c.search_variant != 1) {
        if (staged_ok) nslab = pick_slabs(true);
        staged = nslab > 0;
        if (!staged) nslab = pick_slabs(false);
    }
    const bool tiled = nslab > 0;
    const int lz = (tiled ? nslab : 1) + 2 * g.hz;
    const size_t tile_lds = tiled ? search_tile_lds(lx, ly, g.hz, nslab, K, staged) : 0;
    auto* count = c.out_count ? c.out_count : (int*)ws.get("count", (size_t)n * 4, err);
    auto* nboff = c.out_nboff ? c.out_nboff : (int*)ws.get("nboff", (size_t)n * K * 4, err);
    auto* cl_cells = (int*)ws.get("cl_cells", (size_t)n * 4, err);
    auto* counters = (int*)ws.get("counters", CTR_COUNT * 4, err);
    auto* sorted = (int*)ws.get("sorted", ((size_t)n + 16 * SOLVE_CHUNK) * 4, err);
    auto* bucket_hist = (int*)ws.get("bucket_hist", (size_t)BUCKET_BLOCKS * NBINS * 4, err);
    const int mean_blocks = 1024;
    auto* psum = (double*)ws.get("mean_psum", mean_blocks * 8, err);
    auto* pcnt = (int*)ws.get("mean_pcnt", mean_blocks * 4, err);
    auto* dmean = (double*)ws.get("mean", 8, err);
    // Cluster-path scratch: every candidate comes from one of 27 clusters
    // holding at most `limit` nodes each.
    const int cap = (int)(27 * (h.limit > 0 ? h.limit : 1));
    // Octant-rank cluster search: packed coordinates (11/11/10 bits), the
    // sort word rank << qbits | candidate index, and LDS for the buffers.
    int qbits = 1;
    while ((1LL << qbits) < 27LL * node_cap) ++qbits;
    const int cap1 = c.cluster_cap > 0 && c.cluster_cap < kClusterSortCap ? c.cluster_cap : kClusterSortCap;
    const int cand_cap = (int)std::min<long long>(
        27LL * node_cap, c.cluster_cand > 0 && c.cluster_cand < kClusterCandCap ? c.cluster_cand : kClusterCandCap);
    const size_t cl_lds = ((size_t)cap1 * CL_BLOCK + (size_t)cand_cap) * 4;
    const int cap2 = c.cluster_cap > 0 && c.cluster_cap < kClusterListCap ? c.cluster_cap : kClusterListCap;
    const bool oct_cluster = plan.rank_w && c.solve_variant != 1 && c.nx <= 2048 && c.ny <= 2048 && c.nz <= 1024 &&
                             qbits < 32 && (unsigned long long)h.rank_max < (1ULL << (32 - qbits)) &&
                             cl_lds <= kSearchLdsMax;
    // The full-table search takes what the octant form leaves (its overflow
    // list), or every cluster-path cell without octant tables.
    const int nlanes = c.cluster_lanes > 0 ? c.cluster_lanes : (oct_cluster ? 8192 : 32768);
    auto* skey = (double*)ws.get("skey", (size_t)cap * nlanes * 8, err);
    auto* soff = (int*)ws.get("soff", (size_t)cap * nlanes * 4, err);
    if (!err.empty()) return -1;
    int* n_cl_cells = counters + CTR_N_CL_CELLS;
    int* overflow = counters + CTR_OVERFLOW;
    int* n_fallback = counters + CTR_N_FALLBACK;
    int* n_ovf = counters + CTR_N_OVF;
    int* n_ovf2 = counters + CTR_N_OVF2;
    auto* ovf_cells = (int*)ws.get("ovf_cells", (size_t)n * 4, err);
    auto* ovf2_cells = (int*)ws.get("ovf2_cells", (size_t)n * 4, err);
    auto* fb_cells = (int*)ws.get("fb_cells", (size_t)n * 4, err);
    if (!err.empty()) return -1;

    // Profiling: start/stop events of the search and of the main solve
    // kernel, recorded by the kernel dispatches themselves.
    hipEvent_t ev_t[4] = {nullptr, nullptr, nullptr, nullptr};
    if (timing && timing->valid)
        for (int i = 0; i < 4; ++i) ev_t[i] = timing->ev[i];
    const int B = 256;
    // The padded mask's halo stays zero between calls that share its buffer
    // and geometry; only then is the whole buffer cleared.
    const std::string pm_sig = std::to_string((uintptr_t)pmask) + ":" + std::to_string(g.px) + ":" +
                               std::to_string(g.py) + ":" + std::to_string(g.pz) + ":" + std::to_string(g.hx) + ":" +
                               std::to_string(g.hy) + ":" + std::to_string(g.hz);
    if (ws.tags["pmask"] != pm_sig) {
        hipLaunchKernelGGL(k_fill_u8, dim3(std::min<long long>(4096, blocks_for(P, B))), dim3(B), 0, st, pmask, P,
                           (unsigned char)0);
        ws.tags["pmask"] = pm_sig;
    }
    hipLaunchKernelGGL(k_cluster_build, dim3(ncl), dim3(256), 0, st, g, c.mask, pmask, cl_count, cl_nodes, node_cap,
                       counters, (int)CTR_COUNT);
    hipLaunchKernelGGL(k_cluster_flags, dim3(blocks_for(ncl, B)), dim3(B), 0, st, g, (const int*)cl_count,
                       (long long)h.limit, nbhd);
    // Template deltas for this grid's padding and tile box, kept by the plan.
    if (plan.tkey[0] != g.px || plan.tkey[1] != g.py || plan.tkey[2] != lx || plan.tkey[3] != ly) {
        hipLaunchKernelGGL(k_tdelta, dim3(blocks_for(plan.tpad, B)), dim3(B), 0, st, g, T, plan.tpad,
                           (const int*)plan.tdx, (const int*)plan.tdy, (const int*)plan.tdz, plan.tdelta,
                           plan.tdelta_lds, lx, ly, plan.tpack);
        plan.tkey[0] = g.px;
        plan.tkey[1] = g.py;
        plan.tkey[2] = lx;
        plan.tkey[3] = ly;
    }
    if (c.kind == KIND_SK && c.auto_mean) {
        hipLaunchKernelGGL(k_mean_partial, dim3(mean_blocks), dim3(256), 0, st, n, c.values, c.mask, psum, pcnt);
        hipLaunchKernelGGL(k_mean_final, dim3(1), dim3(64), 0, st, mean_blocks, (const double*)psum,
                           (const int*)pcnt, dmean);
    }

    SearchArgs sa;
    sa.g = g;
    sa.kind = c.kind;
    sa.kmax = K;
    sa.mask = c.mask;
    sa.pmask = pmask;
    sa.tdelta = plan.tdelta;
    sa.tdelta_lds = plan.tdelta_lds;
    sa.lx = lx;
    sa.ly = ly;
    sa.lz = lz;
    sa.tpack = plan.tpack;
    sa.tsize = T;
    sa.cl_nbhd_exceeded = nbhd;
    sa.count = count;
    sa.nboff = nboff;
    sa.cl_cells = cl_cells;
    sa.n_cl_cells = n_cl_cells;
    // With the octant cluster search the cells of within-limit clusters are
    // its own (it classifies and counts them); the full-table form takes the list.
    sa.cluster_list = oct_cluster ? 0 : 1;
    if (tiled) {
        const void* fn = staged ? (const void*)&k_search_tile<true> : (const void*)&k_search_tile<false>;
        if (tile_lds > 65536) HPGLX_CHECK(allow_lds(fn, (int)tile_lds));
        const dim3 grid((c.nx + TILE_X - 1) / TILE_X, (c.ny + TILE_Y - 1) / TILE_Y, (c.nz + nslab - 1) / nslab);
        const dim3 block(TILE_CELLS * nslab);
        const int head = c.search_head > 0 && c.search_head < SEARCH_HEAD ? c.search_head : SEARCH_HEAD;
        if (staged)
            HPGLX_LAUNCH_EV(k_search_tile<true>, grid, block, (uint32_t)tile_lds, st, ev_t[0], ev_t[1], sa, head);
        else
            HPGLX_LAUNCH_EV(k_search_tile<false>, grid, block, (uint32_t)tile_lds, st, ev_t[0], ev_t[1], sa, head);
    } else {
        HPGLX_LAUNCH_EV(k_search_tmpl, dim3(blocks_for(n, B)), dim3(B), 0, st, ev_t[0], ev_t[1], sa);
    }

    ClusterSearchArgs ca;
    ca.g = g;
    ca.kmax = K;
    ca.cl_cells = cl_cells;
    ca.n_cl_cells = n_cl_cells;
    ca.cl_count = cl_count;
    ca.cl_nodes = cl_nodes;
    ca.node_cap = node_cap;
    ca.limit = h.limit;
    ca.ctab = plan.ctab;
    ca.wx = h.W[0];
    ca.wy = h.W[1];
    ca.wz = h.W[2];
    ca.threshold = h.c0() / 100;
    ca.skey = skey;
    ca.soff = soff;
    ca.capacity = cap;
    ca.nlanes = nlanes;
    ca.count = count;
    ca.nboff = nboff;
    ca.overflow = overflow;
    if (oct_cluster) {
        ClusterOctArgs oa;
        oa.g = g;
        oa.kind = c.kind;
        oa.mask = c.mask;
        oa.n_cl_cells = n_cl_cells;
        oa.kmax = K;
        oa.cl_count = cl_count;
        oa.cl_nodes = cl_nodes;
        oa.node_cap = node_cap;
        oa.nbhd_exceeded = nbhd;
        oa.rank_w = plan.rank_w;
        oa.wx1 = h.W[0] + 1;
        oa.wy1 = h.W[1] + 1;
        oa.wz1 = h.W[2] + 1;
        oa.qbits = qbits;
        oa.cap1 = cap1;
        oa.cand_cap = cand_cap;
        oa.count = count;
        oa.nboff = nboff;
        oa.ovf_cells = ovf_cells;
        oa.n_ovf = n_ovf;
        // Candidates come from the 27 clusters around a cell (|d| <= 2r - 1);
        // only a radius beyond the covariance's reach puts some past the table.
        const bool bounded = 2 * g.rx - 1 > h.W[0] || 2 * g.ry - 1 > h.W[1] || 2 * g.rz - 1 > h.W[2];
        const void* fo = bounded ? (const void*)&k_search_cluster_oct<true> : (const void*)&k_search_cluster_oct<false>;
        if (cl_lds > 65536) HPGLX_CHECK(allow_lds(fo, (int)cl_lds));
        if (bounded)
            hipLaunchKernelGGL(k_search_cluster_oct<true>, dim3((unsigned)std::min(ncl, 2048)), dim3(CL_BLOCK), cl_lds,
                               st, oa);
        else
            hipLaunchKernelGGL(k_search_cluster_oct<false>, dim3((unsigned)std::min(ncl, 2048)), dim3(CL_BLOCK), cl_lds,
                               st, oa);
        // its overflow: one lane per cell with a deeper LDS buffer
        const size_t cll_lds = (size_t)cap2 * CL_LIST_BLOCK * 4;
        const void* fl = bounded ? (const void*)&k_search_cluster_list_oct<true>
                                 : (const void*)&k_search_cluster_list_oct<false>;
        if (cll_lds > 65536) HPGLX_CHECK(allow_lds(fl, (int)cll_lds));
        if (bounded)
            hipLaunchKernelGGL(k_search_cluster_list_oct<true>, dim3(1024), dim3(CL_LIST_BLOCK), cll_lds, st, oa,
                               (const int*)ovf_cells, (const int*)n_ovf, cap2, ovf2_cells, n_ovf2);
        else
            hipLaunchKernelGGL(k_search_cluster_list_oct<false>, dim3(1024), dim3(CL_LIST_BLOCK), cll_lds, st, oa,
                               (const int*)ovf_cells, (const int*)n_ovf, cap2, ovf2_cells, n_ovf2);
        ca.cl_cells = ovf2_cells;
        ca.n_cl_cells = n_ovf2;
    }
    hipLaunchKernelGGL(k_search_cluster, dim3(nlanes / B), dim3(B), 0, st, ca);

    SolveArgs so{};
    so.g = g;
    so.kind = c.kind;
    so.kmax = K;
    so.values = c.values;
    so.mask = c.mask;
    so.ind = c.ind;
    so.zfield = c.zfield;
    so.thresholds = c.thresholds;
    so.nind = c.nind;
    so.mean_dev = (c.kind == KIND_SK && c.auto_mean) ? dmean : nullptr;
    so.mean = c.mean;
    // marginal probabilities and categories travel in the kernel arguments
    for (int k = 0; k < 16; ++k) {
        so.marg[k] = (c.kind == KIND_MIK && k < c.nind) ? c.marg_host[k] : 0.0;
        so.cats[k] = (c.cats_host && k < c.nind) ? c.cats_host[k] : 0;
    }
    so.has_cats = c.cats_host ? 1 : 0;
    so.lvm_mean = c.lvm_mean;
    so.count = count;
    so.nboff = nboff;
    so.ctab = plan.ctab;
    so.ctab32 = plan.ctab32;
    so.wx = h.W[0];
    so.wy = h.W[1];
    so.wz = h.W[2];
    so.otab = plan.otab;
    so.otab_w = c.solve_variant == 1 ? nullptr : plan.otab_w;
    so.ex = h.E[0];
    so.ey = h.E[1];
    so.ez = h.E[2];
    so.c0 = h.c0();
    so.pivot_tol = c.pivot_tol;
    so.refine_steps = c.refine_steps;
    so.out_val = c.out_val;
    so.out_mask = c.out_mask;
    so.out_class = c.out_class;
    so.out_prob = c.out_prob;
    so.out_var = c.out_var;
    so.out_est64 = c.out_est64;
    so.out_status = c.out_status;
    so.order_relations = c.order_relations;
    so.cell_list = nullptr;
    so.n_list = nullptr;
    so.cl_cells = cl_cells;
    so.n_cl_cells = n_cl_cells;
    so.fallback_cells = fb_cells;
    so.n_fallback = n_fallback;
    const bool mixed = c.precision == PREC_MIXED;
    if (K > 32) {
        if (dispatch_solve_wide(so, c.kind, ws, ev_t + 2, st, err) != 0) return -1;
        HPGLX_CHECK(hipGetLastError());
        return 0;
    }
    // The LDS table form needs the packed x/y table addresses (both 16-bit
    // fields below 32768: x up to E_x * 8, y up to E_y * (E_x + 1) * 8).
    const bool pack = (long long)h.E[0] * 8 < 32768 && (long long)h.E[1] * (h.E[0] + 1) * 8 < 32768;
    const int nelem = (plan.otab && c.solve_variant == 0 && pack) ? (int)h.otab.size() : 0;
    // (the table must fit one block's LDS)
    const bool lds = nelem > 0 && (size_t)nelem * sizeof(double) <= kSearchLdsMax;
    const int all_global = lds ? 0 : 1;
    so.sorted = sorted;
    so.seg = counters + CTR_SEG;
    hipLaunchKernelGGL(k_bucket_count, dim3(BUCKET_BLOCKS), dim3(256), 0, st, so, all_global, bucket_hist);
    hipLaunchKernelGGL(k_bucket_scan, dim3(NBINS), dim3(256), 0, st, bucket_hist, counters + CTR_BIN_TOTAL);
    hipLaunchKernelGGL(k_bucket_scatter, dim3(BUCKET_BLOCKS), dim3(256), 0, st, so, all_global,
                       (const int*)bucket_hist, (const int*)(counters + CTR_BIN_TOTAL), counters + CTR_SEG, sorted);
    // Buckets up to `top` run in one launch per path; the 25-32 bucket in
    // its own (4 lanes per cell throughout).  The main kernel (events) is the
    // largest bucket's launch on the main path.
    const int top = K <= 8 ? 8 : (K <= 16 ? 16 : 24);
    const bool b32 = K > 24;
    SortedLaunch sl;
    sl.mixed = mixed;
    sl.blocks_per_cu = c.blocks_per_cu;
    // The launches below cover disjoint ranges of the work list: every one
    // after the first may overlap those before it.
    bool first_solve = true;
    for (int path = lds ? 0 : 1; path < 2; ++path) {
        sl.path = path;
        sl.nelem = path == 0 ? nelem : 0;
        const bool main_path = path == (lds ? 0 : 1);
        sl.ev = (main_path && !b32) ? ev_t + 2 : nullptr;
        sl.any_order = c.overlap != 0 && !first_solve;
        first_solve = false;
        SolveArgs sp = so;
        sp.sched = c.dyn_sched ? counters + CTR_SCHED + 2 * path : nullptr;
        // where the LDS form cannot run, the global tables serve the same range
        if (!solve_sorted(top, c.kind, sp, sl, st)) {
            sl.nelem = 0;
            solve_sorted(top, c.kind, sp, sl, st);
        }
        if (b32) {
            sl.nelem = path == 0 ? nelem : 0;
            sl.ev = main_path ? ev_t + 2 : nullptr;
            sl.any_order = c.overlap != 0;
            sp.sched = c.dyn_sched ? counters + CTR_SCHED + 2 * path + 1 : nullptr;
            if (!solve_sorted(32, c.kind, sp, sl, st)) {
                sl.nelem = 0;
                solve_sorted(32, c.kind, sp, sl, st);
            }
        }
    }
    if (mixed) {
        SolveArgs fb = so;
        fb.cell_list = fb_cells;
        fb.n_list = n_fallback;
        fb.fallback_cells = nullptr;
        fb.n_fallback = nullptr;
        solve_fallback(false, c.kind, fb, n, st);
        if (b32) solve_fallback(true, c.kind, fb, n, st);
    }
    HPGLX_CHECK(hipGetLastError());
    return 0;
}

}  // namespace hpglx

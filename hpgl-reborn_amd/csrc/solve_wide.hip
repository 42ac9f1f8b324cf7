// solve_wide.hip -- kriging systems of more than 32 neighbours.
//
// The register solvers hold a cell's system in the registers of 4 or 8
// lanes, which bounds them at 32 neighbours.  The reference takes any
// max_neighbours (its edge tests krige with 1000 on a 10x10x5 grid), so
// larger systems run here: up to 64 neighbours one wavefront per cell with a
// row per lane (solve_wave.h), beyond that one workgroup per cell, the system
// in LDS (up to 141 neighbours) or in a global slot per workgroup, an FP64 Cholesky
// factor column by column (my_kriging_weights.h:60-196 SK, :198-361 OK,
// dpotrf/dpotrs), and the combine, failure policy and outputs of the
// register path (combiner.h:10-25, cont_kriging.h:73-105, median_ik.cpp:60-78,
// indicator_kriging.h:34-59), the float sums taken in neighbour order by one
// lane.
#include <hip/hip_ext.h>

#include <algorithm>
#include <string>

#include "krige_kernels.h"
#include "krige_launch.h"
#include "solve.h"
#include "solve_dispatch.h"
#include "solve_dispatch_impl.h"
#include "solve_wave.h"

namespace hpglx {
namespace {

constexpr int WB = 256;                                  // lanes per cell
constexpr long long kWideLdsBytes = 160 * 1024;          // working sets up to this size live in LDS
constexpr long long kWideScratchBytes = 2LL << 30;       // global slots of all workgroups together

// Working set of one cell with room for k neighbours, in doubles: the lower
// triangle of A row-major with row stride k, b, s, o, then two int arrays
// (table offsets and node deltas) in k doubles.
__host__ __device__ inline long long wide_slot(int k) { return (long long)k * k + 4LL * k; }

template <int KIND>
__global__ __launch_bounds__(WB) void k_solve_wide(SolveArgs a, double* gscratch, int kcap) {
    extern __shared__ double wlds[];
    const int tid = threadIdx.x;
    double* A = gscratch ? gscratch + (long long)blockIdx.x * wide_slot(kcap) : wlds;
    double* bv = A + (long long)kcap * kcap;
    double* sv = bv + kcap;
    double* ov = sv + kcap;
    int* tr = (int*)(ov + kcap);
    int* dn = tr + kcap;
    const GridDims& g = a.g;
    const int SX = 2 * a.wx + 1, SY = 2 * a.wy + 1;
    const int center = (a.wz * SY + a.wy) * SX + a.wx;
    constexpr bool OKK = KIND == 0;
    for (long long cell = blockIdx.x; cell < g.n; cell += gridDim.x) {
        __syncthreads();  // the previous cell's working set is free
        const int cnt = a.count[cell];
        if (KIND != 2 && cnt == CNT_INFORMED) {
            write_informed<1>(a, cell, tid);
            continue;
        }
        const int n = cnt < 0 ? 0 : (cnt & CNT_MASK);
        const int* nb = a.nboff + cell * a.kmax;
        for (int i = tid; i < n; i += WB) {
            int dx, dy, dz;
            unpack_off(nb[i], dx, dy, dz);
            tr[i] = (dz * SY + dy) * SX + dx;
            dn[i] = (dz * g.ny + dy) * g.nx + dx;
        }
        __syncthreads();
        for (int i = tid; i < n; i += WB) {
            bv[i] = a.ctab[center + tr[i]];
            sv[i] = bv[i];
            ov[i] = 1.0;
        }
        for (long long idx = tid; idx < (long long)n * n; idx += WB) {
            const int i = (int)(idx / n), k = (int)(idx - (long long)i * n);
            if (k <= i) A[(long long)i * n + k] = a.ctab[center + tr[i] - tr[k]];
        }
        __syncthreads();
        int status = n == 0 ? ST_NO_NEIGHBOURS : ST_KRIGED;
        // A = L L' in place (lower triangle); fails, as dpotrf, on a pivot
        // that is not positive.
        for (int j = 0; j < n && status == ST_KRIGED; ++j) {
            const double piv = A[(long long)j * n + j];
            __syncthreads();  // every lane has read the pivot
            if (!(piv > 0.0)) {
                status = ST_SINGULAR;
                break;
            }
            const double d = sqrt(piv);
            for (int i = j + 1 + tid; i < n; i += WB) A[(long long)i * n + j] /= d;
            if (tid == 0) A[(long long)j * n + j] = d;
            __syncthreads();
            const int m = n - j - 1;
            for (long long idx = tid; idx < (long long)m * m; idx += WB) {
                const int ii = (int)(idx / m), kk = (int)(idx - (long long)ii * m);
                if (kk <= ii) {
                    const long long i = j + 1 + ii, k = j + 1 + kk;
                    A[i * n + k] -= A[i * n + j] * A[k * n + j];
                }
            }
            __syncthreads();
        }
        if (status == ST_KRIGED) {
            // L y = b (and L u = 1 for OK), then L' x = y
            for (int j = 0; j < n; ++j) {
                if (tid == 0) {
                    const double d = A[(long long)j * n + j];
                    sv[j] /= d;
                    if (OKK) ov[j] /= d;
                }
                __syncthreads();
                for (int i = j + 1 + tid; i < n; i += WB) {
                    const double l = A[(long long)i * n + j];
                    sv[i] -= l * sv[j];
                    if (OKK) ov[i] -= l * ov[j];
                }
                __syncthreads();
            }
            for (int j = n - 1; j >= 0; --j) {
                if (tid == 0) {
                    const double d = A[(long long)j * n + j];
                    sv[j] /= d;
                    if (OKK) ov[j] /= d;
                }
                __syncthreads();
                for (int i = tid; i < j; i += WB) {
                    const double l = A[(long long)j * n + i];
                    sv[i] -= l * sv[j];
                    if (OKK) ov[i] -= l * ov[j];
                }
                __syncthreads();
            }
        }
        if (tid != 0) continue;
        // Weights, then the register path's outputs, on lane 0.
        double mu = 0.0;
        if (status == ST_KRIGED && OKK) {
            double ss = 0.0, so = 0.0;
            for (int i = 0; i < n; ++i) {
                ss += sv[i];
                so += ov[i];
            }
            if (fabs(so) < 1e-15) {
                status = ST_SINGULAR;
            } else {
                mu = (ss - 1) / so;
                for (int i = 0; i < n; ++i) sv[i] -= mu * ov[i];
            }
        }
        if (status != ST_KRIGED) {
            for (int i = 0; i < n; ++i) sv[i] = 0.0;
            mu = 0.0;
        }
        double wb = 0.0;
        for (int i = 0; i < n; ++i) wb += sv[i] * bv[i];
        if (KIND != 2) {
            const double mean = (KIND == 1) ? (a.mean_dev ? a.mean_dev[0] : a.mean) : 0.0;
            const bool lvm = KIND == 1 && a.lvm_mean != nullptr;
            const float mf = lvm ? a.lvm_mean[cell] : (float)mean;
            float acc = mf;
            double pe = 0.0;
            for (int i = 0; i < n; ++i) {
                const long long node = cell + dn[i];
                const float v = a.values[node];
                const float ms = lvm ? a.lvm_mean[node] : mf;
                acc += (float)(sv[i] * (double)(v - ms));
                pe += sv[i] * ((double)v - (double)ms);
            }
            const double var = a.c0 - wb - mu;
            if (status == ST_KRIGED) {
                if (a.out_val) a.out_val[cell] = acc;
                if (a.out_mask) a.out_mask[cell] = 1;
                if (a.out_var) a.out_var[cell] = var < 0 ? 0.0 : var;
                if (a.out_est64) a.out_est64[cell] = (double)mf + pe;
            } else if (KIND == 1) {
                if (a.out_val) a.out_val[cell] = mf;
                if (a.out_mask) a.out_mask[cell] = 1;
                if (a.out_var) a.out_var[cell] = -1.0;
                if (a.out_est64) a.out_est64[cell] = (double)mf;
            } else {
                if (a.out_val) a.out_val[cell] = a.values[cell];
                if (a.out_mask) a.out_mask[cell] = a.mask[cell];
                if (a.out_var) a.out_var[cell] = -1.0;
                if (a.out_est64) a.out_est64[cell] = __builtin_nan("");
            }
            if (a.out_status) a.out_status[cell] = (unsigned char)status;
            continue;
        }
        for (int k = 0; k < a.nind; ++k) {
            const float pf = (float)a.marg[k];
            const float thr = a.zfield ? a.thresholds[k] : 0.f;
            const int cat = a.has_cats ? a.cats[k] : -1;
            double acc = (double)pf, pe = 0.0;
            float accf = pf;
            for (int i = 0; i < n; ++i) {
                const long long node = cell + dn[i];
                float iv;
                if (a.zfield)
                    iv = (a.zfield[node] <= thr) ? 1.f : 0.f;
                else if (cat >= 0)
                    iv = ((int)a.ind[node] == cat) ? 1.f : 0.f;
                else
                    iv = (a.ind[(long long)k * g.n + node] == 1) ? 1.f : 0.f;
                const double term = sv[i] * (double)(iv - pf);
                if (cat >= 0)
                    accf += (float)term;
                else
                    acc += term;
                pe += sv[i] * ((double)iv - (double)pf);
            }
            const long long row = (long long)(cat >= 0 ? cat : k) * g.n + cell;
            if (a.out_est64) a.out_est64[row] = (double)pf + pe;
            if (cat >= 0) {
                if (a.out_prob) a.out_prob[row] = accf;
            } else {
                if (a.out_class) a.out_class[row] = ((float)acc <= 0.5f) ? 0 : 1;
                if (a.out_prob) a.out_prob[row] = (float)acc;
            }
        }
        if (a.order_relations && a.out_prob) {
            float p[16];
            for (int k = 0; k < a.nind; ++k) {
                const float v = a.out_prob[(long long)k * g.n + cell];
                p[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            }
            for (int k = 0; k + 1 < a.nind; ++k)
                if (p[k] > p[k + 1]) {
                    const double avg = (double)(p[k] + p[k + 1]) / 2.0;
                    p[k] = (float)avg;
                    p[k + 1] = (float)avg;
                }
            for (int k = 0; k < a.nind; ++k) a.out_prob[(long long)k * g.n + cell] = p[k];
        }
        if (a.out_mask) a.out_mask[cell] = 1;
        const double var = a.c0 - wb;
        if (a.out_var) a.out_var[cell] = status == ST_KRIGED ? (var < 0 ? 0.0 : var) : -1.0;
        if (a.out_status) a.out_status[cell] = (unsigned char)status;
    }
}

// Systems of up to NMAX (<= 64) neighbours: one wavefront per cell
// (solve_wave.h), four cells per block.
template <int NMAX, int KIND>
__global__ __launch_bounds__(256) void k_solve_wave(SolveArgs a) {
    __shared__ double col[4][64];
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const long long cell = (long long)blockIdx.x * 4 + w;
    if (cell < a.g.n) wave_solve_cell<NMAX, KIND>(a, cell, a.count[cell], lane, col[w]);
}

// Largest neighbour count of the grid (into *out, zeroed by the first lane).
__global__ __launch_bounds__(256) void k_count_max(const int* count, long long n, int* out) {
    __shared__ int wmax[4];
    int m = 0;
    for (long long i = (long long)blockIdx.x * 256 + threadIdx.x; i < n; i += (long long)gridDim.x * 256)
        m = max(m, count_n(count[i]));
    for (int d = 32; d > 0; d >>= 1) m = max(m, __shfl_xor(m, d));
    if ((threadIdx.x & 63) == 0) wmax[threadIdx.x >> 6] = m;
    __syncthreads();
    if (threadIdx.x == 0) atomicMax(out, max(max(wmax[0], wmax[1]), max(wmax[2], wmax[3])));
}

__global__ void k_zero_int(int* p) { *p = 0; }

// Beyond this the wide solver's global slots are given back after the call.
constexpr size_t kWideKeepBytes = 256u << 20;

template <int KIND>
int launch_wide(const SolveArgs& sa, Workspace& ws, const hipEvent_t* ev, hipStream_t st, std::string& err) {
    // The working sets are sized by the largest neighbour count the search
    // found (read back once), not by max_neighbours: a grid smaller than
    // max_neighbours gives every system fewer rows.
    int* dmax = (int*)ws.get("wide_max", 4, err);
    if (!dmax) return -1;
    hipLaunchKernelGGL(k_zero_int, dim3(1), dim3(1), 0, st, dmax);
    hipLaunchKernelGGL(k_count_max, dim3(256), dim3(256), 0, st, sa.count, sa.g.n, dmax);
    int found = 0;
    if (hipMemcpyAsync(&found, dmax, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "reading the largest neighbour count failed";
        return -1;
    }
    const int kcap = std::max(1, std::min(found, sa.kmax));
    const long long n = sa.g.n;
    if (kcap <= 64) {
        const unsigned blocks = (unsigned)((n + 3) / 4);
        if (kcap <= 48)
            HPGLX_LAUNCH_EV((k_solve_wave<48, KIND>), dim3(blocks), dim3(256), 0, st, ev[0], ev[1], sa);
        else
            HPGLX_LAUNCH_EV((k_solve_wave<64, KIND>), dim3(blocks), dim3(256), 0, st, ev[0], ev[1], sa);
        return 0;
    }
    const long long slot = wide_slot(kcap) * 8;
    if (slot <= kWideLdsBytes) {
        const void* fn = (const void*)&k_solve_wide<KIND>;
        if (slot > 65536 && hipFuncSetAttribute(fn, hipFuncAttributeMaxDynamicSharedMemorySize, (int)slot) != hipSuccess) {
            err = "cannot raise the LDS limit of the wide solver";
            return -1;
        }
        const unsigned blocks = (unsigned)std::min<long long>(n, 8192);
        HPGLX_LAUNCH_EV(k_solve_wide<KIND>, dim3(blocks), dim3(WB), (uint32_t)slot, st, ev[0], ev[1], sa,
                              (double*)nullptr, kcap);
        return 0;
    }
    const long long blocks = std::min<long long>({n, 2048LL, kWideScratchBytes / slot});
    if (blocks < 1) {
        err = "max_neighbours is too large for the wide solver's scratch (one system above 2 GiB)";
        return -1;
    }
    auto* scratch = (double*)ws.get("wide_scratch", (size_t)(blocks * slot), err);
    if (!err.empty()) return -1;
    HPGLX_LAUNCH_EV(k_solve_wide<KIND>, dim3((unsigned)blocks), dim3(WB), 0, st, ev[0], ev[1], sa, scratch,
                          kcap);
    if ((size_t)(blocks * slot) > kWideKeepBytes && !ws.arena) {
        if (hipStreamSynchronize(st) != hipSuccess) {
            err = "the wide solver failed";
            return -1;
        }
        ws.drop("wide_scratch");
    }
    return 0;
}

}  // namespace

void sgs_level_48_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<48, 8, 0>(sa, la, st);
}
void sgs_level_48_1(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<48, 8, 1>(sa, la, st);
}
void sgs_level_64_0(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<64, 8, 0>(sa, la, st);
}
void sgs_level_64_1(const SolveArgs& sa, const SgsLevelArgs& la, hipStream_t st) {
    launch_sgs_level<64, 8, 1>(sa, la, st);
}

int dispatch_solve_wide(const SolveArgs& sa, int kind, Workspace& ws, const hipEvent_t* ev, hipStream_t st,
                        std::string& err) {
    if (kind == KIND_OK) return launch_wide<0>(sa, ws, ev, st, err);
    if (kind == KIND_SK) return launch_wide<1>(sa, ws, ev, st, err);
    return launch_wide<2>(sa, ws, ev, st, err);
}

}  // namespace hpglx

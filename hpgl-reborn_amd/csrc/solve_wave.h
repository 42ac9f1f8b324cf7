// solve_wave.h -- one wavefront per cell for systems of 33..64 neighbours.
//
// Lane i owns row i of the FP64 system (NMAX doubles in registers, static
// indices); dead rows (i >= n) are identity rows.  The factor is the
// right-looking Cholesky of the register solvers: at step j the pivot comes
// from lane j, every lane scales its column-j entry, the column is written
// to a per-wave LDS buffer (one double per lane) and every lane takes its
// trailing update from it (reads at one address per instruction are
// broadcasts, free of bank conflicts).  No workgroup barriers: the wave
// writes and reads its own buffer in program order.  Forward substitution is
// column-oriented (the new unknown comes from its owner by readlane), the
// backward one a wave-wide dot product per unknown.  OK takes the Schur
// complement from the forward results (u = L^-1 1, z = L^-1 b: mu = (u.z - 1)
// / u.u, w = L^-T (z - mu u)), one backward pass.
//
// Reference algebra and policies as the register path (solve.h):
// my_kriging_weights.h:60-361 (failure on a pivot that is not positive, as
// dpotrf; OK singular when |1'A^-1 1| < 1e-15), combiner.h:10-25 (float sum
// in neighbour order), cont_kriging.h:73-105, median_ik.cpp:60-78,
// indicator_kriging.h:34-59.
#pragma once

#include "solve.h"

namespace hpglx {

// Lane j's double in every lane (j a compile-time constant after unrolling).
__device__ __forceinline__ double wave_bcast(double v, int j) {
    const int lo = __builtin_amdgcn_readlane(__double2loint(v), j);
    const int hi = __builtin_amdgcn_readlane(__double2hiint(v), j);
    return __hiloint2double(hi, lo);
}
__device__ __forceinline__ float wave_bcast(float v, int j) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), j));
}

// A DPP permutation of a double (both halves moved).
template <int CTRL>
__device__ __forceinline__ double dpp_double(double v) {
    const int lo = __builtin_amdgcn_mov_dpp(__double2loint(v), CTRL, 0xf, 0xf, true);
    const int hi = __builtin_amdgcn_mov_dpp(__double2hiint(v), CTRL, 0xf, 0xf, true);
    return __hiloint2double(hi, lo);
}

// Sum over the 64 lanes; every lane ends with the same bits.  Within each
// row of 16 lanes by DPP (quad exchanges, then half-row and row mirrors),
// then the four row sums by readlane: vector-ALU latencies only.
__device__ __forceinline__ double wave_sum(double v) {
    v += dpp_double<0xB1>(v);   // quad_perm [1,0,3,2]
    v += dpp_double<0x4E>(v);   // quad_perm [2,3,0,1]
    v += dpp_double<0x141>(v);  // row_half_mirror
    v += dpp_double<0x140>(v);  // row_mirror
    return (wave_bcast(v, 0) + wave_bcast(v, 16)) + (wave_bcast(v, 32) + wave_bcast(v, 48));
}

// Kriges one cell with the whole wave (lane = 0..63).  col: this wave's LDS
// buffer of 64 doubles.  FP64 regardless of the precision asked.
template <int NMAX, int KIND>
__device__ __forceinline__ void wave_solve_cell(const SolveArgs& a, long long cell, int cnt, int lane, double* col) {
    static_assert(NMAX <= 64, "one row per lane");
    constexpr bool OKK = KIND == 0;
    const GridDims& g = a.g;
    if (KIND != 2 && cnt == CNT_INFORMED) {
        write_informed<64>(a, cell, lane);
        return;
    }
    const int n = cnt < 0 ? 0 : (cnt & CNT_MASK);
    const int SX = 2 * a.wx + 1, SY = 2 * a.wy + 1;
    const int center = (a.wz * SY + a.wy) * SX + a.wx;
    const bool live = lane < n;
    int tr = 0, dn = 0;
    if (live) {
        int dx, dy, dz;
        unpack_off(a.nboff[cell * a.kmax + lane], dx, dy, dz);
        tr = (dz * SY + dy) * SX + dx;
        dn = (dz * g.ny + dy) * g.nx + dx;
    }
    // Row `lane` of A on and below the diagonal (the rest is never read).
    double arow[NMAX];
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
        const int tc = __builtin_amdgcn_readlane(tr, c);
        double v = (lane == c) ? 1.0 : 0.0;
        if (live && c <= lane) v = a.ctab[center + tr - tc];
        arow[c] = v;
        // bound the loads in flight (registers)
        if (c % 8 == 7) __builtin_amdgcn_sched_barrier(0);
    }
    const double b = live ? a.ctab[center + tr] : 0.0;

    int status = n == 0 ? ST_NO_NEIGHBOURS : ST_KRIGED;
    double w = 0.0, mu = 0.0;
    if (status == ST_KRIGED) {
        double dinv = 1.0;  // 1 / L[lane][lane]
        bool ok = true;
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            const double piv = wave_bcast(arow[j], j);
            ok = ok && (piv > 0.0);
            const double rs = rsqrt_t(piv);
            if (lane == j) dinv = rs;
            arow[j] *= rs;  // lanes below j: L[lane][j]
            col[lane] = arow[j];
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
#pragma unroll
            for (int c = j + 1; c < NMAX; ++c) {
                arow[c] = fmadd(-arow[j], col[c], arow[c]);
                if (c % 16 == 15) __builtin_amdgcn_sched_barrier(0);
            }
            __builtin_amdgcn_wave_barrier();
        }
        if (!ok) {
            status = ST_SINGULAR;
        } else {
            // z = L^-1 b (and u = L^-1 1 for OK), column by column
            double z = b, u = live ? 1.0 : 0.0;
#pragma unroll
            for (int j = 0; j < NMAX; ++j) {
                const double zj = wave_bcast(z * dinv, j);
                const double uj = OKK ? wave_bcast(u * dinv, j) : 0.0;
                if (lane == j) {
                    z = zj;
                    if (OKK) u = uj;
                } else if (lane > j) {
                    z = fmadd(-arow[j], zj, z);
                    if (OKK) u = fmadd(-arow[j], uj, u);
                }
            }
            double y = z;
            if (OKK) {
                const double sum_sk = wave_sum(u * z), sum_ones = wave_sum(u * u);
                if (fabs(sum_ones) < 1e-15) {
                    status = ST_SINGULAR;
                } else {
                    mu = (sum_sk - 1) / sum_ones;
                    y = z - mu * u;
                }
            }
            if (status == ST_KRIGED) {
                // w = L^-T y: unknown i from the rows below it
#pragma unroll
                for (int i = NMAX - 1; i >= 0; --i) {
                    const double p = wave_sum(lane > i ? arow[i] * y : 0.0);
                    if (lane == i) y = (y - p) * dinv;
                }
                w = live ? y : 0.0;
            }
        }
    }
    if (status != ST_KRIGED) {
        w = 0.0;
        mu = 0.0;
    }
    const double wb = wave_sum(w * b);

    if (KIND != 2) {
        const double mean = (KIND == 1) ? (a.mean_dev ? a.mean_dev[0] : a.mean) : 0.0;
        const bool lvm = KIND == 1 && a.lvm_mean != nullptr;
        const float mf = lvm ? a.lvm_mean[cell] : (float)mean;
        const long long node = cell + (live ? dn : 0);
        const float v = live ? a.values[node] : 0.0f;
        const float ms = lvm ? (live ? a.lvm_mean[node] : 0.0f) : mf;
        const float term = (float)(w * (double)(v - ms));
        const double e64 = (double)mf + wave_sum(w * ((double)v - (double)ms));
        float acc = mf;
#pragma unroll
        for (int i = 0; i < NMAX; ++i) {
            const float ti = wave_bcast(term, i);
            if (i < n) acc += ti;
        }
        const double var = a.c0 - wb - mu;
        if (lane == 0) {
            if (status == ST_KRIGED) {
                if (a.out_val) a.out_val[cell] = acc;
                if (a.out_mask) a.out_mask[cell] = 1;
                if (a.out_var) a.out_var[cell] = var < 0 ? 0.0 : var;
                if (a.out_est64) a.out_est64[cell] = e64;
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
        }
        return;
    }

    // Median indicator kriging and indicator categories, as solve_cell.
    const long long node = cell + (live ? dn : 0);
    for (int k = 0; k < a.nind; ++k) {
        const float pf = (float)a.marg[k];
        const float thr = a.zfield ? a.thresholds[k] : 0.f;
        const int cat = a.has_cats ? a.cats[k] : -1;
        float iv;
        if (a.zfield)
            iv = (a.zfield[node] <= thr) ? 1.f : 0.f;
        else if (cat >= 0)
            iv = ((int)a.ind[node] == cat) ? 1.f : 0.f;
        else
            iv = (a.ind[(long long)k * g.n + node] == 1) ? 1.f : 0.f;
        iv = live ? iv : 0.f;
        const double term = w * (double)(iv - pf);
        const double e64 = (double)pf + wave_sum(w * ((double)iv - (double)pf));
        double acc = (double)pf;
        float accf = pf;
#pragma unroll
        for (int i = 0; i < NMAX; ++i) {
            const double ti = wave_bcast(term, i);
            if (i < n) {
                if (cat >= 0)
                    accf += (float)ti;
                else
                    acc += ti;
            }
        }
        const long long row = (long long)(cat >= 0 ? cat : k) * g.n + cell;
        if (lane == 0) {
            if (a.out_est64) a.out_est64[row] = e64;
            if (cat >= 0) {
                if (a.out_prob) a.out_prob[row] = accf;
            } else {
                if (a.out_class) a.out_class[row] = ((float)acc <= 0.5f) ? 0 : 1;
                if (a.out_prob) a.out_prob[row] = (float)acc;
            }
        }
    }
    if (lane == 0) {
        if (a.order_relations && a.out_prob) {
            // indicator_kriging.h:34-59 on this cell's CDF
            constexpr int KMAXIND = 16;
            float p[KMAXIND];
#pragma unroll
            for (int k = 0; k < KMAXIND; ++k) {
                const float v = k < a.nind ? a.out_prob[(long long)k * g.n + cell] : 0.f;
                p[k] = v < 0.0f ? 0.0f : (v > 1.0f ? 1.0f : v);
            }
#pragma unroll
            for (int k = 0; k + 1 < KMAXIND; ++k)
                if (k + 1 < a.nind && p[k] > p[k + 1]) {
                    const double avg = (double)(p[k] + p[k + 1]) / 2.0;  // float sum, as the reference
                    p[k] = (float)avg;
                    p[k + 1] = (float)avg;
                }
#pragma unroll
            for (int k = 0; k < KMAXIND; ++k)
                if (k < a.nind) a.out_prob[(long long)k * g.n + cell] = p[k];
        }
        if (a.out_mask) a.out_mask[cell] = 1;
        const double var = a.c0 - wb;
        if (a.out_var) a.out_var[cell] = status == ST_KRIGED ? (var < 0 ? 0.0 : var) : -1.0;
        if (a.out_status) a.out_status[cell] = (unsigned char)status;
    }
}

}  // namespace hpglx

// solve.h -- per-cell covariance assembly, Cholesky, SK/OK solve and combine.
//
// Layout: a group of G lanes (G = 4 or 8) kriges one cell.  Lane t owns the
// rows r = G*s + t of the n x n system (s = 0 .. NMAX/G - 1 "slots") and keeps,
// for slot s, the columns 0 .. G*s + G - 1 in registers: a ragged lower
// triangle with static indices, about NMAX^2 / (2G) + NMAX/2 values per lane.
//   * The factor is stored negated: the caller fills -A, the factor leaves
//     -L.  Every update is then an add of a product (v_fmac), into which the
//     compiler folds the broadcast operand as DPP, so a broadcast costs no
//     instruction of its own.
//   * Right-looking Cholesky: at step j the owner of row j broadcasts the
//     pivot; every lane scales its column-j entries; then for each c > j the
//     entry L[c][j] is broadcast from its owner and every slot holding column
//     c does one FMA.  One broadcast feeds up to NMAX/G FMAs.
//   * Forward solve: column-oriented, one broadcast per step.
//   * Backward solve: each lane forms the partial dot product of its rows
//     below i with column i, the group sums it, the owner divides.
// Broadcasts and sums use DPP quad permutes (G = 4) or ds_swizzle (G = 8):
// no LDS traffic, no address arithmetic.  Entries above the diagonal in a
// lane's slot width are written with garbage and never read.
//
// FP64 form: everything in double.  Mixed form: the factor and both
// triangular solves in float; residuals b - A x in double from the FP64
// covariance table, corrections from the float factor; a cell whose
// smallest float pivot falls under pivot_tol * C(0), or whose float factor
// fails, is queued for the FP64 form.
//
// Reference algebra (src/geo_bsd/hpgl/my_kriging_weights.h):
//   SK (:60-196)  A w = b, var = C(0) - w.b
//   OK (:198-361) A s = b, A o = 1, mu = (sum s - 1) / sum o, w = s - mu o,
//                 singular when |sum o| < 1e-15, var = C(0) - w.b - mu
//   failure       Cholesky pivot <= 0 (dpotrf info > 0)
// Combine (src/geo_bsd/hpgl/combiner.h:10-25): m + sum (float)(w_i (v_i - m))
// accumulated in float in neighbour order (double for median IK).
#pragma once

#include <type_traits>

#include "krige_kernels.h"

namespace hpglx {

// ------------------------------------------------------- group primitives
template <int G>
struct Grp;

template <>
struct Grp<4> {
    // lane q of each quad to all four lanes (DPP quad_perm [q,q,q,q])
    static __device__ __forceinline__ int bcast(int v, int q) {
        switch (q & 3) {
            case 0: return __builtin_amdgcn_mov_dpp(v, 0x00, 0xf, 0xf, true);
            case 1: return __builtin_amdgcn_mov_dpp(v, 0x55, 0xf, 0xf, true);
            case 2: return __builtin_amdgcn_mov_dpp(v, 0xAA, 0xf, 0xf, true);
            default: return __builtin_amdgcn_mov_dpp(v, 0xFF, 0xf, 0xf, true);
        }
    }
    // value of lane (t ^ m), m in {1, 2}
    static __device__ __forceinline__ int xchg(int v, int m) {
        if (m == 1) return __builtin_amdgcn_mov_dpp(v, 0xB1, 0xf, 0xf, true);
        return __builtin_amdgcn_mov_dpp(v, 0x4E, 0xf, 0xf, true);
    }
};

template <>
struct Grp<8> {
    // ds_swizzle bit mode within 32 lanes: lane' = ((lane & and) | or) ^ xor
    static __device__ __forceinline__ int bcast(int v, int q) {
        switch (q & 7) {
            case 0: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (0 << 5));
            case 1: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (1 << 5));
            case 2: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (2 << 5));
            case 3: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (3 << 5));
            case 4: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (4 << 5));
            case 5: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (5 << 5));
            case 6: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (6 << 5));
            default: return __builtin_amdgcn_ds_swizzle(v, 0x18 | (7 << 5));
        }
    }
    static __device__ __forceinline__ int xchg(int v, int m) {
        if (m == 1) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (1 << 10));
        if (m == 2) return __builtin_amdgcn_ds_swizzle(v, 0x1F | (2 << 10));
        return __builtin_amdgcn_ds_swizzle(v, 0x1F | (4 << 10));
    }
};

template <int G>
__device__ __forceinline__ int gbc(int v, int q) {
    return Grp<G>::bcast(v, q);
}
template <int G>
__device__ __forceinline__ float gbc(float v, int q) {
    return __int_as_float(Grp<G>::bcast(__float_as_int(v), q));
}
template <int G>
__device__ __forceinline__ double gbc(double v, int q) {
    const int lo = Grp<G>::bcast(__double2loint(v), q);
    const int hi = Grp<G>::bcast(__double2hiint(v), q);
    return __hiloint2double(hi, lo);
}
template <int G>
__device__ __forceinline__ float gxchg(float v, int m) {
    return __int_as_float(Grp<G>::xchg(__float_as_int(v), m));
}
template <int G>
__device__ __forceinline__ double gxchg(double v, int m) {
    const int lo = Grp<G>::xchg(__double2loint(v), m);
    const int hi = Grp<G>::xchg(__double2hiint(v), m);
    return __hiloint2double(hi, lo);
}
// Butterfly sum: every lane of the group ends with the same bits.
template <int G, typename T>
__device__ __forceinline__ T gsum(T v) {
    v += gxchg<G>(v, 1);
    v += gxchg<G>(v, 2);
    if (G == 8) v += gxchg<G>(v, 4);
    return v;
}

__device__ __forceinline__ float rsqrt_t(float x) {
    return __builtin_amdgcn_rsqf(x);
}
// FP64 reciprocal square root: hardware estimate + two Newton steps.
__device__ __forceinline__ double rsqrt_t(double x) {
    double y = __builtin_amdgcn_rsq(x);
    double h = x * y;
    double e = __builtin_fma(-h, y, 1.0);
    y = __builtin_fma(0.5 * y, e, y);
    h = x * y;
    e = __builtin_fma(-h, y, 1.0);
    return __builtin_fma(0.5 * y, e, y);
}

__device__ __forceinline__ float fmadd(float a, float b, float c) {
    return __builtin_fmaf(a, b, c);
}
__device__ __forceinline__ double fmadd(double a, double b, double c) {
    return __builtin_fma(a, b, c);
}

// Offset of slot s in the ragged row storage.
template <int G>
__host__ __device__ constexpr int slot_off(int s) {
    return G * s * (s + 1) / 2;
}

// ----------------------------------------------------------- factor/solve
template <int NMAX, int G, typename T>
struct Chol {
    static constexpr int S = NMAX / G;
    static constexpr int TOT = slot_off<G>(S);

    // In-place Cholesky of the negated matrix: on entry the lower triangle
    // holds -A, on return -L (the diagonal entries are not read again) and
    // D[s] = 1 / L[r_s][r_s].  Group-uniform result.
    static __device__ __forceinline__ bool factor(T (&L)[TOT], T (&D)[S], T& minpiv, int t) {
        bool ok = true;
        minpiv = T(3.0e38);
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            const int q = j % G, sj = j / G;
            const T piv = -gbc<G>(L[slot_off<G>(sj) + j], q);
            ok = ok && (piv > T(0));
            minpiv = piv < minpiv ? piv : minpiv;
            const T rs = rsqrt_t(piv);
            if (t == q) D[sj] = rs;
#pragma unroll
            for (int s = sj; s < S; ++s) L[slot_off<G>(s) + j] *= rs;
            // -A[s][c] + L[s][j] L[c][j] = (-L[s][j]) (-L[c][j]) + (-A[s][c])
#pragma unroll
            for (int c = j + 1; c < NMAX; ++c) {
                const T lc = gbc<G>(L[slot_off<G>(c / G) + j], c % G);
#pragma unroll
                for (int s = c / G; s < S; ++s)
                    L[slot_off<G>(s) + c] = fmadd(L[slot_off<G>(s) + j], lc, L[slot_off<G>(s) + c]);
            }
        }
        return ok;
    }

    // Blocked form for float with groups of 4 lanes.  Lane t holds rows
    // 4s+t, so the 4x4 tile (s, b) of the lower triangle is the lane's
    // registers L[slot_off(s) + 4b .. 4b+3] (row 4s+t, columns 4b..4b+3).
    // Per panel of 4 columns the panel is factored with broadcasts as above;
    // every trailing tile then takes the panel's rank-4 update on the matrix
    // cores: v_mfma_f32_4x4x1_16b_f32 forms 16 independent 4x4 outer
    // products per wave, one per group, D[i][j] (register i, lane 4g+j) +=
    // A[i] (lane 4g+i) * B[j] (lane 4g+j).  With A = column entries of tile
    // row b and B = those of tile row s, register i of lane t receives
    // (-L[4b+i][k]) (-L[4s+t][k]), the update of entry (4s+t, 4b+i).  Each
    // entry takes its updates in column order, as in factor().
    static __device__ __forceinline__ bool factor_mfma(T (&L)[TOT], T (&D)[S], T& minpiv, int t) {
        static_assert(G == 4, "4x4 tiles need groups of 4 lanes");
        typedef float f4 __attribute__((ext_vector_type(4)));
        bool ok = true;
        minpiv = T(3.0e38);
#pragma unroll
        for (int p = 0; p < S; ++p) {
#pragma unroll
            for (int q = 0; q < 4; ++q) {
                const int j = 4 * p + q;
                const T piv = -gbc<G>(L[slot_off<G>(p) + j], q);
                ok = ok && (piv > T(0));
                minpiv = piv < minpiv ? piv : minpiv;
                const T rs = rsqrt_t(piv);
                if (t == q) D[p] = rs;
#pragma unroll
                for (int s = p; s < S; ++s) L[slot_off<G>(s) + j] *= rs;
#pragma unroll
                for (int c = j + 1; c < 4 * p + 4; ++c) {
                    const T lc = gbc<G>(L[slot_off<G>(p) + j], c % 4);
#pragma unroll
                    for (int s = p; s < S; ++s)
                        L[slot_off<G>(s) + c] = fmadd(L[slot_off<G>(s) + j], lc, L[slot_off<G>(s) + c]);
                }
            }
#pragma unroll
            for (int b = p + 1; b < S; ++b) {
#pragma unroll
                for (int s = b; s < S; ++s) {
                    T* tile = &L[slot_off<G>(s) + 4 * b];
                    f4 acc = {tile[0], tile[1], tile[2], tile[3]};
#pragma unroll
                    for (int k = 0; k < 4; ++k)
                        acc = __builtin_amdgcn_mfma_f32_4x4x1f32(L[slot_off<G>(b) + 4 * p + k],
                                                                 L[slot_off<G>(s) + 4 * p + k], acc, 0, 0, 0);
                    tile[0] = acc[0];
                    tile[1] = acc[1];
                    tile[2] = acc[2];
                    tile[3] = acc[3];
                }
            }
        }
        return ok;
    }

    // z <- L^-1 z (column-oriented, one broadcast per step)
    template <typename V>
    static __device__ __forceinline__ void forward(const T (&L)[TOT], const T (&D)[S], V (&z)[S], int t) {
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            const int q = j % G, sj = j / G;
            const V yl = z[sj] * (V)D[sj];  // y_j on its owner
            const V yj = gbc<G>(yl, q);
#pragma unroll
            for (int s = sj; s < S; ++s) {
                const V upd = fmadd((V)L[slot_off<G>(s) + j], yj, z[s]);
                if (s == sj)
                    z[s] = (t == q) ? yl : ((t > q) ? upd : z[s]);
                else
                    z[s] = upd;
            }
        }
    }

    // Clears the entries on and above the diagonal of each slot's diagonal
    // tile (the factor leaves them with values it never reads; the diagonal
    // itself lives on as D).  The solves below then need no per-lane tests:
    // a cleared entry adds 0 * y to a row that is not to change.
    static __device__ __forceinline__ void clear_upper(T (&L)[TOT], int t) {
#pragma unroll
        for (int s = 0; s < S; ++s)
#pragma unroll
            for (int i = 0; i < G; ++i) {
                T& e = L[slot_off<G>(s) + G * s + i];
                e = (t > i) ? e : T(0);
            }
    }

    // z <- L^-1 z after clear_upper: every row takes every update (rows not
    // below j through a cleared entry); row r is scaled by D once, at the
    // end, with the value its owner broadcast at step r.
    template <typename V>
    static __device__ __forceinline__ void forward_cu(const T (&L)[TOT], const T (&D)[S], V (&z)[S]) {
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            const int q = j % G, sj = j / G;
            const V yj = gbc<G>(z[sj] * (V)D[sj], q);
#pragma unroll
            for (int s = sj; s < S; ++s) z[s] = fmadd((V)L[slot_off<G>(s) + j], yj, z[s]);
        }
#pragma unroll
        for (int s = 0; s < S; ++s) z[s] *= (V)D[s];
    }

    template <typename V>
    static __device__ __forceinline__ void forward2_cu(const T (&L)[TOT], const T (&D)[S], V (&z)[S], V (&u)[S]) {
#pragma unroll
        for (int j = 0; j < NMAX; ++j) {
            const int q = j % G, sj = j / G;
            const V zj = gbc<G>(z[sj] * (V)D[sj], q), uj = gbc<G>(u[sj] * (V)D[sj], q);
#pragma unroll
            for (int s = sj; s < S; ++s) {
                const V lv = (V)L[slot_off<G>(s) + j];
                z[s] = fmadd(lv, zj, z[s]);
                u[s] = fmadd(lv, uj, u[s]);
            }
        }
#pragma unroll
        for (int s = 0; s < S; ++s) {
            z[s] *= (V)D[s];
            u[s] *= (V)D[s];
        }
    }

    // x <- L^-T x after clear_upper (cleared entries add 0 to the sums).
    template <typename V>
    static __device__ __forceinline__ void backward_cu(const T (&L)[TOT], const T (&D)[S], V (&x)[S], int t) {
#pragma unroll
        for (int i = NMAX - 1; i >= 0; --i) {
            const int q = i % G, si = i / G;
            V p = V(0);
#pragma unroll
            for (int s = si; s < S; ++s) p += (V)L[slot_off<G>(s) + i] * x[s];
            p = gsum<G>(p);
            if (t == q) x[si] = (x[si] + p) * (V)D[si];
        }
    }

    // x <- L^-T x (partial dot products, group sum, owner divides)
    template <typename V>
    static __device__ __forceinline__ void backward(const T (&L)[TOT], const T (&D)[S], V (&x)[S], int t) {
#pragma unroll
        for (int i = NMAX - 1; i >= 0; --i) {
            const int q = i % G, si = i / G;
            V p = V(0);
#pragma unroll
            for (int s = si; s < S; ++s) {
                const V prod = (V)L[slot_off<G>(s) + i] * x[s];
                if (s == si)
                    p += (t > q) ? prod : V(0);
                else
                    p += prod;
            }
            p = gsum<G>(p);
            if (t == q) x[si] = (x[si] + p) * (V)D[si];
        }
    }
};

template <int G>
__device__ __forceinline__ void write_informed(const SolveArgs& a, long long cell, int t) {
    if (t != 0) return;
    const float v = a.values[cell];
    if (a.out_val) a.out_val[cell] = v;
    if (a.out_mask) a.out_mask[cell] = 1;
    if (a.out_var) a.out_var[cell] = 0.0;
    if (a.out_est64) a.out_est64[cell] = v;
    if (a.out_status) a.out_status[cell] = ST_INFORMED;
}

// |a - b| + acc on non-negative ints (v_sad_u32).
__device__ __forceinline__ unsigned adiff(unsigned a, unsigned b, unsigned acc) {
    return (a > b ? a - b : b - a) + acc;
}

// Octant form of the covariance lookup (LTAB): an octant table, staged in
// LDS or read from global memory, indexed by |dx|, |dy|, |dz| of the pair's
// offset difference with row strides ox (x fastest) and oxy.  Each neighbour
// offset is kept as three biased, pre-scaled components
// ((d + OFF_BIAS) * stride * 8), so a pair's byte address is the sum of three
// absolute differences.  In the packed form (PACK, the LDS table of the
// template path, whose offsets satisfy |d| <= E/2) x and y share one word as
// two 16-bit fields ((d + E/2) * stride * 8, each below 32768), and one
// v_sad_u16 forms both their absolute differences.
constexpr int OFF_BIAS = 1024;

// Weights of one cell: w[s] for the lane's rows, mu (OK), status.  Returns
// false when the mixed form hands the cell to the FP64 form.
// FULL: the cell has exactly NMAX neighbours, so no row is dead and no
// liveness test is compiled.
template <int NMAX, int G, bool MIXED, int KIND, bool LTAB, bool PACK, bool FULL>
__device__ __forceinline__ bool cell_weights(const SolveArgs& a, long long cell, int t, int n, const double* ltab,
                                             unsigned ox, unsigned oxy, double (&w)[NMAX / G],
                                             double (&bq_out)[NMAX / G], double& wb, int (&dnode)[NMAX / G],
                                             double& mu, int& status) {
    using T = typename std::conditional<MIXED, float, double>::type;
    typedef Chol<NMAX, G, T> CH;
    constexpr int S = CH::S;
    constexpr bool OKK = KIND == 0;
    const GridDims& g = a.g;
    const int SX = 2 * a.wx + 1, SY = 2 * a.wy + 1;
    const int center = (a.wz * SY + a.wy) * SX + a.wx;
    const int* nb = a.nboff + cell * a.kmax;

    // Table-linear offsets of my rows; column offsets come by broadcast.
    int tr[S];
    unsigned ux[S], uy[S], uz[S];
    int pk[S];
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int r = G * s + t;
        pk[s] = nb[(FULL || r < n) ? r : 0];  // unconditional, so the loads issue together
    }
    const unsigned OX = ox, OXY = oxy;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const int r = G * s + t;
        int dx, dy, dz;
        unpack_off(pk[s], dx, dy, dz);
        if (!FULL && r >= n) dx = dy = dz = 0;
        tr[s] = (dz * SY + dy) * SX + dx;
        dnode[s] = (dz * g.ny + dy) * g.nx + dx;
        if (PACK) {
            ux[s] = (unsigned)(dx + (a.ex >> 1)) * 8u | ((unsigned)(dy + (a.ey >> 1)) * (OX * 8u)) << 16;
            uy[s] = 0u;
        } else {
            ux[s] = (unsigned)(dx + OFF_BIAS) * 8u;
            uy[s] = (unsigned)(dy + OFF_BIAS) * (OX * 8u);
        }
        uz[s] = (unsigned)(dz + OFF_BIAS) * (OXY * 8u);
    }
    const char* lb = (const char*)ltab;
    // The packed form reads LDS only: its 32-bit base is the accumulator of
    // the first absolute difference.
    typedef const double __attribute__((address_space(3)))* lds_dp;
    const unsigned lbase = PACK ? (unsigned)(uintptr_t)(lds_dp)ltab : 0u;
    // Covariance of my slot s with the neighbour (cx, cy, cz).
    auto lcov = [&](int s, unsigned cx, unsigned cy, unsigned cz) -> double {
        if constexpr (PACK)
            return *(lds_dp)(uintptr_t)adiff(uz[s], cz, __builtin_amdgcn_sad_u16(ux[s], cx, lbase));
        else
            return *(const double*)(lb + adiff(ux[s], cx, adiff(uy[s], cy, adiff(uz[s], cz, 0u))));
    };
    T L[CH::TOT];
    T D[S];
    // b, the covariances of my rows with the target: the mixed form keeps it
    // here and returns w.b; the FP64 form hands it to the caller, which
    // forms w.b (each allocates its registers better that way, measured).
    double bq_here[S];
    double(&bq)[S] = MIXED ? bq_here : bq_out;
#pragma unroll
    for (int c = 0; c < NMAX; ++c) {
        int tcol = 0;
        unsigned cx = 0, cy = 0, cz = 0;
        if (LTAB) {
            cx = (unsigned)gbc<G>((int)ux[c / G], c % G);
            cy = PACK ? 0u : (unsigned)gbc<G>((int)uy[c / G], c % G);
            cz = (unsigned)gbc<G>((int)uz[c / G], c % G);
        } else {
            tcol = gbc<G>(tr[c / G], c % G);
        }
        // Every entry that is read has c <= r, so row liveness is enough;
        // dead rows and columns have offset 0, so every address is in range.
#pragma unroll
        for (int s = c / G; s < S; ++s) {
            const int r = G * s + t;
            const bool live = FULL || r < n;
            T v;
            if (LTAB)
                v = (T)lcov(s, cx, cy, cz);
            else if (MIXED)
                v = (T)a.ctab32[center + tr[s] - tcol];
            else
                v = (T)a.ctab[center + tr[s] - tcol];
            L[slot_off<G>(s) + c] = live ? -v : ((r == c) ? T(-1) : T(0));
        }
    }
    const unsigned c0x = PACK ? (unsigned)(a.ex >> 1) * 8u | ((unsigned)(a.ey >> 1) * (OX * 8u)) << 16
                              : (unsigned)OFF_BIAS * 8u;
    const unsigned c0y = (unsigned)OFF_BIAS * (OX * 8u), c0z = (unsigned)OFF_BIAS * (OXY * 8u);
#pragma unroll
    for (int s = 0; s < S; ++s) {
        const double v = LTAB ? lcov(s, c0x, c0y, c0z) : a.ctab[center + tr[s]];
        bq[s] = (FULL || G * s + t < n) ? v : 0.0;
    }
    T minpiv;
    bool ok;
    if constexpr (MIXED && G == 4)
        ok = CH::factor_mfma(L, D, minpiv, t);
    else
        ok = CH::factor(L, D, minpiv, t);
    if (MIXED && (!ok || !((double)minpiv >= a.pivot_tol * a.c0))) {
        if (t == 0) {
            const int pos = atomicAdd(a.n_fallback, 1);
            a.fallback_cells[pos] = (int)cell;
        }
        return false;
    }
    if (!ok) {
        status = ST_SINGULAR;
        return true;
    }
    CH::clear_upper(L, t);
    T u[S];  // OK: L^-1 1
    double sum_ones = 0.0;
    if (OKK) {
        // With z = L^-1 b and u = L^-1 1 (one forward pass, two right-hand
        // sides): 1'A^-1 b = u.z and 1'A^-1 1 = u.u, so the Schur complement
        // needs no second backward pass: mu = (u.z - 1) / u.u and
        // w = A^-1 (b - mu 1) = L^-T (z - mu u).
        T z[S];
#pragma unroll
        for (int s = 0; s < S; ++s) {
            z[s] = (T)bq[s];
            u[s] = (FULL || G * s + t < n) ? T(1) : T(0);
        }
        CH::forward2_cu(L, D, z, u);
        double pzu = 0.0, puu = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            pzu = fmadd((double)z[s], (double)u[s], pzu);
            puu = fmadd((double)u[s], (double)u[s], puu);
        }
        const double sum_sk = gsum<G>(pzu);
        sum_ones = gsum<G>(puu);
        if (fabs(sum_ones) < 1e-15) {
            status = ST_SINGULAR;
            return true;
        }
        mu = (sum_sk - 1) / sum_ones;
#pragma unroll
        for (int s = 0; s < S; ++s) z[s] = (T)((double)z[s] - mu * (double)u[s]);
        CH::backward_cu(L, D, z, t);
#pragma unroll
        for (int s = 0; s < S; ++s) w[s] = (double)z[s];
    } else {
        T z[S];
#pragma unroll
        for (int s = 0; s < S; ++s) z[s] = (T)bq[s];
        CH::forward_cu(L, D, z);
        CH::backward_cu(L, D, z, t);
#pragma unroll
        for (int s = 0; s < S; ++s) w[s] = (double)z[s];
    }
    if (MIXED) {
        // Refinement in FP64 residuals with the float factor.  SK: r = b - A w,
        // w += A^-1 r.  OK, on the bordered system [A 1; 1' 0] [w; mu] =
        // [b; 1]: r = b - A w - mu, r_mu = 1 - sum w; with f = L^-1 r,
        // dmu = (u.f - r_mu) / u.u and dw = A^-1 (r - dmu 1) = L^-T (f - dmu u).
        for (int it = 0; it < a.refine_steps; ++it) {
            double r[S];
            double r_mu = 0.0;
            if (OKK) {
                double pw = 0.0;
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    r[s] = bq[s] - mu;
                    pw += w[s];
                }
                r_mu = 1.0 - gsum<G>(pw);
            } else {
#pragma unroll
                for (int s = 0; s < S; ++s) r[s] = bq[s];
            }
            // r -= A w over every column (dead columns carry w = 0; dead rows
            // are cleared below).
#pragma unroll
            for (int c = 0; c < NMAX; ++c) {
                int tcol = 0;
                unsigned cx = 0, cy = 0, cz = 0;
                if (LTAB) {
                    cx = (unsigned)gbc<G>((int)ux[c / G], c % G);
                    cy = PACK ? 0u : (unsigned)gbc<G>((int)uy[c / G], c % G);
                    cz = (unsigned)gbc<G>((int)uz[c / G], c % G);
                } else {
                    tcol = gbc<G>(tr[c / G], c % G);
                }
                const double xw = gbc<G>(w[c / G], c % G);
#pragma unroll
                for (int s = 0; s < S; ++s) {
                    const double av = LTAB ? lcov(s, cx, cy, cz) : a.ctab[center + tr[s] - tcol];
                    r[s] = __builtin_fma(-av, xw, r[s]);
                }
                // bound the loads the scheduler keeps in flight
                if (c % 2 == 1) __builtin_amdgcn_sched_barrier(0);
            }
            if (!FULL) {
#pragma unroll
                for (int s = 0; s < S; ++s) r[s] = (G * s + t < n) ? r[s] : 0.0;
            }
            T f[S];
#pragma unroll
            for (int s = 0; s < S; ++s) f[s] = (T)r[s];
            CH::forward_cu(L, D, f);
            if (OKK) {
                double pf = 0.0;
#pragma unroll
                for (int s = 0; s < S; ++s) pf = fmadd((double)u[s], (double)f[s], pf);
                const double dmu = (gsum<G>(pf) - r_mu) / sum_ones;
#pragma unroll
                for (int s = 0; s < S; ++s) f[s] = (T)((double)f[s] - dmu * (double)u[s]);
                mu += dmu;
            }
            CH::backward_cu(L, D, f, t);
#pragma unroll
            for (int s = 0; s < S; ++s) w[s] += (double)f[s];
        }
    }
    if (MIXED) {
        double pwb = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) pwb += w[s] * bq[s];
        wb = gsum<G>(pwb);
    }
    return true;
}

// Kriges one cell with the group of G lanes; t = lane index in the group,
// cnt = the cell's count[] entry, ltab/ox/oxy = the octant table and its
// strides (LTAB form).
template <int NMAX, int G, bool MIXED, int KIND, bool LTAB, bool PACK = false, bool FULL = false>
__device__ __forceinline__ void solve_cell(const SolveArgs& a, long long cell, int cnt, int t, const double* ltab,
                                           unsigned ox, unsigned oxy) {
    constexpr int S = NMAX / G;
    const GridDims& g = a.g;
    if (KIND != 2 && cnt == CNT_INFORMED) {
        write_informed<G>(a, cell, t);
        return;
    }
    const int n = cnt < 0 ? 0 : (cnt & CNT_MASK);
    int status = ST_KRIGED;
    double w[S], bq[S];
    int dnode[S];
    double mu = 0.0, wb = 0.0;
#pragma unroll
    for (int s = 0; s < S; ++s) {
        w[s] = 0.0;
        bq[s] = 0.0;
        dnode[s] = 0;
    }
    if (n == 0) {
        status = ST_NO_NEIGHBOURS;
    } else if (!cell_weights<NMAX, G, MIXED, KIND, LTAB, PACK, FULL>(a, cell, t, n, ltab, ox, oxy, w, bq, wb, dnode,
                                                                     mu, status)) {
        return;  // queued for the FP64 form
    }
    if (status != ST_KRIGED) {
#pragma unroll
        for (int s = 0; s < S; ++s) w[s] = 0.0;
        mu = 0.0;
        wb = 0.0;
    }
    if (!MIXED) {
        double pwb = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) pwb += w[s] * bq[s];
        wb = gsum<G>(pwb);
    }

    if (KIND != 2) {
        const double mean = (KIND == 1) ? (a.mean_dev ? a.mean_dev[0] : a.mean) : 0.0;
        // A locally varying mean (lvm_kriging.cpp:16-39) gives the centre
        // and each neighbour its own mean; combiner.h:10-25 either way.
        const bool lvm = KIND == 1 && a.lvm_mean != nullptr;
        const float mf = lvm ? a.lvm_mean[cell] : (float)mean;
        float term[S];
        double pe = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool live = FULL || G * s + t < n;
            const long long node = cell + (live ? dnode[s] : 0);
            const float v0 = a.values[node];
            const float v = live ? v0 : 0.0f;
            const float ms = lvm ? (live ? a.lvm_mean[node] : 0.0f) : mf;
            term[s] = (float)(w[s] * (double)(v - ms));
            pe += w[s] * ((double)v - (double)ms);
        }
        float acc = mf;
#pragma unroll
        for (int i = 0; i < NMAX; ++i) {
            const float ti = gbc<G>(term[i / G], i % G);
            if (FULL || i < n) acc += ti;
        }
        const double e64 = (double)mf + gsum<G>(pe);
        const double var = a.c0 - wb - mu;
        if (t == 0) {
            if (status == ST_KRIGED) {
                if (a.out_val) a.out_val[cell] = acc;
                if (a.out_mask) a.out_mask[cell] = 1;
                if (a.out_var) a.out_var[cell] = var < 0 ? 0.0 : var;
                if (a.out_est64) a.out_est64[cell] = e64;
            } else if (KIND == 1) {
                // mean_on_failure (cont_kriging.h:80-94)
                if (a.out_val) a.out_val[cell] = mf;
                if (a.out_mask) a.out_mask[cell] = 1;
                if (a.out_var) a.out_var[cell] = -1.0;
                if (a.out_est64) a.out_est64[cell] = (double)mf;
            } else {
                // undefined_on_failure: the output keeps the input clone
                if (a.out_val) a.out_val[cell] = a.values[cell];
                if (a.out_mask) a.out_mask[cell] = a.mask[cell];
                if (a.out_var) a.out_var[cell] = -1.0;
                if (a.out_est64) a.out_est64[cell] = __builtin_nan("");
            }
            if (a.out_status) a.out_status[cell] = (unsigned char)status;
        }
        return;
    }

    // Median indicator kriging: one weight set, one combine per threshold
    // (median_ik.cpp:60-78, accumulated in double).  With categories
    // (indicator_kriging.h:88-161) indicator k is (data == cats[k]), the sum
    // is accumulated in float and lands in probability row cats[k].
    for (int k = 0; k < a.nind; ++k) {
        const float pf = (float)a.marg[k];
        const float thr = a.zfield ? a.thresholds[k] : 0.f;
        const int cat = a.has_cats ? a.cats[k] : -1;
        double term[S];
        double pe = 0.0;
#pragma unroll
        for (int s = 0; s < S; ++s) {
            const bool live = FULL || G * s + t < n;
            const long long node = cell + (live ? dnode[s] : 0);
            float iv;
            if (a.zfield)
                iv = (a.zfield[node] <= thr) ? 1.f : 0.f;
            else if (cat >= 0)
                iv = ((int)a.ind[node] == cat) ? 1.f : 0.f;
            else
                iv = (a.ind[(long long)k * g.n + node] == 1) ? 1.f : 0.f;
            iv = live ? iv : 0.f;
            term[s] = w[s] * (double)(iv - pf);
            pe += w[s] * ((double)iv - (double)pf);
        }
        double acc = (double)pf;
        float accf = pf;
        if (cat >= 0) {
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                const double ti = gbc<G>(term[i / G], i % G);
                if (FULL || i < n) accf += (float)ti;
            }
        } else {
#pragma unroll
            for (int i = 0; i < NMAX; ++i) {
                const double ti = gbc<G>(term[i / G], i % G);
                if (FULL || i < n) acc += ti;
            }
        }
        const double e64 = (double)pf + gsum<G>(pe);
        const long long row = (long long)(cat >= 0 ? cat : k) * g.n + cell;
        if (t == 0) {
            if (a.out_est64) a.out_est64[row] = e64;
            if (cat >= 0) {
                if (a.out_prob) a.out_prob[row] = accf;
            } else {
                if (a.out_class) a.out_class[row] = ((float)acc <= 0.5f) ? 0 : 1;
                if (a.out_prob) a.out_prob[row] = (float)acc;
            }
        }
    }
    if (t == 0) {
        if (a.order_relations && a.out_prob) {
            // indicator_kriging.h:34-59 on this cell's CDF (written above by
            // this lane): clamp to [0,1], then one forward averaging pass.
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

// ------------------------------------------------ count-bucketed work list
// The list (k_bucket_scatter) holds the cells to krige ordered by search path
// (template walk 0, cluster search 1), count bucket (largest first) and
// neighbour count (descending); each (path, bucket) segment starts at a
// multiple of SOLVE_CHUNK and gaps hold -1.  A wave takes one chunk of
// SOLVE_CHUNK consecutive entries, so all its cells share a bucket and it
// runs one register-solver size; when every cell of the chunk has exactly
// the bucket's size (and FULL_OK), the solver without liveness tests runs.
// Chunks are dealt to the waves round-robin, so every wave gets a share of
// each bucket.
template <int NB, int G, bool MIXED, int KIND, bool LTAB, bool PACK, bool FULL_OK>
__device__ __forceinline__ void solve_chunk(const SolveArgs& a, int i0, int iend, int lane, bool full,
                                            const double* ltab, unsigned ox, unsigned oxy) {
    constexpr int PER = 64 / G;  // cells per pass of the wave
    const int grp = lane / G, t = lane & (G - 1);
#pragma unroll
    for (int p = 0; p < SOLVE_CHUNK / PER; ++p) {
        const int idx = i0 + p * PER + grp;
        const int cell = idx < iend ? a.sorted[idx] : -1;
        if (cell < 0) continue;
        const int cnt = a.count[cell];
        if (FULL_OK && full)
            solve_cell<NB, G, MIXED, KIND, LTAB, PACK, true>(a, cell, cnt, t, ltab, ox, oxy);
        else
            solve_cell<NB, G, MIXED, KIND, LTAB, PACK, false>(a, cell, cnt, t, ltab, ox, oxy);
    }
}

// Chunks of [lo, hi) for one wave, strided over the grid's waves.  TOP = the
// largest bucket size the range holds (8, 16, 24 with 4 lanes per cell; 32
// with 8); the chunk's first cell has the chunk's largest count.
// With a.sched set, a wave's first chunk is its own and later ones come from
// that counter (dynamic dealing) instead of the fixed stride.
template <int TOP, bool MIXED, int KIND, bool LTAB, bool PACK>
__device__ __forceinline__ void solve_range(const SolveArgs& a, int lo, int hi, long long wave, long long nwaves,
                                            int lane, const double* ltab, unsigned ox, unsigned oxy) {
    for (long long c = lo / SOLVE_CHUNK + wave; c * SOLVE_CHUNK < hi;) {
        const int i0 = (int)(c * SOLVE_CHUNK);
        const int iend = min(i0 + SOLVE_CHUNK, hi);
        // a chunk starts on a cell: segments start at chunk multiples and a
        // gap never reaches the next one
        const int c0 = a.sorted[i0];
        const int n0 = __builtin_amdgcn_readfirstlane(c0 >= 0 ? count_n(a.count[c0]) : 0);
        const int nb = n0 <= 8 ? 8 : (n0 <= 16 ? 16 : (n0 <= 24 ? 24 : 32));
        const int idx = i0 + (lane & (SOLVE_CHUNK - 1));
        const int cell = idx < iend ? a.sorted[idx] : -1;
        const bool full = __all(cell < 0 || count_n(a.count[cell]) == nb);
        if constexpr (TOP == 32) {
            solve_chunk<32, 4, MIXED, KIND, LTAB, PACK, true>(a, i0, iend, lane, full, ltab, ox, oxy);
        } else {
            if (nb == 8)
                solve_chunk<8, 4, MIXED, KIND, LTAB, PACK, TOP == 8>(a, i0, iend, lane, full, ltab, ox, oxy);
            if constexpr (TOP >= 16)
                if (nb == 16)
                    solve_chunk<16, 4, MIXED, KIND, LTAB, PACK, TOP == 16>(a, i0, iend, lane, full, ltab, ox, oxy);
            if constexpr (TOP >= 24)
                if (nb == 24)
                    solve_chunk<24, 4, MIXED, KIND, LTAB, PACK, TOP == 24>(a, i0, iend, lane, full, ltab, ox, oxy);
        }
        if (a.sched) {
            int next = 0;
            if (lane == 0) next = atomicAdd(a.sched, 1);
            c = lo / SOLVE_CHUNK + nwaves + __shfl(next, 0);
        } else {
            c += nwaves;
        }
    }
}

// Range of the work list a launch covers: buckets 0-2 (TOP <= 24) or bucket
// 3 (TOP == 32) of one path.
template <int TOP>
__device__ __forceinline__ void sorted_range(const SolveArgs& a, int path, int& lo, int& hi) {
    const int* sg = a.seg + 8 * path;
    if (TOP == 32) {
        lo = sg[2 * 3];
        hi = sg[2 * 3 + 1];
    } else {
        lo = sg[2 * 2];
        hi = sg[2 * 0 + 1];
    }
}

// Persistent solve over one path's range with the octant covariance table
// staged in LDS (packed addresses).  Exits before staging when the range is
// empty.
template <int TOP, bool MIXED, int KIND, int BLOCK>
__global__ __launch_bounds__(BLOCK) void k_solve_sorted_lds(SolveArgs a, int nelem, int path) {
    extern __shared__ double otab_lds[];
    int lo, hi;
    sorted_range<TOP>(a, path, lo, hi);
    if (lo >= hi) return;
    for (int i = threadIdx.x; i < nelem; i += BLOCK) otab_lds[i] = a.otab[i];
    __syncthreads();
    const long long wave = ((long long)blockIdx.x * BLOCK + threadIdx.x) / 64;
    const long long nwaves = ((long long)gridDim.x * BLOCK) / 64;
    const unsigned ox = (unsigned)(a.ex + 1), oxy = ox * (unsigned)(a.ey + 1);
    solve_range<TOP, MIXED, KIND, true, true>(a, lo, hi, wave, nwaves, threadIdx.x & 63, otab_lds, ox, oxy);
}

// The same over global tables: the +-W octant table when OCT, else the full
// table.  The mixed form is held to 2 waves per SIMD, which it reaches
// without spilling (the FP64 form would spill).
template <int TOP, bool MIXED, int KIND, bool OCT>
__global__ __launch_bounds__(256, (MIXED && TOP != 32) ? 2 : 1) void k_solve_sorted(SolveArgs a, int path) {
    int lo, hi;
    sorted_range<TOP>(a, path, lo, hi);
    if (lo >= hi) return;
    const long long wave = ((long long)blockIdx.x * 256 + threadIdx.x) / 64;
    const long long nwaves = ((long long)gridDim.x * 256) / 64;
    const unsigned ox = (unsigned)(a.wx + 1), oxy = ox * (unsigned)(a.wy + 1);
    solve_range<TOP, MIXED, KIND, OCT, false>(a, lo, hi, wave, nwaves, threadIdx.x & 63, a.otab_w, ox, oxy);
}

// FP64 re-solve of the mixed form's fallbacks (an unordered list): each
// group of 4 lanes takes its cell's own bucket size, so a cell gets the FP64
// form's solver and bits.  B32: the 25-32 bucket, else the others.
template <bool B32, int KIND, bool OCT>
__global__ __launch_bounds__(256) void k_solve_fallback(SolveArgs a) {
    constexpr int G = 4;
    const int t = threadIdx.x & (G - 1);
    const long long gid = ((long long)blockIdx.x * blockDim.x + threadIdx.x) / G;
    const long long ngroups = ((long long)gridDim.x * blockDim.x) / G;
    const long long total = *a.n_list;
    const unsigned ox = (unsigned)(a.wx + 1), oxy = ox * (unsigned)(a.wy + 1);
    for (long long item = gid; item < total; item += ngroups) {
        const int cell = a.cell_list[item];
        const int cnt = a.count[cell];
        const int b = count_bucket(count_n(cnt));
        if constexpr (B32) {
            if (b == 3) solve_cell<32, 4, false, KIND, OCT>(a, cell, cnt, t, a.otab_w, ox, oxy);
        } else {
            if (b == 0) solve_cell<8, 4, false, KIND, OCT>(a, cell, cnt, t, a.otab_w, ox, oxy);
            if (b == 1) solve_cell<16, 4, false, KIND, OCT>(a, cell, cnt, t, a.otab_w, ox, oxy);
            if (b == 2) solve_cell<24, 4, false, KIND, OCT>(a, cell, cnt, t, a.otab_w, ox, oxy);
        }
    }
}

}  // namespace hpglx

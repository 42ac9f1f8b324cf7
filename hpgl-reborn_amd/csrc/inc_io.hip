// inc_io.hip -- INC / GSLIB property text on the GPU.
//
// Reading (read_inc_file.cpp:8-193): the reference walks the file with
// fscanf("%255s") token by token: a token starting with "--" drops the rest
// of its line, a token starting with '/' ends the data (an error before
// `size` values), any other token goes through sscanf("%f") (or "%d").  Here
// the data section is copied to HBM once and walked in parallel:
//   1. k_inc_summary: every lane summarises 32 bytes for both possible
//      states at its start (inside a comment or not): state after, live
//      number tokens, first '/' token, kind of the last live token.  These
//      summaries compose associatively, so a block scan gives each lane's
//      start state and token ordinal within the block's 8 KB chunk.
//   2. k_inc_chunks: one block scans the chunk summaries.
//   3. k_inc_emit: each lane re-walks its bytes with the known start state
//      and parses the number tokens it owns, one token per lane, straight
//      into the output arrays.
// Parsing is correctly rounded like glibc's strtof: exact float or double
// fast paths (one rounding of exact operands, with the double path's rare
// midpoint cases sent on), else a big-integer comparison of the decimal
// value with the binary midpoints around a close candidate.
//
// Writing (property_writer.cpp:18-139): every value's text length is known
// from its sign and class, so a scan of the lengths gives each value's
// offset and one lane per value prints "%E" digits exactly: the 7 significant
// digits come from a double product or quotient by an exact power of ten
// whose rounding is checked with an FMA residual, or from big integers
// outside the exact-power range.
#include "inc_io.h"

#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstdio>
#include <cstring>
#include <string>
#include <vector>

namespace hpglx {

namespace {

constexpr int IO_BLOCK = 256;
constexpr int IO_SEG = 32;                    // bytes per lane
constexpr int IO_CHUNK = IO_BLOCK * IO_SEG;   // bytes per block
constexpr int IO_PAD = 512;                   // spaces after the text
constexpr int MAX_TOKEN = 255;                // fscanf("%255s")

enum TokKind { TK_NONE = 0, TK_NUMBER = 1, TK_COMMENT = 2 };

// Outcome of walking some bytes from a given start state.
struct Branch {
    int out;     // inside a comment after the bytes
    int count;   // live number tokens started in them
    int slash;   // number tokens before the first live '/' token, -1 if none
    int last;    // TokKind of the last live token
};
struct Summ {
    Branch b[2];  // start state: 0 = not in a comment, 1 = in a comment
};

__host__ __device__ inline Summ summ_identity() {
    Summ s;
    for (int k = 0; k < 2; ++k) s.b[k] = Branch{k, 0, -1, TK_NONE};
    return s;
}

__host__ __device__ inline Branch pick(const Summ& B, int state) {
    Branch r;  // B.b[state] by selects, so the summaries stay in registers
    r.out = state ? B.b[1].out : B.b[0].out;
    r.count = state ? B.b[1].count : B.b[0].count;
    r.slash = state ? B.b[1].slash : B.b[0].slash;
    r.last = state ? B.b[1].last : B.b[0].last;
    return r;
}

// A then B.
__host__ __device__ inline Summ compose(const Summ& A, const Summ& B) {
    Summ r;
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const Branch& a = A.b[k];
        const Branch b = pick(B, a.out);
        r.b[k].out = b.out;
        r.b[k].count = a.count + b.count;
        r.b[k].slash = a.slash >= 0 ? a.slash : (b.slash >= 0 ? a.count + b.slash : -1);
        r.b[k].last = b.last != TK_NONE ? b.last : a.last;
    }
    return r;
}

__host__ __device__ inline bool is_space(unsigned c) {
    return c == ' ' || c == '\t' || c == '\n' || c == '\v' || c == '\f' || c == '\r';
}
__host__ __device__ inline bool is_digit(unsigned c) {
    return c >= '0' && c <= '9';
}

// The lane's 32 bytes as 8 dwords, plus the byte before and the byte after.
struct Seg {
    uint32_t w[IO_SEG / 4];
    unsigned prev, next;
    __device__ unsigned at(int j) const { return (w[j >> 2] >> (8 * (j & 3))) & 0xFFu; }
};

template <typename P>
__device__ inline Seg load_seg(P t, long long lo) {
    Seg s;
    const uint4* p = reinterpret_cast<const uint4*>(t + lo);
    const uint4 a = p[0], b = p[1];
    s.w[0] = a.x; s.w[1] = a.y; s.w[2] = a.z; s.w[3] = a.w;
    s.w[4] = b.x; s.w[5] = b.y; s.w[6] = b.z; s.w[7] = b.w;
    s.prev = lo > 0 ? t[lo - 1] : ' ';
    s.next = t[lo + IO_SEG];  // the buffer is padded with spaces
    return s;
}

// Per-byte classes of the segment as bit masks (bit j = byte j).
struct Masks {
    uint32_t ts;     // token starts: a non-space byte after a space
    uint32_t cs;     // token starts of "--" (comment candidates)
    uint32_t slash;  // token starts of '/'
    uint32_t nl;     // newlines
};

__device__ inline Masks seg_masks(const Seg& g) {
    uint32_t sp = 0, nl = 0, dash = 0, sl = 0;
#pragma unroll
    for (int j = 0; j < IO_SEG; ++j) {
        const uint32_t c = g.at(j), bit = 1u << j;
        sp |= (c == ' ' || c - 9u < 5u) ? bit : 0u;
        nl |= c == '\n' ? bit : 0u;
        dash |= c == '-' ? bit : 0u;
        sl |= c == '/' ? bit : 0u;
    }
    Masks m;
    m.ts = ~sp & ((sp << 1) | (is_space(g.prev) ? 1u : 0u));
    m.cs = m.ts & dash & ((dash >> 1) | (g.next == '-' ? 0x80000000u : 0u));
    m.slash = m.ts & sl;
    m.nl = nl;
    return m;
}

// The segment from start state `s`: a comment runs from a live "--" token to
// the next newline.  `numbers` receives the live number-token starts.
__device__ inline Branch walk_bits(const Masks& m, int s, uint32_t& numbers) {
    uint32_t dead = 0, live_cs = 0;
    int pos = 0;
    while (pos < IO_SEG) {
        const uint32_t from = ~0u << pos;
        if (s) {
            const uint32_t nl = m.nl & from;
            if (!nl) {
                dead |= from;
                break;
            }
            const int n = __ffs(nl) - 1;
            dead |= from & ((1u << n) - 1u);
            s = 0;
            pos = n + 1;
        } else {
            const uint32_t cs = m.cs & from;
            if (!cs) break;
            const int c = __ffs(cs) - 1;
            live_cs |= 1u << c;
            s = 1;
            pos = c;
        }
    }
    numbers = m.ts & ~dead & ~m.slash;
    Branch r;
    r.out = s;
    r.count = __popc(numbers);
    const uint32_t live_slash = m.slash & ~dead;
    r.slash = live_slash ? __popc(numbers & ((1u << (__ffs(live_slash) - 1)) - 1u)) : -1;
    const uint32_t ev = numbers | live_cs;
    r.last = ev ? (((numbers >> (31 - __clz(ev))) & 1u) ? TK_NUMBER : TK_COMMENT) : TK_NONE;
    return r;
}

__device__ inline Summ summarise(const Masks& m) {
    Summ s;
    uint32_t unused;
    s.b[0] = walk_bits(m, 0, unused);
    s.b[1] = walk_bits(m, 1, unused);
    return s;
}

__device__ inline Summ shfl_up_summ(const Summ& v, int d) {
    Summ r;
    for (int k = 0; k < 2; ++k) {
        r.b[k].out = __shfl_up(v.b[k].out, d, 64);
        r.b[k].count = __shfl_up(v.b[k].count, d, 64);
        r.b[k].slash = __shfl_up(v.b[k].slash, d, 64);
        r.b[k].last = __shfl_up(v.b[k].last, d, 64);
    }
    return r;
}

// Exclusive scan of the block's lane summaries (BLOCK threads); returns the
// lane's prefix and (through `total`) the block's composition.
template <int BLOCK>
__device__ inline Summ block_exclusive(const Summ& mine, Summ& total) {
    __shared__ Summ wave_tot[BLOCK / 64];
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    Summ inc = mine;
#pragma unroll
    for (int d = 1; d < 64; d <<= 1) {
        const Summ o = shfl_up_summ(inc, d);
        if (lane >= d) inc = compose(o, inc);
    }
    if (lane == 63) wave_tot[wave] = inc;
    __syncthreads();
    Summ before = summ_identity();
    total = summ_identity();
    for (int w = 0; w < BLOCK / 64; ++w) {
        const Summ wt = wave_tot[w];
        if (w < wave) before = compose(before, wt);
        total = compose(total, wt);
    }
    Summ excl = shfl_up_summ(inc, 1);
    if (lane == 0) excl = summ_identity();
    return compose(before, excl);
}

__global__ __launch_bounds__(IO_BLOCK) void k_inc_summary(const unsigned char* t, long long n, Summ* chunk_summ) {
    const long long lo = (long long)blockIdx.x * IO_CHUNK + (long long)threadIdx.x * IO_SEG;
    Summ mine = summ_identity();
    if (lo < n) mine = summarise(seg_masks(load_seg(t, lo)));
    Summ total;
    (void)block_exclusive<IO_BLOCK>(mine, total);
    if (threadIdx.x == 0) chunk_summ[blockIdx.x] = total;
}

// One block: exclusive prefix of every chunk from the state "not in a
// comment" at the start of the data, and the whole text's summary.
__global__ __launch_bounds__(1024) void k_inc_chunks(const Summ* chunk_summ, int nchunks, Branch* chunk_in,
                                                     Branch* whole) {
    const int t = threadIdx.x;
    const int per = (nchunks + 1023) / 1024;
    const int a = min(nchunks, t * per), b = min(nchunks, a + per);
    Summ acc = summ_identity();
    for (int c = a; c < b; ++c) acc = compose(acc, chunk_summ[c]);
    Summ total;
    Summ run = block_exclusive<1024>(acc, total);
    if (t == 0) *whole = total.b[0];
    for (int c = a; c < b; ++c) {
        chunk_in[c] = run.b[0];
        run = compose(run, chunk_summ[c]);
    }
}

// ------------------------------------------------------------ big integers
// Little-endian 32-bit limbs; 40 limbs hold every comparison the parser and
// the printer make (at most about 1,100 bits).
struct Big {
    static constexpr int L = 40;
    uint32_t w[L];
    int n;
};

__device__ inline void big_set(Big& a, uint64_t v) {
    for (int i = 0; i < Big::L; ++i) a.w[i] = 0;
    a.w[0] = (uint32_t)v;
    a.w[1] = (uint32_t)(v >> 32);
    a.n = a.w[1] ? 2 : (a.w[0] ? 1 : 0);
}

__device__ inline void big_mul_add(Big& a, uint32_t m, uint32_t add) {
    uint64_t carry = add;
    for (int i = 0; i < a.n; ++i) {
        const uint64_t v = (uint64_t)a.w[i] * m + carry;
        a.w[i] = (uint32_t)v;
        carry = v >> 32;
    }
    if (carry && a.n < Big::L) a.w[a.n++] = (uint32_t)carry;
    if (a.n == 0 && add) {
        a.w[0] = add;
        a.n = 1;
    }
}

__device__ inline void big_mul_pow5(Big& a, int e) {
    const uint32_t p5[14] = {1u, 5u, 25u, 125u, 625u, 3125u, 15625u, 78125u, 390625u, 1953125u, 9765625u,
                             48828125u, 244140625u, 1220703125u};
    while (e >= 13) {
        big_mul_add(a, p5[13], 0);
        e -= 13;
    }
    if (e > 0) big_mul_add(a, p5[e], 0);
}

__device__ inline void big_shl(Big& a, int bits) {
    if (a.n == 0 || bits <= 0) return;
    const int limbs = bits >> 5, r = bits & 31;
    int nn = a.n + limbs + 1;
    if (nn > Big::L) nn = Big::L;
    for (int i = nn - 1; i >= 0; --i) {
        const int src = i - limbs;
        uint32_t hi = (src >= 0 && src < a.n) ? a.w[src] : 0u;
        uint32_t lo = (src - 1 >= 0 && src - 1 < a.n) ? a.w[src - 1] : 0u;
        a.w[i] = r ? (hi << r) | (lo >> (32 - r)) : hi;
    }
    a.n = nn;
    while (a.n > 0 && a.w[a.n - 1] == 0) --a.n;
}

__device__ inline int big_cmp(const Big& a, const Big& b) {
    if (a.n != b.n) return a.n < b.n ? -1 : 1;
    for (int i = a.n - 1; i >= 0; --i)
        if (a.w[i] != b.w[i]) return a.w[i] < b.w[i] ? -1 : 1;
    return 0;
}

// a /= d, returns the remainder.
__device__ inline uint32_t big_div_small(Big& a, uint32_t d) {
    uint64_t rem = 0;
    for (int i = a.n - 1; i >= 0; --i) {
        const uint64_t cur = (rem << 32) | a.w[i];
        a.w[i] = (uint32_t)(cur / d);
        rem = cur % d;
    }
    while (a.n > 0 && a.w[a.n - 1] == 0) --a.n;
    return (uint32_t)rem;
}

// ----------------------------------------------------------------- parsing
__constant__ float kP10f[11] = {1e0f, 1e1f, 1e2f, 1e3f, 1e4f, 1e5f, 1e6f, 1e7f, 1e8f, 1e9f, 1e10f};
__constant__ double kP10d[23] = {1e0,  1e1,  1e2,  1e3,  1e4,  1e5,  1e6,  1e7,  1e8,  1e9,  1e10, 1e11,
                                 1e12, 1e13, 1e14, 1e15, 1e16, 1e17, 1e18, 1e19, 1e20, 1e21, 1e22};

// Decimal significand of a token: at most 200 significant digits as a big
// integer D, and x = D * 10^e10 (+ something when `sticky`).
struct DecDigits {
    Big D;
    long long e10;
    int nd;
    bool sticky;
};

// Scans digits [. digits] [e[+-]digits] from p[i]; false when no digit.
// The first 19 significant digits also go to *w (when non-null).
__device__ inline bool scan_decimal(const unsigned char* p, int i, int L, uint64_t* w19, int* nd19, bool* sticky19,
                                    long long* e19, DecDigits* full) {
    uint64_t w = 0;
    int nd = 0;
    bool sticky = false, any = false;
    long long e10 = 0;       // exponent for w
    long long e10f = 0;      // exponent for the 200-digit significand
    int ndf = 0;
    bool stickyf = false;
    uint32_t chunk = 0;
    int chunk_n = 0;
    if (full) {
        big_set(full->D, 0);
    }
    auto digit = [&](int d, bool frac) {
        any = true;
        if (nd == 0 && d == 0) {
            if (frac) {
                --e10;
                --e10f;
            }
            return;
        }
        if (nd < 19) {
            w = w * 10 + (uint64_t)d;
            ++nd;
            if (frac) --e10;
        } else {
            if (!frac) ++e10;
            sticky |= d != 0;
        }
        if (full) {
            if (ndf < 200) {
                chunk = chunk * 10 + (uint32_t)d;
                if (++chunk_n == 9) {
                    big_mul_add(full->D, 1000000000u, chunk);
                    chunk = 0;
                    chunk_n = 0;
                }
                ++ndf;
                if (frac) --e10f;
            } else {
                if (!frac) ++e10f;
                stickyf |= d != 0;
            }
        }
    };
    while (i < L && is_digit(p[i])) digit(p[i++] - '0', false);
    if (i < L && p[i] == '.') {
        ++i;
        while (i < L && is_digit(p[i])) digit(p[i++] - '0', true);
    }
    if (!any) return false;
    long long ex = 0;
    if (i < L && (p[i] | 0x20) == 'e') {
        int j = i + 1;
        bool eneg = false;
        if (j < L && (p[j] == '+' || p[j] == '-')) eneg = p[j++] == '-';
        if (j < L && is_digit(p[j])) {
            while (j < L && is_digit(p[j])) {
                if (ex < 1000000) ex = ex * 10 + (p[j] - '0');
                ++j;
            }
            if (eneg) ex = -ex;
        }
    }
    if (w19) *w19 = w;
    if (nd19) *nd19 = nd;
    if (sticky19) *sticky19 = sticky;
    if (e19) *e19 = e10 + ex;
    if (full) {
        if (chunk_n > 0) {
            uint32_t m = 1;
            for (int k = 0; k < chunk_n; ++k) m *= 10u;
            big_mul_add(full->D, m, chunk);
        }
        full->e10 = e10f + ex;
        full->nd = ndf;
        full->sticky = stickyf;
    }
    return true;
}

// Sign of x - K * 2^Q for x = D * 10^e10 (+ sticky).
__device__ inline int cmp_dec_bin(const DecDigits& x, uint32_t K, int Q) {
    Big A = x.D, B;
    big_set(B, K);
    long long shA = 0, shB = 0;
    if (x.e10 >= 0) {
        big_mul_pow5(A, (int)x.e10);
        shA = x.e10;
    } else {
        big_mul_pow5(B, (int)-x.e10);
        shB = -x.e10;
    }
    const long long net = (long long)Q + shB - shA;
    if (net >= 0)
        big_shl(B, (int)net);
    else
        big_shl(A, (int)-net);
    int c = big_cmp(A, B);
    if (c == 0 && x.sticky) c = 1;
    return c;
}

// Midpoints between the positive float `bits` and its neighbours, as K * 2^Q.
__device__ inline void mid_above(uint32_t bits, uint32_t& K, int& Q) {
    const uint32_t ef = bits >> 23, mant = bits & 0x7FFFFFu;
    const uint32_t M = ef ? (mant | 0x800000u) : mant;
    const int E = ef ? (int)ef - 150 : -149;
    K = 2 * M + 1;
    Q = E - 1;
}
__device__ inline void mid_below(uint32_t bits, uint32_t& K, int& Q) {
    const uint32_t ef = bits >> 23, mant = bits & 0x7FFFFFu;
    const uint32_t M = ef ? (mant | 0x800000u) : mant;
    const int E = ef ? (int)ef - 150 : -149;
    if (ef > 1 && mant == 0) {
        K = (1u << 25) - 1;
        Q = E - 2;
    } else {
        K = 2 * M - 1;
        Q = E - 1;
    }
}

// Correctly rounded float of a decimal token whose fast paths did not
// decide: candidate from a double estimate, then exact comparisons with the
// two neighbouring midpoints.
__device__ float decimal_exact(const unsigned char* p, int i, int L, uint64_t w, int nd, long long e10) {
    if (nd + e10 - 1 >= 39) return INFINITY;  // x >= 1e39
    if (nd + e10 <= -46) return 0.0f;         // x < 1e-46 < 2^-150
    DecDigits x;
    scan_decimal(p, i, L, nullptr, nullptr, nullptr, nullptr, &x);
    // (double)w * 10^e10 with e10 in [-65, 39]: relative error near 1e-16
    double pw = 1.0;
    long long e = e10;
    const double ten = 10.0;
    // 10^|e| by squaring from exact powers; far more accurate than needed
    double base = e >= 0 ? ten : 0.1;
    long long k = e >= 0 ? e : -e;
    if (k <= 22 && e >= 0) {
        pw = kP10d[k];
    } else if (k <= 22) {
        pw = 1.0 / kP10d[k];
    } else {
        double b = base;
        while (k) {
            if (k & 1) pw *= b;
            b *= b;
            k >>= 1;
        }
    }
    const double d = (double)w * pw;
    float f = (float)d;
    uint32_t bits = __float_as_uint(f);
    if (bits >= 0x7F800000u) bits = 0x7F7FFFFFu;  // estimate overflowed: start at FLT_MAX
    uint32_t K;
    int Q;
    mid_above(bits, K, Q);
    int c = cmp_dec_bin(x, K, Q);
    if (c > 0 || (c == 0 && (bits & 1u))) return __uint_as_float(bits + 1);
    if (bits > 0) {
        mid_below(bits, K, Q);
        c = cmp_dec_bin(x, K, Q);
        if (c < 0 || (c == 0 && (bits & 1u))) return __uint_as_float(bits - 1);
    }
    return __uint_as_float(bits);
}

// Rounds w * 2^e2 (+ sticky) to the nearest float, ties to even.
__device__ inline float round_bin(uint64_t w, long long e2, bool sticky) {
    if (w == 0) return 0.0f;
    const int b = 63 - __clzll((long long)w);
    const long long eb = b + e2;  // exponent of the leading bit
    if (eb > 127) return INFINITY;
    long long lsb = eb - 23;
    if (lsb < -149) lsb = -149;
    const long long shift = lsb - e2;
    uint64_t M;
    if (shift <= 0) {
        M = w << (-shift);
    } else if (shift >= 64) {
        const bool above = shift == 64 ? (w > (1ull << 63) || (w == (1ull << 63) && sticky)) : false;
        M = above ? 1 : 0;
    } else {
        M = w >> shift;
        const uint64_t rem = w & ((1ull << shift) - 1), half = 1ull << (shift - 1);
        if (rem > half || (rem == half && (sticky || (M & 1)))) ++M;
    }
    if (M == (1ull << 24)) {
        M >>= 1;
        ++lsb;
    }
    if (M == 0) return 0.0f;
    if (M >= (1ull << 23)) {
        const long long ef = lsb + 23 + 127;
        if (ef >= 255) return INFINITY;
        return __uint_as_float((uint32_t)(ef << 23) | (uint32_t)(M & 0x7FFFFFu));
    }
    return __uint_as_float((uint32_t)M);  // subnormal (lsb == -149)
}

__device__ inline int hexval(unsigned c) {
    if (c >= '0' && c <= '9') return (int)(c - '0');
    c |= 0x20;
    if (c >= 'a' && c <= 'f') return (int)(c - 'a' + 10);
    return -1;
}

// "0x" seen at p[i - 2]: hex digits [. hex digits] [p[+-]digits].
__device__ inline float parse_hex(const unsigned char* p, int i, int L) {
    uint64_t w = 0;
    int nd = 0;
    long long e2 = 0;
    bool sticky = false, any = false;
    auto hd = [&](int v, bool frac) {
        any = true;
        if (nd == 0 && v == 0) {
            if (frac) e2 -= 4;
            return;
        }
        if (nd < 15) {
            w = (w << 4) | (uint64_t)v;
            ++nd;
            if (frac) e2 -= 4;
        } else {
            if (!frac) e2 += 4;
            sticky |= v != 0;
        }
    };
    while (i < L && hexval(p[i]) >= 0) hd(hexval(p[i++]), false);
    if (i < L && p[i] == '.') {
        ++i;
        while (i < L && hexval(p[i]) >= 0) hd(hexval(p[i++]), true);
    }
    if (!any) return 0.0f;  // "0x." reads as 0
    if (i < L && (p[i] | 0x20) == 'p') {
        int j = i + 1;
        bool eneg = false;
        if (j < L && (p[j] == '+' || p[j] == '-')) eneg = p[j++] == '-';
        if (j < L && is_digit(p[j])) {
            long long ex = 0;
            while (j < L && is_digit(p[j])) {
                if (ex < 1000000) ex = ex * 10 + (p[j] - '0');
                ++j;
            }
            e2 += eneg ? -ex : ex;
        }
    }
    return round_bin(w, e2, sticky);
}

// sscanf(token, "%f"): 1 converted, 0 no conversion.
__device__ int parse_float_tok(const unsigned char* p, int L, float& out) {
    int i = 0;
    bool neg = false;
    if (i < L && (p[i] == '+' || p[i] == '-')) neg = p[i++] == '-';
    if (i >= L) return 0;
    const unsigned c0 = p[i] | 0x20;
    if (c0 == 'i' || c0 == 'n') {
        // glibc's scanf: "inf" may continue only as "infinity" (a partial
        // "infinit" fails); "nan" may be followed by anything, its payload
        // is not kept
        const char* word = c0 == 'i' ? "infinity" : "nan";
        if (i + 2 >= L || (p[i + 1] | 0x20) != (unsigned)word[1] || (p[i + 2] | 0x20) != (unsigned)word[2]) return 0;
        if (c0 == 'i' && i + 3 < L && (p[i + 3] | 0x20) == 'i') {
            for (int k = 3; k < 8; ++k)
                if (i + k >= L || (p[i + k] | 0x20) != (unsigned)word[k]) return 0;
        }
        const float v = c0 == 'i' ? INFINITY : __uint_as_float(0x7FC00000u);
        out = neg ? -v : v;
        return 1;
    }
    float f;
    if (p[i] == '0' && i + 1 < L && (p[i + 1] | 0x20) == 'x') {
        // "0x" must go on with a hex digit or a point ("0x." reads as 0)
        if (i + 2 >= L || (hexval(p[i + 2]) < 0 && p[i + 2] != '.')) return 0;
        f = parse_hex(p, i + 2, L);
    } else {
        uint64_t w;
        int nd;
        bool sticky;
        long long e10;
        if (!scan_decimal(p, i, L, &w, &nd, &sticky, &e10, nullptr)) return 0;
        if (nd == 0) {
            f = 0.0f;
        } else if (!sticky && w <= (1ull << 24) && e10 >= -10 && e10 <= 10) {
            // both operands exact: one IEEE rounding
            const float fw = (float)w;
            f = e10 >= 0 ? __fmul_rn(fw, kP10f[e10]) : __fdiv_rn(fw, kP10f[-e10]);
        } else if (!sticky && w < (1ull << 53) && e10 >= -22 && e10 <= 22) {
            // correctly rounded double; its float rounding is the true one
            // unless it landed exactly on a float midpoint
            const double dw = (double)w;
            const double d = e10 >= 0 ? __dmul_rn(dw, kP10d[e10]) : __ddiv_rn(dw, kP10d[-e10]);
            f = __double2float_rn(d);
            bool ambiguous = false;
            if (isinf(f)) {
                ambiguous = d == 0x1.ffffffp127;  // midpoint between FLT_MAX and 2^128
            } else if ((double)f != d) {
                const float g = nextafterf(f, d > (double)f ? INFINITY : 0.0f);
                ambiguous = !isinf(g) && d == ((double)f + (double)g) * 0.5;
            }
            if (ambiguous) f = decimal_exact(p, i, L, w, nd, e10);
        } else {
            f = decimal_exact(p, i, L, w, nd, e10);
        }
    }
    out = neg ? -f : f;
    return 1;
}

// sscanf(token, "%d") as glibc does it: strtol (saturating at the long
// range), then the long is stored to an int.
__device__ int parse_int_tok(const unsigned char* p, int L, int& out) {
    int i = 0;
    bool neg = false;
    if (i < L && (p[i] == '+' || p[i] == '-')) neg = p[i++] == '-';
    if (i >= L || !is_digit(p[i])) return 0;
    unsigned long long v = 0;
    bool ovf = false;
    while (i < L && is_digit(p[i])) {
        const unsigned d = p[i++] - '0';
        if (!ovf) {
            if (v > (0xFFFFFFFFFFFFFFFFull - d) / 10) ovf = true;
            else v = v * 10 + d;
        }
    }
    long long lv;
    if (neg)
        lv = (ovf || v > 9223372036854775808ull) ? (long long)0x8000000000000000ull
                                                  : (long long)(0ull - v);
    else
        lv = (ovf || v > 9223372036854775807ull) ? 0x7FFFFFFFFFFFFFFFll : (long long)v;
    out = (int)(uint32_t)(uint64_t)lv;
    return 1;
}

struct EmitArgs {
    const unsigned char* t;
    long long n;
    const Branch* chunk_in;
    const Branch* whole;
    int size;
    int as_int;
    float undef_f;
    int undef_i;
    float* out_f;
    unsigned char* out_b;
    unsigned char* mask;
    int* err_ord;   // smallest ordinal of a token that did not parse
    int* long_ord;  // smallest ordinal of a token longer than 255 characters
};

__device__ inline void emit_token(const EmitArgs& a, const unsigned char* p, int v) {
    int L = 0;
    while (L <= MAX_TOKEN && !is_space(p[L])) ++L;  // padded buffer: a space always follows
    if (L > MAX_TOKEN) {
        atomicMin(a.long_ord, v);
        return;
    }
    if (a.as_int) {
        int x;
        if (!parse_int_tok(p, L, x)) {
            atomicMin(a.err_ord, v);
            return;
        }
        a.out_b[v] = (unsigned char)x;
        a.mask[v] = x == a.undef_i ? 0 : 1;
    } else {
        float x;
        if (!parse_float_tok(p, L, x)) {
            atomicMin(a.err_ord, v);
            return;
        }
        a.out_f[v] = x;
        a.mask[v] = x == a.undef_f ? 0 : 1;
    }
}

// The chunk plus what a token starting in it can reach (256 characters and
// the byte after), staged in LDS.
constexpr int EMIT_STAGE = IO_CHUNK + 272;

__global__ __launch_bounds__(IO_BLOCK) void k_inc_emit(EmitArgs a) {
    __shared__ uint4 stage[EMIT_STAGE / 16];
    const long long base_byte = (long long)blockIdx.x * IO_CHUNK;
    const uint4* src = reinterpret_cast<const uint4*>(a.t + base_byte);
    for (int i = threadIdx.x; i < EMIT_STAGE / 16; i += IO_BLOCK) stage[i] = src[i];
    __syncthreads();
    const unsigned char* lds = reinterpret_cast<const unsigned char*>(stage);
    const int off = threadIdx.x * IO_SEG;
    const long long lo = base_byte + off;
    Masks msk{0u, 0u, 0u, 0u};
    Summ mine = summ_identity();
    if (lo < a.n) {
        Seg g;
        const uint4* p = reinterpret_cast<const uint4*>(lds + off);
        const uint4 x = p[0], y = p[1];
        g.w[0] = x.x; g.w[1] = x.y; g.w[2] = x.z; g.w[3] = x.w;
        g.w[4] = y.x; g.w[5] = y.y; g.w[6] = y.z; g.w[7] = y.w;
        g.prev = off > 0 ? lds[off - 1] : (base_byte > 0 ? a.t[base_byte - 1] : (unsigned)' ');
        g.next = lds[off + IO_SEG];
        msk = seg_masks(g);
        mine = summarise(msk);
    }
    Summ total;
    const Summ excl = block_exclusive<IO_BLOCK>(mine, total);
    if (lo >= a.n) return;
    const Branch cin = a.chunk_in[blockIdx.x];
    const Branch lane_in = pick(excl, cin.out);  // this lane's start, relative to the chunk
    const int state = lane_in.out;
    const int base = cin.count + lane_in.count;
    // values past the first '/' or past `size` are never stored
    const int slash = a.whole->slash;
    const int limit = slash >= 0 && slash < a.size ? slash : a.size;
    uint32_t starts;
    (void)walk_bits(msk, state, starts);
    int v = base;
    while (starts) {
        const int j = __ffs(starts) - 1;
        starts &= starts - 1;
        if (v < limit) emit_token(a, lds + off + j, v);
        ++v;
    }
}

// End of file before `size` values and no '/': fscanf fails and leaves the
// last token in its buffer, which the reference parses again for every
// remaining value (read_inc_file.cpp:55-60).
__global__ void k_inc_tail(const Branch* whole, int size, int as_int, float* out_f, unsigned char* out_b,
                           unsigned char* mask) {
    const int cnt = whole->count;
    if (whole->slash >= 0 || cnt >= size || cnt < 1 || whole->last != TK_NUMBER) return;
    for (long long v = (long long)cnt + blockIdx.x * (long long)blockDim.x + threadIdx.x; v < size;
         v += (long long)gridDim.x * blockDim.x) {
        if (as_int)
            out_b[v] = out_b[cnt - 1];
        else
            out_f[v] = out_f[cnt - 1];
        mask[v] = mask[cnt - 1];
    }
}

__global__ void k_inc_fill_space(unsigned char* p, long long n) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i < n) p[i] = ' ';
}

__global__ void k_remap_bytes(unsigned char* data, const unsigned char* mask, int size, const unsigned char* values,
                              int nv) {
    const long long i = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= size || mask[i] == 0) return;
    const unsigned char d = data[i];
    for (int j = 0; j < nv; ++j)
        if (d == values[j]) {
            data[i] = (unsigned char)j;
            return;
        }
}

// ---------------------------------------------------------------- printing
// Seven significant digits of a > 0 at decimal exponent E: N = a / 10^(E-6)
// rounded half to even, exactly.
__device__ uint32_t digits_at(double a, uint32_t fbits, int E) {
    const int k = 6 - E;
    if (k >= -22 && k <= 22) {
        double q, r;
        bool q_above;  // the double quotient is above the exact one
        if (k >= 0) {
            q = __dmul_rn(a, kP10d[k]);
            r = fma(a, kP10d[k], -q);  // exact q_true - q
            q_above = r < 0;
        } else {
            q = __ddiv_rn(a, kP10d[-k]);
            r = fma(q, kP10d[-k], -a);  // sign of q - q_true
            q_above = r > 0;
        }
        const double fl = floor(q), frac = q - fl;
        uint32_t N = (uint32_t)fl;
        if (frac > 0.5 || (frac == 0.5 && r != 0.0 && !q_above) || (frac == 0.5 && r == 0.0 && (N & 1u))) ++N;
        return N;
    }
    // a = M * 2^e exactly
    const uint32_t ef = fbits >> 23, mant = fbits & 0x7FFFFFu;
    const uint64_t M = ef ? (mant | 0x800000u) : mant;
    const int e = ef ? (int)ef - 150 : -149;
    Big B;
    big_set(B, M);
    if (k > 0) {
        // q = M * 5^k * 2^(k + e)
        big_mul_pow5(B, k);
        const int sh = k + e;
        if (sh >= 0) {
            big_shl(B, sh);
            return B.n ? B.w[0] : 0u;
        }
        // shift right by -sh with round half to even
        const int s = -sh;
        Big H;  // 2^(s-1)
        big_set(H, 1);
        big_shl(H, s - 1);
        // quotient and remainder: q = B >> s, rem = B mod 2^s
        Big Qt = B;
        // right shift
        {
            const int limbs = s >> 5, r = s & 31;
            for (int i = 0; i < Big::L; ++i) {
                const int src = i + limbs;
                const uint32_t lo = src < Big::L ? Qt.w[src] : 0u;
                const uint32_t hi = src + 1 < Big::L ? Qt.w[src + 1] : 0u;
                Qt.w[i] = r ? (lo >> r) | (hi << (32 - r)) : lo;
            }
            Qt.n = Big::L;
            while (Qt.n > 0 && Qt.w[Qt.n - 1] == 0) --Qt.n;
        }
        Big back = Qt;
        big_shl(back, s);
        // rem = B - back (B >= back); compare rem with H via B vs back + H
        Big bh = back;
        {
            uint64_t carry = 0;
            const int nn = max(bh.n, H.n) + 1;
            for (int i = 0; i < nn && i < Big::L; ++i) {
                const uint64_t v = (uint64_t)(i < bh.n ? bh.w[i] : 0u) + (i < H.n ? H.w[i] : 0u) + carry;
                bh.w[i] = (uint32_t)v;
                carry = v >> 32;
            }
            bh.n = min(nn, Big::L);
            while (bh.n > 0 && bh.w[bh.n - 1] == 0) --bh.n;
        }
        uint32_t N = Qt.n ? Qt.w[0] : 0u;
        const int c = big_cmp(B, bh);  // rem vs half
        if (c > 0 || (c == 0 && (N & 1u))) ++N;
        return N;
    }
    // k < -22: a is an integer; q = a / 10^(-k)
    big_shl(B, e);
    uint32_t first = 0;
    bool rest = false;
    for (int i = 0; i < -k; ++i) {
        const uint32_t d = big_div_small(B, 10);
        if (i == -k - 1)
            first = d;
        else
            rest |= d != 0;
    }
    uint32_t N = B.n ? B.w[0] : 0u;
    if (first > 5 || (first == 5 && (rest || (N & 1u)))) ++N;
    return N;
}

// printf("%E", (double)v); returns the length.
__device__ int format_E(float v, char* s) {
    const uint32_t bits = __float_as_uint(v);
    const bool neg = bits >> 31;
    int n = 0;
    if (neg) s[n++] = '-';
    if (isnan(v) || isinf(v)) {
        s[n++] = isnan(v) ? 'N' : 'I';
        s[n++] = isnan(v) ? 'A' : 'N';
        s[n++] = isnan(v) ? 'N' : 'F';
        return n;
    }
    const uint32_t ab = bits & 0x7FFFFFFFu;
    uint32_t N = 0;
    int E = 0;
    if (ab != 0) {
        const double a = (double)__uint_as_float(ab);
        E = (int)floor(log10(a));
        for (int it = 0; it < 3; ++it) {
            N = digits_at(a, ab, E);
            if (N >= 10000000u)
                ++E;
            else if (N < 1000000u)
                --E;
            else
                break;
        }
    }
    char d[7];
    for (int k = 6; k >= 0; --k) {
        d[k] = (char)('0' + N % 10);
        N /= 10;
    }
    s[n++] = d[0];
    s[n++] = '.';
    for (int k = 1; k < 7; ++k) s[n++] = d[k];
    s[n++] = 'E';
    s[n++] = E < 0 ? '-' : '+';
    int ae = E < 0 ? -E : E;
    if (ae >= 100) {
        s[n++] = (char)('0' + ae / 100);
        ae %= 100;
    }
    s[n++] = (char)('0' + ae / 10);
    s[n++] = (char)('0' + ae % 10);
    return n;
}

// printf("%d", (int)(char)b) -- char is signed on this ABI.
__device__ inline int format_byte(unsigned char b, char* s) {
    int v = (int)(signed char)b;
    int n = 0;
    if (v < 0) {
        s[n++] = '-';
        v = -v;
    }
    if (v >= 100) s[n++] = (char)('0' + v / 100);
    if (v >= 10) s[n++] = (char)('0' + (v / 10) % 10);
    s[n++] = (char)('0' + v % 10);
    return n;
}

struct FmtArgs {
    const float* f;
    const unsigned char* b;
    const unsigned char* mask;
    int size;
    float undef_f;
    unsigned char undef_b;
    const unsigned char* remap;
    int remap_count;
};

__device__ inline int fmt_value(const FmtArgs& a, long long i, char* s) {
    const bool informed = a.mask[i] != 0;
    if (a.f) return format_E(informed ? a.f[i] : a.undef_f, s);
    unsigned char v = a.undef_b;
    if (informed) {
        const unsigned char d = a.b[i];
        v = d < a.remap_count ? a.remap[d] : d;
    }
    return format_byte(v, s);
}

// Length of the value's line, newline included.
__device__ inline int fmt_len(const FmtArgs& a, long long i) {
    const bool informed = a.mask[i] != 0;
    if (a.f) {
        const float v = informed ? a.f[i] : a.undef_f;
        const int sign = (__float_as_uint(v) >> 31) ? 1 : 0;
        return (isnan(v) || isinf(v) ? 3 : 12) + sign + 1;
    }
    unsigned char b = a.undef_b;
    if (informed) {
        const unsigned char d = a.b[i];
        b = d < a.remap_count ? a.remap[d] : d;
    }
    const int v = (int)(signed char)b, av = v < 0 ? -v : v;
    return (v < 0 ? 1 : 0) + (av >= 100 ? 3 : (av >= 10 ? 2 : 1)) + 1;
}

constexpr int SCAN_BLOCK = 1024;

__global__ __launch_bounds__(SCAN_BLOCK) void k_fmt_len(FmtArgs a, long long* len, long long* block_sum) {
    __shared__ long long part[SCAN_BLOCK];
    const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    const long long l = i < a.size ? fmt_len(a, i) : 0;
    part[threadIdx.x] = l;
    __syncthreads();
    for (int d = 1; d < SCAN_BLOCK; d <<= 1) {
        const long long o = threadIdx.x >= (unsigned)d ? part[threadIdx.x - d] : 0;
        __syncthreads();
        part[threadIdx.x] += o;
        __syncthreads();
    }
    if (i < a.size) len[i] = part[threadIdx.x] - l;  // exclusive within the block
    if (threadIdx.x == SCAN_BLOCK - 1) block_sum[blockIdx.x] = part[threadIdx.x];
}

__global__ __launch_bounds__(1024) void k_scan_blocks(long long* block_sum, int nblocks, long long* total) {
    __shared__ long long part[1024];
    const int t = threadIdx.x;
    const int per = (nblocks + 1023) / 1024;
    const int a = min(nblocks, t * per), b = min(nblocks, a + per);
    long long acc = 0;
    for (int c = a; c < b; ++c) acc += block_sum[c];
    part[t] = acc;
    __syncthreads();
    if (t == 0) {
        long long run = 0;
        for (int i = 0; i < 1024; ++i) {
            const long long x = part[i];
            part[i] = run;
            run += x;
        }
        *total = run;
    }
    __syncthreads();
    long long run = part[t];
    for (int c = a; c < b; ++c) {
        const long long x = block_sum[c];
        block_sum[c] = run;
        run += x;
    }
}

// Each lane prints its line into LDS at its block-local offset; the block
// then copies its text out with consecutive lanes on consecutive bytes.
__global__ __launch_bounds__(SCAN_BLOCK) void k_fmt_write(FmtArgs a, const long long* len, const long long* block_off,
                                                          char* out) {
    __shared__ char buf[SCAN_BLOCK * 16];
    const long long i = (long long)blockIdx.x * SCAN_BLOCK + threadIdx.x;
    if (i < a.size) {
        char* s = buf + len[i];
        const int n = fmt_value(a, i, s);
        s[n] = '\n';
    }
    __syncthreads();
    const long long o = block_off[blockIdx.x];
    const int nbytes = (int)(block_off[blockIdx.x + 1] - o);
    for (int k = threadIdx.x; k < nbytes; k += SCAN_BLOCK) out[o + k] = buf[k];
}

bool host_isalpha(unsigned c) {
    return (c >= 'A' && c <= 'Z') || (c >= 'a' && c <= 'z');
}

}  // namespace

void IncText::release() {
    if (host) (void)hipHostFree(host);
    host = nullptr;
    host_cap = 0;
}

long long inc_load_file(IncText& io, const char* filename, std::string& err) {
    FILE* f = std::fopen(filename, "rb");
    if (!f) {
        err = std::string("Error opening file:") + filename + ".";
        return -1;
    }
    std::fseek(f, 0, SEEK_END);
    const long long n = std::ftell(f);
    std::fseek(f, 0, SEEK_SET);
    if (n < 0) {
        std::fclose(f);
        err = "Error reading file.";
        return -1;
    }
    const size_t need = (size_t)n + IO_PAD;
    if (need > io.host_cap) {
        io.release();
        if (hipHostMalloc((void**)&io.host, need, hipHostMallocDefault) != hipSuccess) {
            io.host = nullptr;
            std::fclose(f);
            err = "hipHostMalloc failed";
            return -1;
        }
        io.host_cap = need;
    }
    const size_t got = n ? std::fread(io.host, 1, (size_t)n, f) : 0;
    std::fclose(f);
    if ((long long)got != n) {
        err = "Error reading file.";
        return -1;
    }
    return n;
}

long long inc_data_offset(const char* text, long long n, std::string& name, std::string& err) {
    // fscanf(file, "%1023[^\n]\n", line): up to 1023 characters other than
    // '\n', then any run of whitespace; a line whose first character is not
    // a letter is skipped (with its 1023-character continuations).
    long long pos = 0;
    auto read_line = [&](std::string& line) -> bool {
        line.clear();
        while (pos < n && text[pos] != '\n' && line.size() < 1023) line.push_back(text[pos++]);
        if (line.empty()) return false;
        while (pos < n && is_space((unsigned char)text[pos])) ++pos;
        return true;
    };
    std::string line;
    for (;;) {
        // a leading blank line reads as an empty match: skip it
        while (pos < n && is_space((unsigned char)text[pos])) ++pos;
        if (!read_line(line)) {
            err = "Property name not found.";
            return -1;
        }
        const bool named = host_isalpha((unsigned char)line[0]);
        if (named) name = line;
        std::string more;
        while (line.size() == 1023) {
            if (!read_line(more)) break;
            if (named) name += more;
            line = more;
        }
        if (named) return pos;
    }
}

int inc_parse(const char* text, long long n, int size, int as_int, float undef_f, int undef_i, float* d_f,
              unsigned char* d_b, unsigned char* d_mask, Workspace& ws, hipStream_t st, std::string& err) {
    if (size <= 0) return 0;
    const long long nchunks = (n + IO_CHUNK - 1) / IO_CHUNK;
    if (nchunks > 0x7FFFFFFFLL) {
        err = "file too large";
        return -1;
    }
    const long long padded = nchunks * IO_CHUNK + IO_PAD;
    auto* d_t = (unsigned char*)ws.get("inc_text", (size_t)padded, err);
    auto* d_cs = (Summ*)ws.get("inc_chunk_summ", (size_t)std::max<long long>(1, nchunks) * sizeof(Summ), err);
    auto* d_ci = (Branch*)ws.get("inc_chunk_in", (size_t)std::max<long long>(1, nchunks) * sizeof(Branch), err);
    auto* d_misc = (int*)ws.get("inc_misc", 64, err);
    if (!err.empty()) return -1;
    Branch* d_whole = reinterpret_cast<Branch*>(d_misc);
    int* d_err = d_misc + 8;
    int* d_long = d_misc + 9;
    const int init[2] = {0x7FFFFFFF, 0x7FFFFFFF};
    if (hipMemcpyAsync(d_t, text, (size_t)n, hipMemcpyHostToDevice, st) != hipSuccess ||
        hipMemcpyAsync(d_err, init, sizeof init, hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "host to device copy failed";
        return -1;
    }
    const long long npad = padded - n;
    hipLaunchKernelGGL(k_inc_fill_space, dim3((unsigned)((npad + 255) / 256)), dim3(256), 0, st, d_t + n, npad);
    if (nchunks > 0) {
        hipLaunchKernelGGL(k_inc_summary, dim3((unsigned)nchunks), dim3(IO_BLOCK), 0, st, (const unsigned char*)d_t,
                           n, d_cs);
    }
    hipLaunchKernelGGL(k_inc_chunks, dim3(1), dim3(1024), 0, st, (const Summ*)d_cs, (int)nchunks, d_ci, d_whole);
    EmitArgs a;
    a.t = d_t;
    a.n = n;
    a.chunk_in = d_ci;
    a.whole = d_whole;
    a.size = size;
    a.as_int = as_int;
    a.undef_f = undef_f;
    a.undef_i = undef_i;
    a.out_f = d_f;
    a.out_b = d_b;
    a.mask = d_mask;
    a.err_ord = d_err;
    a.long_ord = d_long;
    if (nchunks > 0) hipLaunchKernelGGL(k_inc_emit, dim3((unsigned)nchunks), dim3(IO_BLOCK), 0, st, a);
    hipLaunchKernelGGL(k_inc_tail, dim3(256), dim3(256), 0, st, (const Branch*)d_whole, size, as_int, d_f, d_b,
                       d_mask);
    int h[10];
    if (hipGetLastError() != hipSuccess || hipMemcpyAsync(h, d_misc, sizeof h, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "inc parse kernels failed";
        return -1;
    }
    Branch whole;
    std::memcpy(&whole, h, sizeof whole);
    const int bad = std::min(h[8], h[9]);
    const int limit = whole.slash >= 0 && whole.slash < size ? whole.slash : size;
    if (bad < limit) {
        // the token's text, found by walking the host copy (error path only)
        long long pos = 0;
        int v = -1, comment = 0;
        std::string tok;
        for (long long i = 0; i < n; ++i) {
            const unsigned c = (unsigned char)text[i];
            if (c == '\n') {
                comment = 0;
                continue;
            }
            if (is_space(c) || comment) continue;
            if (i > 0 && !is_space((unsigned char)text[i - 1])) continue;
            if (c == '-' && i + 1 < n && text[i + 1] == '-') {
                comment = 1;
                continue;
            }
            if (c == '/') continue;
            if (++v == bad) {
                pos = i;
                break;
            }
        }
        while (pos < n && !is_space((unsigned char)text[pos]) && tok.size() < 255) tok.push_back(text[pos++]);
        if (bad == h[9] && h[9] <= h[8])
            err = "token longer than 255 characters at value " + std::to_string(bad) + " (not supported)";
        else
            err = "Error parsing '" + tok + "' string.";
        return -1;
    }
    if (whole.slash >= 0 && whole.slash < size) {
        err = "Unexpected end of data.";
        return -1;
    }
    if (whole.count < size && (whole.count < 1 || whole.last != TK_NUMBER)) {
        // the reference loops forever (or reads an uninitialised buffer) here
        err = "Unexpected end of file.";
        return -1;
    }
    return 0;
}

int inc_remap_bytes(unsigned char* d_data, const unsigned char* d_mask, int size, const unsigned char* values,
                    int values_count, Workspace& ws, hipStream_t st, std::string& err) {
    if (size <= 0 || values_count <= 0) return 0;
    auto* d_v = (unsigned char*)ws.get("inc_remap", (size_t)values_count, err);
    if (!d_v) return -1;
    if (hipMemcpyAsync(d_v, values, (size_t)values_count, hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "host to device copy failed";
        return -1;
    }
    hipLaunchKernelGGL(k_remap_bytes, dim3((unsigned)((size + 255) / 256)), dim3(256), 0, st, d_data, d_mask, size,
                       (const unsigned char*)d_v, values_count);
    if (hipGetLastError() != hipSuccess) {
        err = "remap launch failed";
        return -1;
    }
    return 0;
}

long long inc_format(const float* d_f, const unsigned char* d_b, const unsigned char* d_mask, int size, float undef_f,
                     unsigned char undef_b, const unsigned char* remap, int remap_count, IncText& io, Workspace& ws,
                     hipStream_t st, std::string& err) {
    if (size <= 0) return 0;
    const long long nblocks = (size + SCAN_BLOCK - 1) / SCAN_BLOCK;
    auto* d_len = (long long*)ws.get("fmt_len", (size_t)size * 8, err);
    auto* d_bs = (long long*)ws.get("fmt_block", (size_t)nblocks * 8 + 8, err);
    auto* d_remap = (unsigned char*)ws.get("fmt_remap", 256, err);
    const size_t max_bytes = (size_t)size * 16;
    auto* d_out = (char*)ws.get("fmt_out", max_bytes, err);
    if (!err.empty()) return -1;
    FmtArgs a;
    a.f = d_f;
    a.b = d_b;
    a.mask = d_mask;
    a.size = size;
    a.undef_f = undef_f;
    a.undef_b = undef_b;
    a.remap = d_remap;
    a.remap_count = remap ? std::min(remap_count, 256) : 0;
    if (a.remap_count > 0 &&
        hipMemcpyAsync(d_remap, remap, (size_t)a.remap_count, hipMemcpyHostToDevice, st) != hipSuccess) {
        err = "host to device copy failed";
        return -1;
    }
    hipLaunchKernelGGL(k_fmt_len, dim3((unsigned)nblocks), dim3(SCAN_BLOCK), 0, st, a, d_len, d_bs);
    hipLaunchKernelGGL(k_scan_blocks, dim3(1), dim3(1024), 0, st, d_bs, (int)nblocks, d_bs + nblocks);
    hipLaunchKernelGGL(k_fmt_write, dim3((unsigned)nblocks), dim3(SCAN_BLOCK), 0, st, a, (const long long*)d_len,
                       (const long long*)d_bs, d_out);
    long long total = 0;
    if (hipGetLastError() != hipSuccess ||
        hipMemcpyAsync(&total, d_bs + nblocks, 8, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "format kernels failed";
        return -1;
    }
    if ((size_t)total > io.host_cap) {
        io.release();
        if (hipHostMalloc((void**)&io.host, (size_t)total, hipHostMallocDefault) != hipSuccess) {
            io.host = nullptr;
            err = "hipHostMalloc failed";
            return -1;
        }
        io.host_cap = (size_t)total;
    }
    if (hipMemcpyAsync(io.host, d_out, (size_t)total, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess) {
        err = "device to host copy failed";
        return -1;
    }
    return total;
}

}  // namespace hpglx

// Host-side logic of the engine: length selection, execution plan, carry bound,
// conversions between the reference's canonical digits and the device representation.
// Header-only; shared by the C ABI (lnt_api.hip) and the CPU emulation tests.
#pragma once
#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "gl64.hpp"

namespace lnt {

typedef unsigned __int128 u128;
typedef __int128 i128;

// NttLengths::GetNttSecuredLength (ntt_lengths.cu:88-98) in exact integer arithmetic:
// L = ceil(log2(p / 32)); w = ceil(p / 2^L); if (2^w - 1)^2 >= P / 2^(L+1): L++.
// The double-precision original agrees for every p >= 33 (checked against the oracle on
// a dense sweep in tests/test_host.py).
inline uint32_t secured_log2_length(uint32_t p) {
    if (p <= 32) return 0;
    uint32_t L = 0;
    while ((32ull << L) < p) L++;
    uint64_t w = ((uint64_t)p + (1ull << L) - 1) >> L;
    u128 m = (w >= 64) ? ~(u128)0 : (((u128)1 << w) - 1);
    u128 lhs = m * m;
    // m^2 >= P / 2^(L+1)  <=>  m^2 * 2^(L+1) >= P   (real division)
    bool bump = (lhs >> (127 - (L + 1))) != 0 ? true : ((lhs << (L + 1)) >= (u128)GL_P);
    if (bump) L++;
    return L;
}

inline uint32_t ilog2(uint64_t v) {
    uint32_t l = 0;
    while ((1ull << l) < v) l++;
    return l;
}

// Line shape for a line of 2^a words: B1 = 2^ceil(a/2), B2 = 2^floor(a/2).
struct Line {
    uint32_t B1 = 0, B2 = 0;
};
inline Line line_for(uint32_t a) {
    Line l;
    l.B1 = 1u << ((a + 1) / 2);
    l.B2 = 1u << (a / 2);
    return l;
}

struct Plan {
    bool small = false;     // single-workgroup path (n <= 4096)
    uint32_t n = 0, log_n = 0;
    uint32_t C = 0, R = 0;  // large path: row length, row count
    Line row, col;          // row line (also the small path's line), column line
    uint32_t K = 0;         // columns per column tile
    uint32_t G = 0;         // rows per row workgroup
    uint32_t row3 = 0;      // 3-phase rows: C = row3^3 (0: 2-phase rows)
    uint32_t row3m = 0;     // mixed-radix 3-phase rows: C = row3m * rBB * rBC (0: none)
    uint32_t rBB = 0, rBC = 0;
    uint32_t col3 = 0;      // 3-phase columns: first radix BA (0: 2-phase columns)
    uint32_t cBB = 0, cBC = 0;  // 3-phase columns: second and third radix
    uint32_t M = 0;         // carry halo (compile-time), >= levels
    uint32_t levels = 0;    // carry levels
    bool bias = false;      // large path: every carried digit |d| < 2^31, whatever the input
};

// Two plans run the same kernel instantiations over the same table layouts (only the
// exponent-dependent carry level count may differ).
inline bool same_kernels(const Plan &a, const Plan &b) {
    return a.small == b.small && a.n == b.n && a.C == b.C && a.R == b.R && a.row.B1 == b.row.B1 &&
           a.row.B2 == b.row.B2 && a.col.B1 == b.col.B1 && a.col.B2 == b.col.B2 && a.K == b.K && a.G == b.G &&
           a.row3 == b.row3 && a.row3m == b.row3m && a.rBB == b.rBB && a.rBC == b.rBC && a.col3 == b.col3 &&
           a.cBB == b.cBB && a.cBC == b.cBC && a.M == b.M;
}

// Two plans order the transform domain the same way (the column tile width, rows per
// workgroup and carry halo do not change where a frequency is stored).
inline bool same_slots(const Plan &a, const Plan &b) {
    Plan x = a;
    x.K = b.K;
    x.G = b.G;
    x.M = b.M;
    return same_kernels(x, b);
}

// Carry levels m such that digits stay bounded by D with every pre-carry value below P/2
// in steady state, in exact integer arithmetic.  0 if no m <= 8 works.
//  * Digits |d| <= D give convolution outputs |C| <= 2 n D^2 (the IBDWT factors
//    2^(e_i + e_k - e_j - p t) are 1 or 2), plus 2 for the LL "-2": V_0 = 2 n D^2 + 2,
//    which must stay <= (P - 1) / 2 to be read back as a signed value.
//  * A balanced split at width w >= wmin leaves |lo| <= 2^(w-1) and a carry
//    |h| <= V / 2^w + 1/2, so |h| <= floor(V / 2^wmin) + 1 and one level maps V to
//    2^(wmax-1) + floor(V / 2^wmin) + 1.
//  * m levels work when, from D_0 = 2^(wmax-1) + 2, iterating D <- V_m + 1 reaches a D
//    that V_m + 1 does not exceed (a fixed point of the bound).
inline uint32_t carry_levels(uint32_t p, uint32_t n) {
    const uint32_t L = ilog2(n);
    const uint32_t wmin = p >> L, wmax = (p + n - 1) >> L;
    if (wmin == 0 || wmax > 62) return 0;
    const u128 half_p = (u128)((GL_P - 1) / 2);
    for (uint32_t m = 1; m <= 8; m++) {
        u128 D = ((u128)1 << (wmax - 1)) + 2;
        for (int round = 0; round < 200; round++) {
            if (D >> 40) break;  // 2 n D^2 would not fit below P/2 anyway
            const u128 V0 = 2 * (u128)n * D * D + 2;
            if (V0 > half_p) break;
            u128 V = V0;
            for (uint32_t l = 0; l < m; l++) V = ((u128)1 << (wmax - 1)) + (V >> wmin) + 1;
            const u128 Dn = V + 1;
            if (Dn <= D) return m;
            D = Dn;
        }
    }
    return 0;
}

// After `levels` carry levels, is every digit below 2^31 in magnitude for any input value
// |v| < 2^63 (a field element read as a signed integer, corrupted ones included)?  Each level
// maps a bound V to 2^(wmax-1) + floor(V / 2^wmin) + 1.  Then d + 2^31 fits 32 bits and the
// weight multiply can take the biased digit (gl_mul_biased).
inline bool digits_fit_31(uint32_t p, uint32_t n, uint32_t levels) {
    const uint32_t L = ilog2(n);
    const uint32_t wmin = p >> L, wmax = (p + n - 1) >> L;
    if (wmin == 0 || wmax > 32) return false;
    uint64_t V = 1ull << 63;
    for (uint32_t l = 0; l < levels; l++) V = (1ull << (wmax - 1)) + (V >> wmin) + 1;
    return V < (1ull << 31);
}

// Row/column split of a large residue: log2 C and log2 R.
inline bool split_for(uint32_t log_n, uint32_t &lc, uint32_t &lr) {
    if (log_n < 13 || log_n > 26) return false;
    // rows of 4096 words (3-phase row kernel) from n = 2^16 on: measured 1.2-2x faster than
    // the square-ish splits with 2-phase rows of 256 .. 2048 words.  From 2^24 the rows
    // are 8192 words (2^24: 0.550 against 0.581 ms per iteration with 4096 x 4096), and
    // 16384 at 2^26 (mixed-radix 3-phase rows), over columns of at most 4096 words
    lc = log_n >= 24 ? (log_n == 26 ? 14 : 13) : log_n >= 16 ? 12 : (log_n + 1) / 2;
    lr = log_n - lc;
    return lr <= 12 && lr >= 4;
}

// Choices the tuner may vary (0 = default).
struct PlanOpts {
    uint32_t lc = 0;    // log2 of the row length
    uint32_t K = 0;     // column tile width
    int row3 = -1;      // 1: 3-phase rows when C is 512 or 4096, 0: 2-phase, -1: default
                        // (rows of 8192 and 16384 words are always 3-phase: 1 and 2
                        // select 32 or 16 words per thread)
    int col3 = -1;      // 1: 3-phase columns (R = 512 .. 4096), 0: 2-phase, -1: default
};

// Column tile widths of the 3-phase column kernels.
inline bool col3_tile_ok(uint32_t R, uint32_t K) {
    static const uint32_t shapes[][2] = {{512, 16}, {512, 8}, {1024, 16}, {1024, 8}, {2048, 8}, {2048, 4}, {4096, 4}};
    for (auto &s : shapes)
        if (s[0] == R && s[1] == K) return true;
    return false;
}

// Column tile widths that fit the column kernels' LDS (R * K words of 8 bytes <= 128 KiB).
inline bool col_tile_ok(uint32_t R, uint32_t K) {
    static const uint32_t shapes[][2] = {{16, 16},  {32, 16},  {64, 16}, {128, 16}, {256, 16}, {512, 16},
                                         {1024, 16}, {1024, 8}, {2048, 8}, {2048, 4}, {4096, 4}};
    for (auto &s : shapes)
        if (s[0] == R && s[1] == K) return true;
    return false;
}

// Every size a kernel and its grid assume agrees with the plan's lengths: the launchers
// refuse a plan that fails this (a kernel of another shape would index out of bounds).
inline bool plan_consistent(const Plan &pl) {
    if (pl.n < 8 || (pl.n & (pl.n - 1)) || (1ull << pl.log_n) != pl.n) return false;
    if (pl.levels < 1 || pl.levels > pl.M || (pl.M != 4 && pl.M != 8)) return false;
    if (pl.small) {
        const uint64_t line = pl.row3 ? (uint64_t)pl.row3 * pl.row3 * pl.row3 : (uint64_t)pl.row.B1 * pl.row.B2;
        return line == pl.n && pl.C == pl.n && pl.R == 1 && !pl.col3 && !pl.row3m;
    }
    if ((uint64_t)pl.R * pl.C != pl.n) return false;
    const uint64_t col = pl.col3 ? (uint64_t)pl.col3 * pl.cBB * pl.cBC : (uint64_t)pl.col.B1 * pl.col.B2;
    if (col != pl.R) return false;
    if (!pl.K || pl.C % pl.K) return false;
    uint64_t row;
    if (pl.row3m) row = (uint64_t)pl.row3m * pl.rBB * pl.rBC;
    else if (pl.row3) row = (uint64_t)pl.row3 * pl.row3 * pl.row3;
    else row = (uint64_t)pl.row.B1 * pl.row.B2;
    if (row != pl.C) return false;
    return pl.G && pl.R % pl.G == 0 && (!pl.row3m || pl.G == 1);
}

inline std::string make_plan_fields(uint32_t p, uint64_t n, Plan &pl, PlanOpts opts);

// Build the plan; returns an empty string on success, else the reason.  pl is rebuilt from
// scratch on every call (a caller may retry with other options on the same Plan).
inline std::string make_plan(uint32_t p, uint64_t n, Plan &pl, PlanOpts opts = PlanOpts()) {
    pl = Plan();
    std::string why = make_plan_fields(p, n, pl, opts);
    if (why.empty() && !plan_consistent(pl)) why = "internal error: inconsistent plan";
    if (!why.empty()) pl = Plan();
    return why;
}

inline std::string make_plan_fields(uint32_t p, uint64_t n, Plan &pl, PlanOpts opts) {
    const uint32_t force_lc = opts.lc;
    if (n < 8 || (n & (n - 1))) return "NTT length must be a power of two >= 8";
    pl.n = (uint32_t)n;
    pl.log_n = ilog2(n);
    if (pl.log_n > 26) return "NTT lengths above 2^26 are not supported by this engine";
    if ((uint64_t)p < 2 * n) return "exponent too small for this NTT length (needs p >= 2n)";
    if ((uint64_t)p / n > 31) return "NTT length too small for this exponent (digits wider than 31 bits)";
    pl.levels = carry_levels(p, (uint32_t)n);
    if (pl.levels == 0) return "NTT length too small for this exponent (convolution would overflow)";
    pl.bias = digits_fit_31(p, (uint32_t)n, pl.levels);
    pl.M = pl.levels <= 4 ? 4 : 8;
    // n <= 4096 runs in one workgroup except at n = 2048, where the one-workgroup kernel
    // holds 64 words in each of 32 threads (121 us per iteration) and 16 rows of 128 words
    // over three launches take 21.5 us; a row length may also be asked for at n = 4096
    if (pl.log_n <= 10 || (pl.log_n == 12 && !force_lc)) {
        pl.small = true;
        pl.row = line_for(pl.log_n);
        pl.C = pl.n;
        pl.R = 1;
        // 3-phase single-workgroup kernel at n = 4096 = 16^3 and 512 = 8^3 unless asked not to
        if (opts.row3 != 0 && (pl.n == 4096 || pl.n == 512)) pl.row3 = pl.n == 4096 ? 16 : 8;
        return "";
    }
    uint32_t lc, lr;
    if (force_lc) {
        lc = force_lc;
        lr = pl.log_n - lc;
        if (lc < 7 || lc > 14 || lr < 4 || lr > 12) return "unsupported row/column split";
    } else if (pl.log_n == 11) {
        lc = 7;
        lr = 4;
    } else if (!split_for(pl.log_n, lc, lr)) {
        return "unsupported NTT length";
    }
    pl.C = 1u << lc;
    pl.R = 1u << lr;
    pl.row = line_for(lc);
    pl.col = line_for(lr);
    pl.K = pl.R >= 4096 ? 4 : pl.R >= 2048 ? 8 : 16;
    // 3-phase columns by default for R >= 512 (measured faster there)
    const int col3 = opts.col3 >= 0 ? opts.col3 : (pl.R >= 512 ? 1 : 0);
    if (col3 == 1) {
        if (pl.R < 512 || pl.R > 4096) return "3-phase columns need R = 512 .. 4096";
        pl.col3 = pl.R >= 1024 ? 16 : 8;
        const uint32_t A = pl.R / pl.col3;
        pl.cBB = 1u << ((ilog2(A) + 1) / 2);
        pl.cBC = A / pl.cBB;
        // 1024-thread blocks: K = 16 at R = 1024, 8 at R = 2048, 4 at R = 4096 (measured
        // faster than the 512-thread tiles)
        pl.K = pl.R >= 4096 ? 4 : pl.R >= 2048 ? 8 : pl.R >= 1024 ? 16 : 8;
    }
    if (opts.K) {
        if (!(pl.col3 ? col3_tile_ok(pl.R, opts.K) : col_tile_ok(pl.R, opts.K))) return "unsupported column tile";
        pl.K = opts.K;
    }
    pl.G = 64 / pl.row.B2;
    // 3-phase rows by default at C = 4096 (measured faster than the 2-phase rows there)
    if (pl.C > 4096) {
        // mixed-radix 3-phase rows, one row per workgroup of rBB * rBC threads:
        // row3 = 1: 32 words per thread (8192 = 32*16*16, 16384 = 32*32*16);
        // row3 = 2: 16 words per thread (8192 = 16*32*16, 16384 = 16*32*32).
        // Default: 16 words (row kernel at n = 2^25: 594 -> 491 us; at 2^26: 1304 -> 1120 us)
        if (opts.row3 == 0) return "rows above 4096 words need the 3-phase row kernel";
        const bool w16 = opts.row3 != 1;
        pl.row3m = w16 ? 16 : 32;
        pl.rBB = (pl.C == 8192 && !w16) ? 16 : 32;
        pl.rBC = (pl.C == 16384 && w16) ? 32 : 16;
        pl.G = 1;
        return "";
    }
    const int row3 = opts.row3 >= 0 ? opts.row3 : (pl.C == 4096 ? 1 : 0);
    if (row3 == 2) return "16-word mixed-radix rows are for rows above 4096 words";
    if (row3 == 1) {
        if (pl.C != 4096 && pl.C != 512) return "3-phase rows need C = 512 or 4096";
        pl.row3 = pl.C == 4096 ? 16 : 8;
        pl.G = pl.C == 4096 ? 1 : 4;
    }
    return "";
}

// Digit widths w_i = ceil(p(i+1)/n) - ceil(p i/n).
inline void widths(uint32_t p, uint32_t n, std::vector<uint8_t> &w) {
    w.resize(n);
    uint32_t L = ilog2(n);
    for (uint32_t i = 0; i < n; i++) {
        uint64_t e0 = ((uint64_t)p * i + n - 1) >> L, e1 = ((uint64_t)p * (i + 1) + n - 1) >> L;
        w[i] = (uint8_t)(e1 - e0);
    }
}

// Canonical digits -> balanced signed digits (field-encoded), exact value preserved
// mod 2^p - 1.  Any non-negative input digits are accepted.
inline void to_device_digits(const std::vector<uint8_t> &w, const uint64_t *d, uint64_t *x) {
    const size_t n = w.size();
    std::vector<int64_t> b(n);
    i128 carry = 0;
    for (size_t i = 0; i < n; i++) b[i] = 0;
    // first pass with the raw digits, later passes only push the wrapped carry
    for (int pass = 0; pass < 64; pass++) {
        for (size_t i = 0; i < n; i++) {
            i128 t = (pass == 0 ? (i128)d[i] : (i128)b[i]) + carry;
            int wi = w[i];
            i128 hi = (t + ((i128)1 << (wi - 1))) >> wi;
            b[i] = (int64_t)(t - (hi << wi));
            carry = hi;
            if (pass > 0 && carry == 0) break;
        }
        if (carry == 0) break;
        if (carry >= -1 && carry <= 1) {  // fold a unit wrap into digit 0 (slack of one)
            b[0] += (int64_t)carry;
            carry = 0;
            break;
        }
    }
    for (size_t i = 0; i < n; i++) x[i] = gl_from_i64(b[i]);
}

// Device words (pre-carry signed values) -> canonical digits in [0, 2^w).  Returns true
// if the value is 0 mod 2^p - 1.
inline bool to_canonical_digits(const std::vector<uint8_t> &w, const uint64_t *x, bool sub2,
                                uint64_t *d) {
    const size_t n = w.size();
    std::vector<int64_t> v(n);
    for (size_t i = 0; i < n; i++) v[i] = gl_to_i64(x[i]);
    if (sub2) v[0] -= 2;
    i128 carry = 0;
    bool first = true;
    for (int pass = 0; pass < 1000; pass++) {
        size_t i = 0;
        for (; i < n; i++) {
            i128 t = (first ? (i128)v[i] : (i128)d[i]) + carry;
            int wi = w[i];
            i128 q = t >> wi;  // floor
            d[i] = (uint64_t)(t - (q << wi));
            carry = q;
            if (!first && carry == 0) break;
        }
        first = false;
        if (carry == 0) break;
    }
    // all-ones is 2^p - 1 = 0 (mod 2^p - 1): canonical 0
    bool all_ones = true, all_zero = true;
    for (size_t i = 0; i < n; i++) {
        if (d[i] != (1ull << w[i]) - 1) all_ones = false;
        if (d[i] != 0) all_zero = false;
    }
    if (all_ones) std::memset(d, 0, n * sizeof(uint64_t));
    return all_ones || all_zero;
}

inline uint64_t res64(const std::vector<uint8_t> &w, const uint64_t *d) {
    uint64_t r = 0;
    uint32_t e = 0;
    for (size_t i = 0; i < w.size() && e < 64; i++) {
        r += d[i] << e;
        e += w[i];
    }
    return r;
}

}  // namespace lnt

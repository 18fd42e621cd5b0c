// Kernel bodies for one LL/PRP iteration, x <- x^2 (- 2) mod 2^p - 1, as an
// irrational-base weighted NTT over GF(2^64 - 2^32 + 1).
//
// Residue layout (n = R * C words, digit i = c + C * r): natural digit order, each word a
// field element that encodes a signed integer |v| < P/2.  Between iterations the words
// hold the *pre-carry* convolution outputs of the last squaring; the carry (and the LL
// "-2") of iteration k is applied at the start of iteration k+1, fused into the first
// kernel.  Readout normalises on the host.
//
// Transform decomposition (a "line" of L = B1 * B2 words, B1 >= B2, both <= 64):
//   phase 1: thread t < B2 holds words t + B2*j (j < B1), a B1-point register DFT,
//            then the inter-block twiddle omega_L^(t * k2);
//   phase 2: after an LDS transpose, B2-point register DFTs over t.
//   Output X[B1*k1 + k2].  All roots of order <= 64 are powers of two, so the only
//   general twiddle inside a line is the one between the phases.  3-phase lines
//   (rows of 16^3 words, columns of BA * BB * BC words) add one more phase and twiddle.
// Large residues (n >= 2^13) are a 2-D split: C-word rows (contiguous) and R-word
// columns (stride C), with the 4-step twiddle omega_n^(c * kcol) between the column and
// the row transforms.
//
// Per iteration (large path), state x and work buffer y (the carry of one column tile
// reads halo digits of the neighbouring tile, so col_fwd must not write x):
//   col_fwd : x -> y  carry(+ -2) of the previous pre-carry values, weight, column NTT
//   row_sq  : y -> y  4-step twiddle, row NTT, pointwise square, inverse row NTT,
//                     inverse 4-step twiddle
//   col_inv : y -> x  inverse column NTT, unweight (includes 1/n)
// These play the roles of the reference's r2b + m2 + w_fftc_tw (kernels/r2b.cu:25-50,
// kernels/m2.cu:23-41, kernels/w_fftc_tw__r2_8to4k.cu:20-46), fftr_sqr_ifftr
// (kernels/fftr_sqr_ifftr__r2_8to4k.cu:20-56) and tw_ifftc_uw
// (kernels/tw_ifftc_uw__r2_8to4k.cu:20-46).  Small residues (n <= 4096) run the whole
// iteration loop inside one workgroup (the role of w_fft_sqr_ifft_uw__*,
// kernels/w_fft_sqr_ifft_uw__r2_8to4k.cu:20-55, plus r2b and m2).
//
// Carry: "m-level" carry-save normalisation.  Each level splits every value into a
// balanced remainder |lo| <= 2^(w-1) and a carry hi, and adds the carry of the previous
// digit.  After m levels digit j depends only on the pre-carry values j-m..j, so a
// workgroup that owns a run of digits needs just m "halo" values before the run, read
// from memory - no inter-workgroup ordering.  The value sum v_j 2^e_j mod 2^p - 1 is
// preserved exactly (the carry out of digit n-1 enters digit 0 with weight 2^p = 1).
#pragma once
#include "dft.hpp"
#include "simt.hpp"

struct LntParams {
    GView<u64> x;            // n words: residue state (pre-carry values between iterations)
    GView<u64> y;            // n words: transform-domain work buffer
    GView<const u64> w, uw;  // IBDWT weights, unweights * n^-1: n words in digit order (one-workgroup
                             // path: the weights as pairs (w, -2^31 w) at 2i, the unweights at i);
                             // large path: row factors indexed by lnt_wslot, the weights as
                             // pairs at 2 * slot, the unweights at slot
    GView<const u64> tw4, itw4;  // 4-step twiddles (large path), natural slots, times the column
                                 // weight factor (unweight factor * n^-1)
    GView<const u64> twr, itwr;  // row-line twiddles (C words)
    GView<const u64> twr2, itwr2;  // second-level row twiddles (3-phase rows: B^2 words)
    GView<const u64> twc, itwc;  // column-line twiddles (R words)
    GView<const u64> twc2, itwc2;  // second-level column twiddles (3-phase columns: R / BA words)
    GView<const u64> f;      // n words: the other operand of a product, transform domain
    u32 p, n, log_n, C, R;
    u32 levels;              // carry levels m
    u32 sub2;                // batched launches: this test's LL "-2" is due at the next carry
    u32 bias;                // large path: carried digits go to the weight multiply as d + 2^31
};

// Weight exponents on the large path.  Digit i = c + C r has weight B^e_i, e_i = (-p i) mod n
// (B^n = 2), and e_i = e_c + f_r - n [e_c + f_r >= n] with e_c = (-p c) mod n and
// f_r = (-p C r) mod n.  So w_i = B^e_c * B^f_r * 2^-[wrap]: the column factor B^e_c is
// constant down a column and commutes with the column transform, so the 4-step twiddles
// carry it; the row factor B^f_r, or B^f_r / 2 on a wrap, sits in a table of 2R words at
// slot 2r + wrap (unweights likewise: B^-f_r and 2 B^-f_r, with B^-e_c n^-1 in itw4).
GL_HD u32 lnt_wexp(u32 p, u32 n, u64 i) { return (0u - (u32)((u64)p * i)) & (n - 1); }
// mC = -p C mod 2^32 (so f_r = (mC r) mod n, n being a power of two), ec = e_c.
GL_HD u32 lnt_wslot(u32 ec, u32 mC, u32 n, u32 r) {
    const u32 fr = (mC * r) & (n - 1);
    return 2 * r + (ec + fr >= n ? 1u : 0u);
}

// Balanced split of v by width w: v = lo + hi * 2^w, -2^(w-1) <= lo < 2^(w-1).
GL_HD int64_t lnt_split(int64_t v, int w, int64_t &lo) {
    int64_t hi = (v + (int64_t(1) << (w - 1))) >> w;
    lo = (int64_t)((u64)v - ((u64)hi << w));
    return hi;
}

// m-level carry over a run: v[0..M) are the M digits before the run (halo), v[M..M+K)
// the run itself, starting at digit g0.  On return v[M..M+K) hold final digits.
template <int M, int K>
GL_HD void lnt_carry_run(int64_t (&v)[M + K], u32 p, u32 log_n, u64 g0, u32 levels,
                         bool sub2) {
    const u64 n = 1ull << log_n;
    // widths: with R_g = (p g + n - 1) mod n, the width of digit g is
    // floor((R_g + p) / n) = p / n + [R_g + (p mod n) >= n] and R_(g+1) = (R_g + p) mod n;
    // R_n = R_0, so the recurrence runs across the wrap from digit n-1 to digit 0
    const u32 nm = (u32)n - 1, pm = p & nm, wmin = p >> log_n;
    u32 R = ((u32)p * (u32)((g0 + n - M) & (n - 1)) + nm) & nm;
    int w[M + K];
#pragma unroll
    for (int j = 0; j < M + K; j++) {
        const u32 s = R + pm;
        w[j] = (int)(wmin + (s >> log_n));
        R = s & nm;
    }
    // LL "-2" on digit 0, at window position M - g0 when g0 <= M: in the run itself
    // (g0 = 0) or in the halo of a run just after it (every window that holds digit 0 must
    // see the same value); a select per position, no indexing by a run-time value
    if (sub2 && g0 <= (u64)M) {
#pragma unroll
        for (int j = 0; j <= M; j++) v[j] -= (g0 + (u64)j == (u64)M) ? 2 : 0;
    }
    for (u32 l = 0; l < levels; l++) {
        int64_t hprev = 0;
#pragma unroll
        for (int j = 0; j < M + K; j++) {
            int64_t lo;
            int64_t h = lnt_split(v[j], w[j], lo);
            v[j] = lo + hprev;
            hprev = h;
        }
    }
}

constexpr int LNT_PAD = 1;

// ------------------------------------------------------------------ column forward
// grid: C / K workgroups (column tile c0 = bid * K), block: K * B2 threads.
// LDS words: lnt_col_fwd_lds<B1, B2, K>().
// Row stride of the carry staging layout: one word of padding keeps the per-row carry
// accesses (one thread per row) on distinct LDS banks while the stage fits in 144 KiB.
template <int B1, int B2, int K>
constexpr int lnt_col_ks() { return B1 * B2 * (K + 1) <= 18432 ? K + 1 : K; }
template <int B1, int B2, int K>
constexpr u32 lnt_col_fwd_lds() {
    return (u32)(B1 * B2 * lnt_col_ks<B1, B2, K>() > B1 * (B2 + LNT_PAD) * K ? B1 * B2 * lnt_col_ks<B1, B2, K>()
                                                                           : B1 * (B2 + LNT_PAD) * K);
}

template <int B1, int B2, int K, int M>
LNT_KFN void lnt_col_fwd_body(int tid, int bid, LView lds, const LntParams &P, bool sub2) {
    constexpr int T = K * B2;
    constexpr int RR = B1 * B2;
    constexpr int RP = B1 / B2;
    constexpr int S = B2 + LNT_PAD;
    constexpr int KS = lnt_col_ks<B1, B2, K>();
    constexpr int RT = (RR + T - 1) / T;  // carry rows per thread (the last may be idle)
    const u64 C = P.C;
    const u64 n = P.n;
    const u64 c0 = (u64)bid * K;
    const int cl = tid % K, t = tid / K;
    const u64 c = c0 + cl;
    const u32 ec = lnt_wexp(P.p, P.n, c), mC = 0u - P.p * P.C;  // weight slot parameters

    // 0. loads: the tile (K adjacent lanes read one row run: 8K contiguous bytes), the M
    //    digits before each of this thread's carry rows, and the weights of phase 1.
    // Tables are prefetched into registers when a thread's line fits (B1 <= 32).
    // The halo digits are prefetched too when there are few of them per thread.
    constexpr bool PF = B1 <= 32, HP = RT * M <= 16;
    u64 a[B1], wv[PF ? B1 : 1], h[HP ? RT : 1][M];
#pragma unroll
    for (int j = 0; j < B1; j++) a[j] = P.x[c + C * (u64)(t + B2 * j)];
    if (HP) {
#pragma unroll
        for (int q = 0; q < RT; q++) {
            const u64 g0 = C * (u64)(tid + T * q) + c0;
#pragma unroll
            for (int j = 0; j < M; j++)
                h[HP ? q : 0][j] = (tid + T * q < RR) ? P.x[(g0 + n - M + j) & (n - 1)] : 0;
        }
    }
    if (PF) {
#pragma unroll
        for (int j = 0; j < B1; j++) wv[PF ? j : 0] = P.w[2 * lnt_wslot(ec, mC, P.n, (u32)(t + B2 * j))];
    }
#pragma unroll
    for (int j = 0; j < B1; j++) lds[(u32)((t + B2 * j) * KS + cl)] = a[j];
    LNT_SYNC();

    // 1. carry + (-2): one thread per row run of K digits (plus the M halo digits)
#pragma unroll
    for (int q = 0; q < RT; q++) {
        const int r = tid + T * q;
        if (r >= RR) break;
        int64_t v[M + K];
        const u64 g0 = C * (u64)r + c0;
#pragma unroll
        for (int j = 0; j < M; j++) v[j] = gl_to_i64(HP ? h[HP ? q : 0][j] : P.x[(g0 + n - M + j) & (n - 1)]);
#pragma unroll
        for (int j = 0; j < K; j++) v[M + j] = gl_to_i64(lds[(u32)(r * KS + j)]);
        lnt_carry_run<M, K>(v, P.p, P.log_n, g0, P.levels, sub2);
        // the digit as d + 2^31 for the biased weight multiply, else field-encoded
#pragma unroll
        for (int j = 0; j < K; j++)
            lds[(u32)(r * KS + j)] = P.bias ? (u64)(u32)(v[M + j] + (int64_t(1) << 31)) : gl_from_i64(v[M + j]);
    }
    LNT_SYNC();

    // 2. phase 1: weight, B1-point DFT down the column, inter-block twiddle
    if (P.bias) {
#pragma unroll
        for (int j = 0; j < B1; j++) {
            const u32 sl = lnt_wslot(ec, mC, P.n, (u32)(t + B2 * j));
            a[j] = gl_mul_biased((u32)lds[(u32)((t + B2 * j) * KS + cl)], PF ? wv[PF ? j : 0] : P.w[2 * sl],
                                 P.w[2 * sl + 1]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < B1; j++)
            a[j] = gl_mul(lds[(u32)((t + B2 * j) * KS + cl)],
                          PF ? wv[PF ? j : 0] : P.w[2 * lnt_wslot(ec, mC, P.n, (u32)(t + B2 * j))]);
    }
    dft_fwd<B1>(a);
    if (B2 > 1) {
#pragma unroll
        for (int rho = 1; rho < B1; rho++) a[rho] = gl_mul(a[rho], P.twc[(u64)rho * B2 + t]);
    }
    LNT_SYNC();
#pragma unroll
    for (int rho = 0; rho < B1; rho++)
        lds[(u32)((brev_c<B1>(rho) * S + t) * K + cl)] = a[rho];
    LNT_SYNC();

    // 3. phase 2: B2-point DFTs, store (row r' = t*B1 + q*B2 + rho'); the 4-step twiddle
    //    is applied by the row pass as it loads the row
#pragma unroll
    for (int q = 0; q < RP; q++) {
        const int k2 = t * RP + q;
#pragma unroll
        for (int i1 = 0; i1 < B2; i1++) a[q * B2 + i1] = lds[(u32)((k2 * S + i1) * K + cl)];
        dft_fwd<B2>(a + q * B2);
    }
#pragma unroll
    for (int i = 0; i < B1; i++) P.y[c + C * ((u64)t * B1 + i)] = a[i];
}

// ------------------------------------------------------------------ column inverse
template <int B1, int B2, int K>
LNT_KFN void lnt_col_inv_body(int tid, int bid, LView lds, const LntParams &P) {
    constexpr int RP = B1 / B2;
    constexpr int S = B2 + LNT_PAD;
    const u64 C = P.C;
    const int cl = tid % K, t = tid / K;
    const u64 c = (u64)bid * K + cl;
    const u32 ec = lnt_wexp(P.p, P.n, c), mC = 0u - P.p * P.C;  // weight slot parameters
    constexpr bool PF = B1 <= 32;  // prefetch the unweighting factors (fits when B1 <= 32)
    u64 a[B1], uv[PF ? B1 : 1];
    // (the inverse 4-step twiddle was applied by the row pass as it stored the row)
#pragma unroll
    for (int i = 0; i < B1; i++) a[i] = P.y[c + C * ((u64)t * B1 + i)];
    if (PF) {
#pragma unroll
        for (int j = 0; j < B1; j++) uv[PF ? j : 0] = P.uw[lnt_wslot(ec, mC, P.n, (u32)(t + B2 * j))];
    }
#pragma unroll
    for (int q = 0; q < RP; q++) {
        dft_inv<B2>(a + q * B2);
        const int k2 = t * RP + q;
#pragma unroll
        for (int i1 = 0; i1 < B2; i1++) lds[(u32)((k2 * S + i1) * K + cl)] = a[q * B2 + i1];
    }
    LNT_SYNC();
#pragma unroll
    for (int rho = 0; rho < B1; rho++) {
        u64 v = lds[(u32)((brev_c<B1>(rho) * S + t) * K + cl)];
        a[rho] = (B2 > 1 && rho > 0) ? gl_mul(v, P.itwc[(u64)rho * B2 + t]) : v;
    }
    dft_inv<B1>(a);
#pragma unroll
    for (int j = 0; j < B1; j++) {
        const u64 idx = c + C * ((u64)t + (u64)B2 * j);
        P.x[idx] = gl_mul(a[j], PF ? uv[PF ? j : 0] : P.uw[lnt_wslot(ec, mC, P.n, (u32)(t + B2 * j))]);
    }
}

// ------------------------------------------------------------------ 3-phase columns
// Column line R = BA * BB * BC handled by A = BB * BC threads per column (K columns per
// workgroup, block K * A threads, BA words per thread).  Row r = a + A j (a < A, j < BA):
//  phase 1: thread a, BA-point DFT over j -> k1, twiddle w_R^(a k1)
//  phase 2: items (k1, u), u < BC (BA / BB per thread): BB-point DFT over v of the words
//           a = u + BC v -> k2, twiddle w_A^(u k2)
//  phase 3: items (k1, k2) (BA / BC per thread): BC-point DFT over u -> k3
// Frequency k1 + BA (k2 + BB k3) ends in slot t*BA + q*BC + rho (k3 = brev(rho)).
// LDS word (row, column) at row * KP + column, rows k1 * (A + 1) + a.
template <int BA, int BB, int BC, int K>
struct LntCol3 {
    static constexpr int A = BB * BC, R = BA * A;
    static constexpr int QB = BA / BB, QC = BA / BC;
    static constexpr int RS = A + 1;
    static constexpr int KP = BA * RS * (K + 1) <= 18432 ? K + 1 : K;  // padded when it fits 144 KiB
    static constexpr int KS = R * (K + 1) <= 18432 ? K + 1 : K;        // carry staging row width
    static constexpr u32 LDS = (u32)(BA * RS * KP > R * KS ? BA * RS * KP : R * KS);
    static_assert(BA % BB == 0 && BA % BC == 0 && BB >= BC, "3-phase column radices");
};

template <int BA, int BB, int BC, int K, int M>
LNT_KFN void lnt_col3_fwd_body(int tid, int bid, LView lds, const LntParams &P, bool sub2) {
    using S = LntCol3<BA, BB, BC, K>;
    constexpr int A = S::A, RR = S::R, RS = S::RS, KP = S::KP, KS = S::KS;
    constexpr int NT = K * A;
    constexpr int RT = (RR + NT - 1) / NT;  // carry rows per thread
    constexpr bool HP = RT * M <= 16;
    const u64 C = P.C;
    const u64 n = P.n;
    const u64 c0 = (u64)bid * K;
    const int cl = tid % K, t = tid / K;
    const u64 c = c0 + cl;
    const u32 ec = lnt_wexp(P.p, P.n, c), mC = 0u - P.p * P.C;  // weight slot parameters

    // 0. loads: rows t + A j of column c (K adjacent lanes read one row run), the halo
    //    digits of this thread's carry rows, the weights (prefetched into registers only
    //    in blocks of <= 512 threads, whose register budget allows it)
    constexpr bool PW = NT <= 512;
    u64 a[BA], wv[PW ? BA : 1], wb[PW ? BA : 1], h[HP ? RT : 1][M];
#pragma unroll
    for (int j = 0; j < BA; j++) a[j] = P.x[c + C * (u64)(t + A * j)];
    if (HP) {
#pragma unroll
        for (int q = 0; q < RT; q++) {
            const u64 g0 = C * (u64)(tid + NT * q) + c0;
#pragma unroll
            for (int j = 0; j < M; j++)
                h[HP ? q : 0][j] = (tid + NT * q < RR) ? P.x[(g0 + n - M + j) & (n - 1)] : 0;
        }
    }
    if (PW) {
#pragma unroll
        for (int j = 0; j < BA; j++) {
            const u32 sl = lnt_wslot(ec, mC, P.n, (u32)(t + A * j));
            wv[PW ? j : 0] = P.w[2 * sl];
            wb[PW ? j : 0] = P.w[2 * sl + 1];
        }
    }
#pragma unroll
    for (int j = 0; j < BA; j++) lds[(u32)((t + A * j) * KS + cl)] = a[j];
    LNT_SYNC();

    // 1. carry + (-2): one thread per row run of K digits (plus the M halo digits)
#pragma unroll
    for (int q = 0; q < RT; q++) {
        const int r = tid + NT * q;
        if (r >= RR) break;
        // streamed in pieces of PK digits, each piece's halo being the M pre-carry values
        // before it: the whole 16-digit run in one piece where it fits the 128 registers of
        // a 1024-thread block (BA = 16, M = 4), pieces of 8 where one piece would spill
        constexpr int PK = (NT >= 1024 && K > 8 && (M > 4 || BA < 16)) ? 8 : K;
        const u64 g0 = C * (u64)r + c0;
        int64_t halo[M];
#pragma unroll
        for (int j = 0; j < M; j++) halo[j] = gl_to_i64(HP ? h[HP ? q : 0][j] : P.x[(g0 + n - M + j) & (n - 1)]);
#pragma unroll
        for (int s = 0; s < K / PK; s++) {
            int64_t v[M + PK];
#pragma unroll
            for (int j = 0; j < M; j++) v[j] = halo[j];
#pragma unroll
            for (int j = 0; j < PK; j++) v[M + j] = gl_to_i64(lds[(u32)(r * KS + s * PK + j)]);
#pragma unroll
            for (int j = 0; j < M; j++) halo[j] = v[PK + j];
            lnt_carry_run<M, PK>(v, P.p, P.log_n, g0 + s * PK, P.levels, sub2);
            // the digit as d + 2^31 for the biased weight multiply, else field-encoded
#pragma unroll
            for (int j = 0; j < PK; j++)
                lds[(u32)(r * KS + s * PK + j)] =
                    P.bias ? (u64)(u32)(v[M + j] + (int64_t(1) << 31)) : gl_from_i64(v[M + j]);
        }
    }
    LNT_SYNC();

    // 2. phase 1: weight, BA-point DFT, twiddle
    if (P.bias) {
#pragma unroll
        for (int j = 0; j < BA; j++) {
            const u32 sl = PW ? 0 : lnt_wslot(ec, mC, P.n, (u32)(t + A * j));
            a[j] = gl_mul_biased((u32)lds[(u32)((t + A * j) * KS + cl)], PW ? wv[PW ? j : 0] : P.w[2 * sl],
                                 PW ? wb[PW ? j : 0] : P.w[2 * sl + 1]);
        }
    } else {
#pragma unroll
        for (int j = 0; j < BA; j++)
            a[j] = gl_mul(lds[(u32)((t + A * j) * KS + cl)],
                          PW ? wv[PW ? j : 0] : P.w[2 * lnt_wslot(ec, mC, P.n, (u32)(t + A * j))]);
    }
    dft_fwd<BA>(a);
#pragma unroll
    for (int rho = 1; rho < BA; rho++) a[rho] = gl_mul(a[rho], P.twc[(u64)rho * A + t]);
    LNT_SYNC();
#pragma unroll
    for (int rho = 0; rho < BA; rho++) lds[(u32)((brev_c<BA>(rho) * RS + t) * KP + cl)] = a[rho];
    LNT_SYNC();

    // 3. phase 2 (each item rewrites exactly the LDS rows it read: no barrier inside)
#pragma unroll
    for (int q = 0; q < S::QB; q++) {
        const int I = q * A + t, k1 = I / BC, u = I % BC;
        u64 b[BB];
#pragma unroll
        for (int v = 0; v < BB; v++) b[v] = lds[(u32)((k1 * RS + u + BC * v) * KP + cl)];
        dft_fwd<BB>(b);
#pragma unroll
        for (int rho = 0; rho < BB; rho++) lds[(u32)((k1 * RS + BC * brev_c<BB>(rho) + u) * KP + cl)] = b[rho];
    }
    LNT_SYNC();

    // 4. phase 3 with the phase-2 twiddle w_A^(u k2) at the read, store slot
    //    t*BA + q*BC + rho (the row pass applies the 4-step twiddle)
#pragma unroll
    for (int q = 0; q < S::QC; q++) {
        const int I = q * A + t, k1 = I / BB, k2 = I % BB;
        const u64 tb = (u64)brev_c<BB>(k2) * BC;
#pragma unroll
        for (int u = 0; u < BC; u++)
            a[q * BC + u] = gl_mul(lds[(u32)((k1 * RS + BC * k2 + u) * KP + cl)], P.twc2[tb + u]);
        dft_fwd<BC>(a + q * BC);
    }
#pragma unroll
    for (int i = 0; i < BA; i++) P.y[c + C * ((u64)t * BA + i)] = a[i];
}

template <int BA, int BB, int BC, int K>
LNT_KFN void lnt_col3_inv_body(int tid, int bid, LView lds, const LntParams &P) {
    using S = LntCol3<BA, BB, BC, K>;
    constexpr int A = S::A, RS = S::RS, KP = S::KP;
    const u64 C = P.C;
    const int cl = tid % K, t = tid / K;
    const u64 c = (u64)bid * K + cl;
    const u32 ec = lnt_wexp(P.p, P.n, c), mC = 0u - P.p * P.C;  // weight slot parameters
    u64 a[BA], uv[BA];
#pragma unroll
    for (int i = 0; i < BA; i++) a[i] = P.y[c + C * ((u64)t * BA + i)];
#pragma unroll
    for (int j = 0; j < BA; j++) uv[j] = P.uw[lnt_wslot(ec, mC, P.n, (u32)(t + A * j))];
    // phase 3 inverse, then the inverse phase-2 twiddle
#pragma unroll
    for (int q = 0; q < S::QC; q++) {
        const int I = q * A + t, k1 = I / BB, k2 = I % BB;
        const u64 tb = (u64)brev_c<BB>(k2) * BC;
        dft_inv<BC>(a + q * BC);
#pragma unroll
        for (int u = 0; u < BC; u++)
            lds[(u32)((k1 * RS + BC * k2 + u) * KP + cl)] = gl_mul(a[q * BC + u], P.itwc2[tb + u]);
    }
    LNT_SYNC();
    // phase 2 inverse
#pragma unroll
    for (int q = 0; q < S::QB; q++) {
        const int I = q * A + t, k1 = I / BC, u = I % BC;
        u64 b[BB];
#pragma unroll
        for (int rho = 0; rho < BB; rho++) b[rho] = lds[(u32)((k1 * RS + BC * brev_c<BB>(rho) + u) * KP + cl)];
        dft_inv<BB>(b);
#pragma unroll
        for (int v = 0; v < BB; v++) lds[(u32)((k1 * RS + u + BC * v) * KP + cl)] = b[v];
    }
    LNT_SYNC();
    // phase 1 inverse, unweight, store
#pragma unroll
    for (int rho = 0; rho < BA; rho++) {
        u64 v = lds[(u32)((brev_c<BA>(rho) * RS + t) * KP + cl)];
        a[rho] = rho ? gl_mul(v, P.itwc[(u64)rho * A + t]) : v;
    }
    dft_inv<BA>(a);
#pragma unroll
    for (int j = 0; j < BA; j++) {
        const u64 idx = c + C * ((u64)t + (u64)A * j);
        P.x[idx] = gl_mul(a[j], uv[j]);
    }
}

// ------------------------------------------------------------------ row pass
// grid: R / G workgroups, block: G * B2 threads (G rows per workgroup).
// LDS words: G * B1 * (B2 + PAD).
// MODE 0: square (one iteration).  MODE 1: forward only; the transform-domain words go
// to slot base + t*B1 + i of y.  MODE 2: forward, multiply by the other operand's words f
// (same slots, written by a MODE 1 pass), inverse.
constexpr int LNT_SQUARE = 0, LNT_FORWARD = 1, LNT_MULTIPLY = 2;
template <int B1, int B2, int G, int MODE = LNT_SQUARE>
LNT_KFN void lnt_row_body(int tid, int bid, LView lds, const LntParams &P) {
    constexpr int RP = B1 / B2;
    constexpr int S = B2 + LNT_PAD;
    constexpr int LINE = B1 * S;
    const int t = tid % B2, g = tid / B2;
    const u64 base = (u64)P.C * ((u64)bid * G + g);
    const u32 lb = (u32)(g * LINE);
    u64 a[B1];
#pragma unroll
    for (int j = 0; j < B1; j++) {
        const u64 idx = base + t + (u64)B2 * j;
        a[j] = gl_mul(P.y[idx], P.tw4[idx]);  // 4-step twiddle of the column pass
    }
    dft_fwd<B1>(a);
    if (B2 > 1) {
#pragma unroll
        for (int rho = 1; rho < B1; rho++) a[rho] = gl_mul(a[rho], P.twr[(u64)rho * B2 + t]);
    }
#pragma unroll
    for (int rho = 0; rho < B1; rho++) lds[lb + (u32)(brev_c<B1>(rho) * S + t)] = a[rho];
    LNT_SYNC();
#pragma unroll
    for (int q = 0; q < RP; q++) {
        const int k2 = t * RP + q;
#pragma unroll
        for (int i1 = 0; i1 < B2; i1++) a[q * B2 + i1] = lds[lb + (u32)(k2 * S + i1)];
        dft_fwd<B2>(a + q * B2);
        if (MODE == LNT_FORWARD) {
#pragma unroll
            for (int i1 = 0; i1 < B2; i1++) P.y[base + (u64)t * B1 + q * B2 + i1] = a[q * B2 + i1];
            continue;
        }
        if (MODE == LNT_MULTIPLY) {
#pragma unroll
            for (int i1 = 0; i1 < B2; i1++)
                a[q * B2 + i1] = gl_mul(a[q * B2 + i1], P.f[base + (u64)t * B1 + q * B2 + i1]);
        } else {
#pragma unroll
            for (int i1 = 0; i1 < B2; i1++) a[q * B2 + i1] = gl_sqr(a[q * B2 + i1]);
        }
        dft_inv<B2>(a + q * B2);
#pragma unroll
        for (int i1 = 0; i1 < B2; i1++) lds[lb + (u32)(k2 * S + i1)] = a[q * B2 + i1];
    }
    if (MODE == LNT_FORWARD) return;
    LNT_SYNC();
#pragma unroll
    for (int rho = 0; rho < B1; rho++) {
        u64 v = lds[lb + (u32)(brev_c<B1>(rho) * S + t)];
        a[rho] = (B2 > 1 && rho > 0) ? gl_mul(v, P.itwr[(u64)rho * B2 + t]) : v;
    }
    dft_inv<B1>(a);
#pragma unroll
    for (int j = 0; j < B1; j++) {
        const u64 idx = base + t + (u64)B2 * j;
        P.y[idx] = gl_mul(a[j], P.itw4[idx]);  // inverse 4-step twiddle
    }
}

// ------------------------------------------------------------------ 3-phase row pass
// Rows of C = B^3 words, B^2 threads per row each holding B words (more waves per SIMD
// than the 2-phase row: B registers instead of B1 = 64).  With i = i1 + B^2 i2
// (i1 < B^2) and i1 = a + B b:
//   phase 1: thread i1: B-point DFT over i2 -> k2, twiddle omega_C^(i1 k2)      (twr)
//   phase 2: thread (k2, a): B-point DFT over b -> kb, twiddle omega_B^2^(a kb) (twr2)
//   phase 3: thread (k2, kb): B-point DFT over a -> ka; X[B (B ka + kb) + k2]
// then the pointwise square (or the product modes of lnt_row_body) and the inverse in
// reverse.  Within the LDS row of k2 a word (r, a) sits at r * (B + 1) + a, so every
// phase reads and writes conflict-free strides.  Transform-domain slot for MODE 1/2:
// base + k2 * B^2 + kb * B + i (thread-contiguous).
// LDS layout: word (k2, a) of a row at k2 * RS + a with runs of B words, unpadded by default:
// the phase-3 runs are then 16-byte aligned and pair into 128-bit LDS accesses (at n = 2^22
// the unpadded 4096-word row ran 53 against 55 us with one pad word per run).
template <int B, int G, int MODE = LNT_SQUARE, int PAD = 0>
LNT_KFN void lnt_row3_body(int tid, int bid, LView lds, const LntParams &P) {
    constexpr int T = B * B;        // threads per row
    constexpr int RS = B * (B + PAD); // LDS words per k2 row
    constexpr int LINE = B * RS;    // LDS words per row
    const int t = tid % T, g = tid / T;
    const u64 base = (u64)P.C * ((u64)bid * G + g);
    const u32 lb = (u32)(g * LINE);
    const int hi = t / B, lo = t % B;  // phases 2/3: (k2, a) or (k2, kb)
    u64 a[B];
    // phase 1
#pragma unroll
    for (int j = 0; j < B; j++) {
        const u64 idx = base + t + (u64)T * j;
        a[j] = gl_mul(P.y[idx], P.tw4[idx]);  // 4-step twiddle of the column pass
    }
    // the inverse 4-step twiddles of the final store, fetched now so their latency hides
    // behind the transforms
    u64 iv[MODE == LNT_FORWARD ? 1 : B];
    if (MODE != LNT_FORWARD) {
#pragma unroll
        for (int j = 0; j < B; j++) iv[MODE == LNT_FORWARD ? 0 : j] = P.itw4[base + t + (u64)T * j];
    }
    dft_fwd<B>(a);
#pragma unroll
    for (int rho = 1; rho < B; rho++) a[rho] = gl_mul(a[rho], P.twr[(u64)rho * T + t]);
#pragma unroll
    for (int rho = 0; rho < B; rho++) lds[lb + (u32)(brev_c<B>(rho) * RS + (t / B) * (B + PAD) + t % B)] = a[rho];
    LNT_SYNC();
    // phase 2: thread (k2 = hi, a = lo) over b
#pragma unroll
    for (int b = 0; b < B; b++) a[b] = lds[lb + (u32)(hi * RS + b * (B + PAD) + lo)];
    // (this thread rewrites exactly the slots it just read: no barrier needed)
    dft_fwd<B>(a);
#pragma unroll
    for (int rho = 1; rho < B; rho++) a[rho] = gl_mul(a[rho], P.twr2[(u64)rho * B + lo]);
#pragma unroll
    for (int rho = 0; rho < B; rho++) lds[lb + (u32)(hi * RS + brev_c<B>(rho) * (B + PAD) + lo)] = a[rho];
    LNT_SYNC();
    // phase 3: thread (k2 = hi, kb = lo) over a
#pragma unroll
    for (int i = 0; i < B; i++) a[i] = lds[lb + (u32)(hi * RS + lo * (B + PAD) + i)];
    dft_fwd<B>(a);
    if (MODE == LNT_FORWARD) {
#pragma unroll
        for (int i = 0; i < B; i++) P.y[base + (u64)t * B + i] = a[i];
        return;
    }
    if (MODE == LNT_MULTIPLY) {
#pragma unroll
        for (int i = 0; i < B; i++) a[i] = gl_mul(a[i], P.f[base + (u64)t * B + i]);
    } else {
#pragma unroll
        for (int i = 0; i < B; i++) a[i] = gl_sqr(a[i]);
    }
    dft_inv<B>(a);
#pragma unroll
    for (int i = 0; i < B; i++) lds[lb + (u32)(hi * RS + lo * (B + PAD) + i)] = a[i];
    LNT_SYNC();
    // phase 2 inverse: thread (k2, a): registers rho hold kb = brev(rho)
#pragma unroll
    for (int rho = 0; rho < B; rho++) {
        u64 v = lds[lb + (u32)(hi * RS + brev_c<B>(rho) * (B + PAD) + lo)];
        a[rho] = rho ? gl_mul(v, P.itwr2[(u64)rho * B + lo]) : v;
    }
    dft_inv<B>(a);
#pragma unroll
    for (int b = 0; b < B; b++) lds[lb + (u32)(hi * RS + b * (B + PAD) + lo)] = a[b];
    LNT_SYNC();
    // phase 1 inverse: thread i1 = t, registers rho hold k2 = brev(rho)
#pragma unroll
    for (int rho = 0; rho < B; rho++) {
        u64 v = lds[lb + (u32)(brev_c<B>(rho) * RS + (t / B) * (B + PAD) + t % B)];
        a[rho] = rho ? gl_mul(v, P.itwr[(u64)rho * T + t]) : v;
    }
    dft_inv<B>(a);
#pragma unroll
    for (int j = 0; j < B; j++) P.y[base + t + (u64)T * j] = gl_mul(a[j], iv[MODE == LNT_FORWARD ? 0 : j]);
}

// ------------------------------------------------------------------ mixed-radix 3-phase rows
// Rows of C = BA * BB * BC words (8192 = 32*16*16 or 16*32*16, 16384 = 32*32*16 or
// 16*32*32), one row per workgroup of A = BB * BC threads holding BA words each.  The
// phases are those of the 3-phase columns (phase 1 over j of words t + A j; phase 2 items
// (k1, u), BA * BC of them; phase 3 items (k1, k2), BA * BB of them), with the 4-step
// twiddles at load and store and the square (or product) in the middle.  The phase-2
// twiddle is applied where phase 3 reads (and its inverse where inverse phase 3 writes).
// With BA = 16 a phase may have fewer items than threads (the rest wait at the barrier):
// 16 words per thread keep the register count low enough for twice the waves of BA = 32.
// LDS word of (k1, a = u + BC v): k1 * RS + v * (BC + 1) + u (padded runs).
template <int BA, int BB, int BC>
struct LntRow3m {
    static constexpr int A = BB * BC, C = BA * A;
    static constexpr int N2 = BA * BC, N3 = BA * BB;  // phase 2 / phase 3 items
    static constexpr int Q2 = (N2 + A - 1) / A, Q3 = (N3 + A - 1) / A;  // items per thread
    static constexpr int RS = BB * (BC + 1) + 1;
    static constexpr u32 LDS = (u32)(BA * RS);
    static_assert(N2 % A == 0 || A % N2 == 0, "phase 2 items");
    static_assert(N3 % A == 0 || A % N3 == 0, "phase 3 items");
};

template <int BA, int BB, int BC, int MODE = LNT_SQUARE>
LNT_KFN void lnt_row3m_body(int tid, int bid, LView lds, const LntParams &P) {
    using S = LntRow3m<BA, BB, BC>;
    constexpr int A = S::A, RS = S::RS, N2 = S::N2, N3 = S::N3, Q2 = S::Q2, Q3 = S::Q3;
    const int t = tid;
    const u64 base = (u64)P.C * (u64)bid;
    auto pos = [](int k1, int a) { return (u32)(k1 * RS + (a / BC) * (BC + 1) + a % BC); };
    u64 a[BA];
#pragma unroll
    for (int j = 0; j < BA; j++) {
        const u64 idx = base + t + (u64)A * j;
        a[j] = gl_mul(P.y[idx], P.tw4[idx]);  // 4-step twiddle of the column pass
    }
    // phase 1
    dft_fwd<BA>(a);
#pragma unroll
    for (int rho = 1; rho < BA; rho++) a[rho] = gl_mul(a[rho], P.twr[(u64)rho * A + t]);
#pragma unroll
    for (int rho = 0; rho < BA; rho++) lds[pos(brev_c<BA>(rho), t)] = a[rho];
    LNT_SYNC();
    // phase 2 (each item rewrites exactly the LDS words it read)
#pragma unroll
    for (int q = 0; q < Q2; q++) {
        const int I = q * A + t, k1 = I / BC, u = I % BC;
        if (N2 >= A || I < N2) {
            u64 b[BB];
#pragma unroll
            for (int v = 0; v < BB; v++) b[v] = lds[pos(k1, u + BC * v)];
            dft_fwd<BB>(b);
#pragma unroll
            for (int rho = 0; rho < BB; rho++) lds[pos(k1, u + BC * brev_c<BB>(rho))] = b[rho];
        }
    }
    LNT_SYNC();
    // phase 3 (with the phase-2 twiddle w_A^(u k2) at the read, where every thread has an
    // item: phase 2 keeps only its DFT registers live), the pointwise step, inverse phase 3
    // and the inverse phase-2 twiddle.  Item q of thread t: slots (q * A + t) * BC + rho.
#pragma unroll
    for (int q = 0; q < Q3; q++) {
        const int I = q * A + t, k1 = I / BB, k2 = I % BB;
        if (N3 >= A || I < N3) {
            const u64 tb = (u64)brev_c<BB>(k2) * BC;  // twr2 row of k2
            u64 e[BC];
#pragma unroll
            for (int u = 0; u < BC; u++) e[u] = gl_mul(lds[pos(k1, u + BC * k2)], P.twr2[tb + u]);
            dft_fwd<BC>(e);
            const u64 so = base + (u64)I * BC;
            if (MODE == LNT_FORWARD) {
#pragma unroll
                for (int i = 0; i < BC; i++) P.y[so + i] = e[i];
                continue;
            }
            if (MODE == LNT_MULTIPLY) {
#pragma unroll
                for (int i = 0; i < BC; i++) e[i] = gl_mul(e[i], P.f[so + i]);
            } else {
#pragma unroll
                for (int i = 0; i < BC; i++) e[i] = gl_sqr(e[i]);
            }
            dft_inv<BC>(e);
#pragma unroll
            for (int u = 0; u < BC; u++) lds[pos(k1, u + BC * k2)] = gl_mul(e[u], P.itwr2[tb + u]);
        }
    }
    if (MODE == LNT_FORWARD) return;
    LNT_SYNC();
    // inverse phase 2 (its twiddle was applied by inverse phase 3)
#pragma unroll
    for (int q = 0; q < Q2; q++) {
        const int I = q * A + t, k1 = I / BC, u = I % BC;
        if (N2 >= A || I < N2) {
            u64 b[BB];
#pragma unroll
            for (int rho = 0; rho < BB; rho++) b[rho] = lds[pos(k1, u + BC * brev_c<BB>(rho))];
            dft_inv<BB>(b);
#pragma unroll
            for (int v = 0; v < BB; v++) lds[pos(k1, u + BC * v)] = b[v];
        }
    }
    LNT_SYNC();
    // inverse phase 1, inverse 4-step twiddle, store
#pragma unroll
    for (int rho = 0; rho < BA; rho++) {
        u64 v = lds[pos(brev_c<BA>(rho), t)];
        a[rho] = rho ? gl_mul(v, P.itwr[(u64)rho * A + t]) : v;
    }
    dft_inv<BA>(a);
#pragma unroll
    for (int j = 0; j < BA; j++) {
        const u64 idx = base + t + (u64)A * j;
        P.y[idx] = gl_mul(a[j], P.itw4[idx]);
    }
}

// ------------------------------------------------------------------ small residues
// One workgroup of B2 threads runs `iters` whole iterations on n = B1 * B2 words kept
// in LDS (natural order, field elements).  LDS words: max(n, B1 * (B2 + PAD)).
// MODE 0: `iters` squarings.  MODE 1 (iters = 1): forward transform of x into y (slot
// t*B1 + i), x unchanged.  MODE 2 (iters = 1): x <- x * (the operand whose transform is f).
template <int B1, int B2, int M, int MODE = LNT_SQUARE>
LNT_KFN void lnt_small_body(int tid, LView lds, const LntParams &P, u32 iters, bool sub2_first,
                            bool ll) {
    constexpr int RP = B1 / B2;
    constexpr int S = B2 + LNT_PAD;
    constexpr u32 N = B1 * B2;
    const int t = tid;
#pragma unroll
    for (int j = 0; j < B1; j++) lds[(u32)(t + B2 * j)] = P.x[(u64)t + B2 * j];
    LNT_SYNC();
    for (u32 it = 0; it < iters; it++) {
        // tables re-read each iteration (from cache) instead of living in registers
        const u64 *pw = P.w.p, *puw = P.uw.p, *ptw = P.twr.p, *pitw = P.itwr.p;
        LNT_OPAQUE(pw);
        LNT_OPAQUE(puw);
        LNT_OPAQUE(ptw);
        LNT_OPAQUE(pitw);
        const GView<const u64> Tw{pw, P.w.n}, Tuw{puw, P.uw.n}, Ttw{ptw, P.twr.n}, Titw{pitw, P.itwr.n};
        // carry over this thread's chunk [t*B1, (t+1)*B1), streamed in pieces of PK
        // digits; each piece's halo is the M pre-carry values just before it.
        {
            constexpr int PK = B1 < 16 ? B1 : 16;
            const u32 g0 = (u32)t * B1;
            int64_t halo[M];
#pragma unroll
            for (int j = 0; j < M; j++) halo[j] = gl_to_i64(lds[(g0 + N - M + j) & (N - 1)]);
            LNT_SYNC();  // every halo read before any chunk is overwritten
            const bool sub2 = it == 0 ? sub2_first : ll;
            for (int s = 0; s < B1 / PK; s++) {
                int64_t v[M + PK];
#pragma unroll
                for (int j = 0; j < M; j++) v[j] = halo[j];
#pragma unroll
                for (int j = 0; j < PK; j++) v[M + j] = gl_to_i64(lds[g0 + s * PK + j]);
#pragma unroll
                for (int j = 0; j < M; j++) halo[j] = v[PK + j];
                lnt_carry_run<M, PK>(v, P.p, P.log_n, g0 + s * PK, P.levels, sub2);
#pragma unroll
                for (int j = 0; j < PK; j++)
                    lds[g0 + s * PK + j] =
                        P.bias ? (u64)(u32)(v[M + j] + (int64_t(1) << 31)) : gl_from_i64(v[M + j]);
            }
            LNT_SYNC();
        }
        u64 a[B1];
#pragma unroll
        for (int j = 0; j < B1; j++) {
            const u32 i = t + B2 * j;
            a[j] = P.bias ? gl_mul_biased((u32)lds[i], Tw[2 * i], Tw[2 * i + 1]) : gl_mul(lds[i], Tw[2 * i]);
        }
        dft_fwd<B1>(a);
        if (B2 > 1) {
#pragma unroll
            for (int rho = 1; rho < B1; rho++) a[rho] = gl_mul(a[rho], Ttw[(u64)rho * B2 + t]);
        }
        LNT_SYNC();
#pragma unroll
        for (int rho = 0; rho < B1; rho++) lds[(u32)(brev_c<B1>(rho) * S + t)] = a[rho];
        LNT_SYNC();
#pragma unroll
        for (int q = 0; q < RP; q++) {
            const int k2 = t * RP + q;
#pragma unroll
            for (int i1 = 0; i1 < B2; i1++) a[q * B2 + i1] = lds[(u32)(k2 * S + i1)];
            dft_fwd<B2>(a + q * B2);
            if (MODE == LNT_FORWARD) {
#pragma unroll
                for (int i1 = 0; i1 < B2; i1++) P.y[(u64)t * B1 + q * B2 + i1] = a[q * B2 + i1];
                continue;
            }
            if (MODE == LNT_MULTIPLY) {
#pragma unroll
                for (int i1 = 0; i1 < B2; i1++)
                    a[q * B2 + i1] = gl_mul(a[q * B2 + i1], P.f[(u64)t * B1 + q * B2 + i1]);
            } else {
#pragma unroll
                for (int i1 = 0; i1 < B2; i1++) a[q * B2 + i1] = gl_sqr(a[q * B2 + i1]);
            }
            dft_inv<B2>(a + q * B2);
#pragma unroll
            for (int i1 = 0; i1 < B2; i1++) lds[(u32)(k2 * S + i1)] = a[q * B2 + i1];
        }
        if (MODE == LNT_FORWARD) return;
        LNT_SYNC();
#pragma unroll
        for (int rho = 0; rho < B1; rho++) {
            u64 v = lds[(u32)(brev_c<B1>(rho) * S + t)];
            a[rho] = (B2 > 1 && rho > 0) ? gl_mul(v, Titw[(u64)rho * B2 + t]) : v;
        }
        dft_inv<B1>(a);
        LNT_SYNC();
#pragma unroll
        for (int j = 0; j < B1; j++) {
            const u32 i = t + B2 * j;
            lds[i] = gl_mul(a[j], Tuw[i]);
        }
        LNT_SYNC();
    }
#pragma unroll
    for (int j = 0; j < B1; j++) P.x[(u64)t + B2 * j] = lds[(u32)(t + B2 * j)];
}

// ------------------------------------------------------------------ small residues, 3 phases
// One workgroup of B^2 threads runs `iters` iterations on n = B^3 words (B = 16: 4096,
// B = 8: 512), B words per thread in every phase (four waves at n = 4096 where the
// 2-phase kernel runs one).  LDS: the residue in natural order (n words), then the
// transform region (B * B * (B + 1) words, the 3-phase row layout).  MODEs as above.
template <int B>
constexpr u32 lnt_small3_lds() { return (u32)(B * B * B + B * B * (B + 1)); }

template <int B, int M, int MODE = LNT_SQUARE>
LNT_KFN void lnt_small3_body(int tid, LView lds, const LntParams &P, u32 iters, bool sub2_first, bool ll) {
    constexpr int T = B * B;
    constexpr u32 N = B * B * B;
    constexpr int RS = B * (B + 1);
    constexpr u32 Z = N;  // transform region
    const int t = tid, hi = t / B, lo = t % B;
#pragma unroll
    for (int j = 0; j < B; j++) lds[(u32)(t + T * j)] = P.x[(u64)t + T * j];
    LNT_SYNC();
    for (u32 it = 0; it < iters; it++) {
        // tables re-read each iteration (from cache) instead of living in registers
        const u64 *pw = P.w.p, *puw = P.uw.p, *ptw = P.twr.p, *pitw = P.itwr.p, *ptw2 = P.twr2.p,
                  *pitw2 = P.itwr2.p;
        LNT_OPAQUE(pw);
        LNT_OPAQUE(puw);
        LNT_OPAQUE(ptw);
        LNT_OPAQUE(pitw);
        LNT_OPAQUE(ptw2);
        LNT_OPAQUE(pitw2);
        const GView<const u64> Tw{pw, P.w.n}, Tuw{puw, P.uw.n}, Ttw{ptw, P.twr.n}, Titw{pitw, P.itwr.n},
            Ttw2{ptw2, P.twr2.n}, Titw2{pitw2, P.itwr2.n};
        // carry over this thread's chunk [t*B, (t+1)*B); halo = the M pre-carry words before it
        {
            const u32 g0 = (u32)t * B;
            int64_t v[M + B];
#pragma unroll
            for (int j = 0; j < M; j++) v[j] = gl_to_i64(lds[(g0 + N - M + j) & (N - 1)]);
#pragma unroll
            for (int j = 0; j < B; j++) v[M + j] = gl_to_i64(lds[g0 + j]);
            LNT_SYNC();  // every halo read before any chunk is overwritten
            lnt_carry_run<M, B>(v, P.p, P.log_n, g0, P.levels, it == 0 ? sub2_first : ll);
#pragma unroll
            for (int j = 0; j < B; j++)
                lds[g0 + j] = P.bias ? (u64)(u32)(v[M + j] + (int64_t(1) << 31)) : gl_from_i64(v[M + j]);
            LNT_SYNC();
        }
        u64 a[B];
        // phase 1: weight, B-point DFT over j (stride T), twiddle w_n^(t k1)
#pragma unroll
        for (int j = 0; j < B; j++) {
            const u32 i = (u32)(t + T * j);
            a[j] = P.bias ? gl_mul_biased((u32)lds[i], Tw[2 * i], Tw[2 * i + 1]) : gl_mul(lds[i], Tw[2 * i]);
        }
        dft_fwd<B>(a);
#pragma unroll
        for (int rho = 1; rho < B; rho++) a[rho] = gl_mul(a[rho], Ttw[(u64)rho * T + t]);
#pragma unroll
        for (int rho = 0; rho < B; rho++) lds[Z + (u32)(brev_c<B>(rho) * RS + hi * (B + 1) + lo)] = a[rho];
        LNT_SYNC();
        // phase 2: thread (hi, lo) over the middle index
#pragma unroll
        for (int b = 0; b < B; b++) a[b] = lds[Z + (u32)(hi * RS + b * (B + 1) + lo)];
        dft_fwd<B>(a);
#pragma unroll
        for (int rho = 1; rho < B; rho++) a[rho] = gl_mul(a[rho], Ttw2[(u64)rho * B + lo]);
#pragma unroll
        for (int rho = 0; rho < B; rho++) lds[Z + (u32)(hi * RS + brev_c<B>(rho) * (B + 1) + lo)] = a[rho];
        LNT_SYNC();
        // phase 3, pointwise square (or product), inverse phase 3
#pragma unroll
        for (int i = 0; i < B; i++) a[i] = lds[Z + (u32)(hi * RS + lo * (B + 1) + i)];
        dft_fwd<B>(a);
        if (MODE == LNT_FORWARD) {
#pragma unroll
            for (int i = 0; i < B; i++) P.y[(u64)t * B + i] = a[i];
            return;
        }
        if (MODE == LNT_MULTIPLY) {
#pragma unroll
            for (int i = 0; i < B; i++) a[i] = gl_mul(a[i], P.f[(u64)t * B + i]);
        } else {
#pragma unroll
            for (int i = 0; i < B; i++) a[i] = gl_sqr(a[i]);
        }
        dft_inv<B>(a);
#pragma unroll
        for (int i = 0; i < B; i++) lds[Z + (u32)(hi * RS + lo * (B + 1) + i)] = a[i];
        LNT_SYNC();
        // inverse phase 2
#pragma unroll
        for (int rho = 0; rho < B; rho++) {
            u64 v = lds[Z + (u32)(hi * RS + brev_c<B>(rho) * (B + 1) + lo)];
            a[rho] = rho ? gl_mul(v, Titw2[(u64)rho * B + lo]) : v;
        }
        dft_inv<B>(a);
#pragma unroll
        for (int b = 0; b < B; b++) lds[Z + (u32)(hi * RS + b * (B + 1) + lo)] = a[b];
        LNT_SYNC();
        // inverse phase 1, unweight, back to the residue region
#pragma unroll
        for (int rho = 0; rho < B; rho++) {
            u64 v = lds[Z + (u32)(brev_c<B>(rho) * RS + hi * (B + 1) + lo)];
            a[rho] = rho ? gl_mul(v, Titw[(u64)rho * T + t]) : v;
        }
        dft_inv<B>(a);
#pragma unroll
        for (int j = 0; j < B; j++) {
            const u32 i = (u32)(t + T * j);
            lds[i] = gl_mul(a[j], Tuw[i]);
        }
        LNT_SYNC();
    }
#pragma unroll
    for (int j = 0; j < B; j++) P.x[(u64)t + T * j] = lds[(u32)(t + T * j)];
}

// ------------------------------------------------------------------ tables
// Twiddles of a line of L = B1 * B2 words: tw[rho * B2 + t] = omega_L^(t * brev_B1(rho)).
GL_HD u64 lnt_line_tw(u64 omega_L, u32 B1, u32 B2, u32 idx) {
    u32 rho = idx / B2, t = idx % B2;
    u32 br = 0;
    for (u32 b = 1, v = rho; b < B1; b <<= 1, v >>= 1) br = (br << 1) | (v & 1);
    return gl_pow(omega_L, (u64)t * br);
}

// 4-step twiddle for slot idx = c + C * r', r' = t*B1 + q*B2 + rho' (column line B1 x B2):
// omega_n^(c * kcol), kcol = B1 * brev_B2(rho') + t*(B1/B2) + q.
GL_HD u64 lnt_tw4(u64 omega_n, u32 C, u32 B1, u32 B2, u64 idx) {
    u32 c = (u32)(idx % C);
    u32 rr = (u32)(idx / C);
    u32 t = rr / B1, q = (rr % B1) / B2, rp = rr % B2;
    u32 br = 0;
    for (u32 b = 1, v = rp; b < B2; b <<= 1, v >>= 1) br = (br << 1) | (v & 1);
    u64 kcol = (u64)B1 * br + (u64)t * (B1 / B2) + q;
    return gl_pow(omega_n, (u64)c * kcol);
}

// 4-step twiddle of a 3-phase column (radices BA, BB, BC): slot s = t*BA + q*BC + rho
// holds frequency k1 + BA (k2 + BB brev_BC(rho)), (k1, k2) = (I / BB, I % BB), I = q*A + t.
GL_HD u64 lnt_tw4_3(u64 omega_n, u32 C, u32 BA, u32 BB, u32 BC, u64 idx) {
    const u32 c = (u32)(idx % C), s = (u32)(idx / C), A = BB * BC;
    const u32 t = s / BA, q = (s % BA) / BC, rho = s % BC;
    const u32 I = q * A + t, k1 = I / BB, k2 = I % BB;
    u32 br = 0;
    for (u32 b = 1, v = rho; b < BC; b <<= 1, v >>= 1) br = (br << 1) | (v & 1);
    const u64 kcol = k1 + (u64)BA * (k2 + (u64)BB * br);
    return gl_pow(omega_n, (u64)c * kcol);
}

// Large-path weight tables (the factorisation at lnt_wexp): row slot 2r + h holds B^f_r,
// times m when h = 1 (B = weight base, m = 1/2; or B = its inverse, m = 2 for unweights).
GL_HD u64 lnt_row_factor(u64 B, u64 m, u32 p, u32 n, u32 C, u32 slot) {
    const u64 v = gl_pow(B, lnt_wexp(p, n, (u64)C * (slot >> 1)));
    return (slot & 1) ? gl_mul(v, m) : v;
}
// Word i of the large-path weight table: the pair (w, -2^31 w) of slot i / 2
// (gl_mul_biased takes both).
GL_HD u64 lnt_weight_pair(u64 Bw, u32 p, u32 n, u32 C, u32 i) {
    const u64 w = lnt_row_factor(Bw, (GL_P + 1) / 2, p, n, C, i >> 1);
    return (i & 1) ? gl_mul(w, GL_P - (1ull << 31)) : w;
}
// Column factor B^e_c times s (s = 1 for weights, n^-1 for unweights), c = idx mod C.
GL_HD u64 lnt_col_factor(u64 B, u64 s, u32 p, u32 n, u32 C, u64 idx) {
    return gl_mul(gl_pow(B, lnt_wexp(p, n, idx % C)), s);
}

// IBDWT weight of digit i (common/pre_calculation.cu:23-36): B^(n - (p i mod n)) for
// i > 0 and 1 for i = 0, with B = 7^((P-1)/192/n*5) (so B^n = 2).
GL_HD u64 lnt_weight(u64 Bw, u32 p, u32 n, u32 i) {
    if (i == 0) return 1;
    u32 r = (u32)((u64)p * i) & (n - 1);
    return gl_pow(Bw, (u64)(n - r));
}

/*
 * TEST INFRASTRUCTURE ONLY.  This file is the oracle: a plain-C restatement of the
 * reference's per-iteration squaring mod 2^p - 1 (LucasNTT/cuda).  Only tests/,
 * __graft_entry__.smoke() and bench.py's cpu_baseline leg load it, and only as the
 * checker (or as the timed CPU baseline).  The product never links it.
 *
 * Every function cites the reference file:line it restates.  Pinning (see tests/):
 *   - field ops           : Python-int arithmetic mod P on random and edge operands
 *   - secured length      : exact integer restatement of the overflow bound, dense sweep
 *   - weights/widths      : B^n = 2, w*uw*n = 1, sum(widths) = p, closed-form ceil(p*i/n)
 *                           (the reference's double expression equals the integer ceil
 *                           while p*(i+1) <= 2^53; above that, at n >= 2^25, it rounds and
 *                           lo_set_exact_widths switches to the integer form, pinned by GMP)
 *   - iterate (LL / PRP)  : GMP mpz residues (oracle/gmp_ll.c fixtures), brute-force
 *                           Python-int runs on tiny p, known Mersenne primes
 *   - carry / minus-two   : brute force vs V mod M_p, including the wrap and borrow cases
 * A fault-injection knob (lo_set_fault) lets tests plant a mistake here and check
 * that the pins above catch it.
 */
#include <math.h>
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#ifdef _OPENMP
#include <omp.h>
#endif

#define P_MOD 0xFFFFFFFF00000001ULL /* common/arithmetics.cuh:26 */

typedef unsigned __int128 u128;

/* ---------------------------------------------------------------- field ops */

/* common/arithmetics.cuh:37-44 (x + y computed as y - (P - x)) */
uint64_t lo_add(uint64_t x, uint64_t y) {
    x = P_MOD - x;
    uint64_t d = y - x;
    if (y < x) d += P_MOD;
    return d;
}

/* common/arithmetics.cuh:46-52 */
uint64_t lo_sub(uint64_t x, uint64_t y) {
    uint64_t d = x - y;
    if (x < y) d += P_MOD;
    return d;
}

/* common/arithmetics.cuh:66-87: 128-bit product, 2^64 = 2^32 - 1 and 2^96 = -1 (mod P).
 * Restated with the compiler's 128-bit type for the product. */
uint64_t lo_mul(uint64_t x, uint64_t y) {
    u128 r = (u128)x * y;
    uint64_t reslo = (uint64_t)r, reshi = (uint64_t)(r >> 64);
    if (reslo >= P_MOD) reslo -= P_MOD;
    uint64_t a = (uint32_t)reshi + (reshi >> 32);
    uint64_t d = a - (reshi << 32);
    if (a < (reshi << 32)) d += P_MOD;
    a = reslo - d;
    if (reslo < d) a += P_MOD;
    return a;
}

/* common/arithmetics.cuh:161-171 */
uint64_t lo_pow(uint64_t a, uint64_t e) {
    uint64_t r = 1;
    while (e) {
        if (e & 1) r = lo_mul(r, a);
        e >>= 1;
        a = lo_mul(a, a);
    }
    return r;
}

/* common/arithmetics.cuh:182-185 (Fermat) */
uint64_t lo_inv(uint64_t a) { return lo_pow(a, P_MOD - 2); }

/* --------------------------------------------------------- length selection */

/* ntt_lengths.cu:88-98, restated with the same double-precision operations. */
uint32_t lo_secured_length(uint32_t exponent) {
    double log_n = ceil(log2(1.0 * exponent / 32));
    double m = exp2(ceil(1.0 * exponent / exp2(log_n))) - 1;
    if (m * m >= 1.0 * (double)P_MOD / (2 * exp2(log_n))) log_n++;
    return (uint32_t)log_n;
}

/* ntt_lengths.cu:100-128.  Writes the candidate strings, ';'-separated, exactly as
 * the reference builds them (std::to_string(double) gives "%f" formatting). */
int lo_candidates(uint32_t exponent, char *buf, int buflen) {
    uint32_t l = lo_secured_length(exponent);
    char tmp[1024];
    int k = 0;
    k += snprintf(tmp + k, sizeof tmp - k, "%f", exp2(l));
    if (l >= 10 && l <= 24) {
        uint32_t a = l / 2;
        k += snprintf(tmp + k, sizeof tmp - k, ";%f:%f", exp2(a), exp2(l - a));
        if (l - a != a) k += snprintf(tmp + k, sizeof tmp - k, ";%f:%f", exp2(l - a), exp2(a));
    }
    if (l == 23) k += snprintf(tmp + k, sizeof tmp - k, ";256:256:128;512:256:64");
    if (l == 24) k += snprintf(tmp + k, sizeof tmp - k, ";512:256:128;512:512:64");
    if (l == 25) k += snprintf(tmp + k, sizeof tmp - k, ";512:512:128;1024:256:128;1024:512:64");
    if (l == 26) k += snprintf(tmp + k, sizeof tmp - k, ";512:512:256;1024:256:256;1024:512:128");
    if (k + 1 > buflen) return -1;
    memcpy(buf, tmp, k + 1);
    return k;
}

/* ------------------------------------------------------------------- tables */

/* common/pre_calculation.cu:23-36: IBDWT weights B^(n - (p*i mod n)) with
 * B = 7^((P-1)/192/n*5), unweights (w*n)^-1, and digit widths
 * ceil(p(i+1)/n) - ceil(p*i/n) (same double expression as the reference). */
void lo_tables(uint32_t exponent, uint32_t n, uint64_t *weights, uint64_t *unweights, uint8_t *widths) {
#pragma omp parallel for schedule(static) if (n >= (1u << 16))
    for (int64_t tid = 0; tid < (int64_t)n; tid++) {
        if (weights) {
            if (tid == 0) {
                weights[0] = 1;
                unweights[0] = lo_inv(n);
            } else {
                uint32_t r = (uint32_t)(exponent * (uint32_t)tid) & (n - 1);
                uint64_t e = ((P_MOD - 1) / 192 / n) * 5ULL * (uint64_t)(n - r);
                weights[tid] = lo_pow(7, e);
                unweights[tid] = lo_inv(lo_mul(weights[tid], n));
            }
        }
        widths[tid] = (uint8_t)(ceil(1.0 * exponent * (tid + 1) / n) - ceil(1.0 * exponent * tid / n));
    }
}

/* Digit widths by integer arithmetic: ceil(p(i+1)/n) - ceil(p i/n) exactly.  Equal to
 * the reference's double expression above whenever p * n <= 2^53; above that (possible
 * only at n >= 2^25) the double products round, the expression can return a width one
 * off at digits with tiny p*i mod n, and the weights (exact integers) then no longer
 * match the digit layout, so the squaring is not x^2 mod M_p.  Not a reference
 * restatement: the exact form the weights assume, checked against GMP. */
void lo_exact_widths(uint32_t exponent, uint32_t n, uint8_t *widths) {
#pragma omp parallel for schedule(static) if (n >= (1u << 16))
    for (int64_t tid = 0; tid < (int64_t)n; tid++) {
        uint64_t a = ((uint64_t)exponent * (uint64_t)(tid + 1) + n - 1) / n;
        uint64_t b = ((uint64_t)exponent * (uint64_t)tid + n - 1) / n;
        widths[tid] = (uint8_t)(a - b);
    }
}

/* ------------------------------------------------------------------ context */

typedef struct lo_ctx {
    uint32_t p, n, n0, n1, log_n0, log_n1;
    int prp;
    int fault;
    uint64_t *x;                 /* n digits, natural order (lucas_prp.cu:91) */
    uint64_t *weights, *unweights;
    uint8_t *widths;
    uint64_t *roots0, *iroots0;  /* omega_{n0}^j, j < n0/2 */
    uint64_t *roots1, *iroots1;  /* omega_{n1}^j, j < n1/2 */
    uint64_t *tw, *itw;          /* 4-step twiddles, n entries (kernel.cu:199-206) */
} lo_ctx;

static uint32_t ilog2(uint32_t v) { uint32_t l = 0; while ((1u << l) < v) l++; return l; }

static uint32_t brev(uint32_t v, uint32_t bits) {
    uint32_t r = 0;
    for (uint32_t b = 0; b < bits; b++) r |= ((v >> b) & 1u) << (bits - 1 - b);
    return r;
}

/* common/pre_calculation.cu:61-66: omega_L = 554^((P-1)/L). */
static void make_roots(uint32_t L, uint64_t **r, uint64_t **ir) {
    uint32_t h = L / 2 ? L / 2 : 1;
    *r = malloc(sizeof(uint64_t) * h);
    *ir = malloc(sizeof(uint64_t) * h);
    uint64_t w = lo_pow(554, (P_MOD - 1) / L);
    uint64_t iw = lo_inv(w);
    uint64_t a = 1, b = 1;
    for (uint32_t j = 0; j < h; j++) { (*r)[j] = a; (*ir)[j] = b; a = lo_mul(a, w); b = lo_mul(b, iw); }
}

/* fault kinds planted by tests; 0 = none */
enum { F_NONE = 0, F_TWIDDLE_INDEX = 1, F_NO_WRAP = 2, F_NO_MINUS2 = 3, F_WEIGHT_INDEX = 4 };
void lo_set_fault(lo_ctx *c, int kind) { c->fault = kind; }

/* n0 = 1 selects the one-transform chain (FFT_1SMEM / FFT_GMEM: kernel_pool.cu:23-38);
 * otherwise n = n0 * n1 is the 2SMEM chain (kernel_pool.cu:39-48). */
lo_ctx *lo_create(uint32_t p, uint32_t n0, uint32_t n1, int prp) {
    lo_ctx *c = calloc(1, sizeof *c);
    c->p = p; c->n0 = n0; c->n1 = n1; c->n = n0 * n1; c->prp = prp;
    c->log_n0 = ilog2(n0); c->log_n1 = ilog2(n1);
    uint32_t n = c->n;
    c->x = calloc(n, 8);
    c->weights = malloc(8ull * n); c->unweights = malloc(8ull * n); c->widths = malloc(n);
    lo_tables(p, n, c->weights, c->unweights, c->widths);
    make_roots(n1, &c->roots1, &c->iroots1);
    if (n0 > 1) {
        make_roots(n0, &c->roots0, &c->iroots0);
        /* common/pre_calculation.cu:69-77 with stride = n1, length = n0:
         * tw[t] = omega_n^((t mod n1) * brev_{log n0}(t / n1)). */
        c->tw = malloc(8ull * n); c->itw = malloc(8ull * n);
        uint64_t w = lo_pow(554, (P_MOD - 1) / n);
#pragma omp parallel for schedule(static) if (n >= (1u << 16))
        for (int64_t t = 0; t < (int64_t)n; t++) {
            uint32_t i = (uint32_t)t & (n1 - 1);
            uint32_t j = brev((uint32_t)t / n1, c->log_n0);
            c->tw[t] = lo_pow(w, (uint64_t)i * j);
            c->itw[t] = lo_inv(c->tw[t]);
        }
    }
    c->x[0] = prp ? 3 : 4; /* lucas_prp.cu:127-134 */
    return c;
}

void lo_destroy(lo_ctx *c) {
    if (!c) return;
    free(c->x); free(c->weights); free(c->unweights); free(c->widths);
    free(c->roots0); free(c->iroots0); free(c->roots1); free(c->iroots1);
    free(c->tw); free(c->itw);
    free(c);
}

uint32_t lo_n(const lo_ctx *c) { return c->n; }
/* Switch a fresh context to exact integer widths (call before setting digits). */
void lo_set_exact_widths(lo_ctx *c) { lo_exact_widths(c->p, c->n, c->widths); }
void lo_set_digits(lo_ctx *c, const uint64_t *d) { memcpy(c->x, d, 8ull * c->n); }
void lo_get_digits(const lo_ctx *c, uint64_t *d) { memcpy(d, c->x, 8ull * c->n); }
void lo_get_widths(const lo_ctx *c, uint8_t *w) { memcpy(w, c->widths, c->n); }

/* --------------------------------------------------------------- transforms */

/* Radix-2 forward transform, natural order in, bit-reversed order out, the output
 * order of fft_radix2_pease_inplace (core/fft_r2.cuh:17-66).  roots[j] = omega_L^j. */
static void ntt_fwd(uint64_t *a, uint32_t L, const uint64_t *roots, int fault) {
    for (uint32_t len = L; len >= 2; len >>= 1) {
        uint32_t half = len >> 1, step = L / len;
        for (uint32_t s = 0; s < L; s += len)
            for (uint32_t j = 0; j < half; j++) {
                uint64_t u = a[s + j], v = a[s + j + half];
                a[s + j] = lo_add(u, v);
                uint32_t ti = j * step + (fault == F_TWIDDLE_INDEX && j > 0 ? 1 : 0);
                a[s + j + half] = lo_mul(lo_sub(u, v), roots[ti]);
            }
    }
}

/* Radix-2 inverse transform, bit-reversed in, natural out, no 1/L
 * (ifft_radix2_pease_inplace, core/fft_r2.cuh:68-107).  iroots[j] = omega_L^-j. */
static void ntt_inv(uint64_t *a, uint32_t L, const uint64_t *iroots) {
    for (uint32_t len = 2; len <= L; len <<= 1) {
        uint32_t half = len >> 1, step = L / len;
        for (uint32_t s = 0; s < L; s += len)
            for (uint32_t j = 0; j < half; j++) {
                uint64_t u = a[s + j], v = lo_mul(a[s + j + half], iroots[j * step]);
                a[s + j] = lo_add(u, v);
                a[s + j + half] = lo_sub(u, v);
            }
    }
}

/* kernels/r2b.cu:25-50 restated sequentially.  The reference's atomic ripple keeps
 * every digit below 2^width and folds the carry out of the top digit into digit 0
 * (2^p = 1 mod M_p).  Its end state depends only on the value S of the input: the
 * carries stop as soon as the value drops below 2^p, so the result is S mod M_p in
 * (0, M_p] for S > 0 (all-ones when S = 0 mod M_p) and 0 for S = 0.  A sequential
 * ripple reaches the same unique state. */
static void carry(lo_ctx *c) {
    uint64_t *x = c->x;
    const uint8_t *w = c->widths;
    uint32_t n = c->n;
    uint64_t cy = 0;
    for (uint32_t i = 0; i < n; i++) {
        u128 v = (u128)x[i] + cy;
        x[i] = (uint64_t)(v & ((1ull << w[i]) - 1));
        cy = (uint64_t)(v >> w[i]);
    }
    if (c->fault == F_NO_WRAP) return;
    uint32_t i = 0;
    while (cy) {
        u128 v = (u128)x[i] + cy;
        x[i] = (uint64_t)(v & ((1ull << w[i]) - 1));
        cy = (uint64_t)(v >> w[i]);
        i = (i + 1 == n) ? 0 : i + 1;
    }
}

/* kernels/m2.cu:23-41: subtract 2 from digit 0 with a borrow ripple that wraps. */
static void minus_two(lo_ctx *c) {
    uint64_t *x = c->x;
    const uint8_t *w = c->widths;
    uint32_t n1 = c->n - 1;
    uint64_t a = x[0];
    x[0] = a - 2;
    if (a < 2) {
        x[0] &= (1ull << w[0]) - 1;
        uint32_t id = 0;
        int cf = 1;
        while (cf) {
            id = (id == n1) ? 0 : id + 1;
            uint64_t m = (1ull << w[id]) - 1;
            a = x[id];
            x[id] = (a - 1) & m;
            cf = a == 0;
        }
    }
}

/* One iteration, LucasPRP::OneIteration (lucas_prp.cu:227-229) over the chain. */
static void one_iteration(lo_ctx *c) {
    uint32_t n = c->n, n0 = c->n0, n1 = c->n1;
    uint64_t *x = c->x;
    int fault = c->fault;
    if (n0 == 1) {
        /* kernels/w_fft_sqr_ifft_uw__r2_8to4k.cu:20-55: weight, NTT, square, INTT, unweight */
        for (uint32_t i = 0; i < n; i++)
            x[i] = lo_mul(x[i], c->weights[fault == F_WEIGHT_INDEX ? (i ^ 1) : i]);
        ntt_fwd(x, n, c->roots1, fault);
        for (uint32_t i = 0; i < n; i++) x[i] = lo_mul(x[i], x[i]);
        ntt_inv(x, n, c->iroots1);
        for (uint32_t i = 0; i < n; i++) x[i] = lo_mul(x[i], c->unweights[i]);
    } else {
        /* kernels/w_fftc_tw__r2_8to4k.cu:20-46: column c (stride n1): weight, NTT_n0, twiddle */
#pragma omp parallel if (n >= (1u << 16))
        {
            uint64_t *col = malloc(8ull * n0);
#pragma omp for schedule(static)
            for (int64_t cc = 0; cc < (int64_t)n1; cc++) {
                for (uint32_t r = 0; r < n0; r++) {
                    uint64_t idx = cc + (uint64_t)n1 * r;
                    col[r] = lo_mul(x[idx], c->weights[fault == F_WEIGHT_INDEX ? (idx ^ 1) : idx]);
                }
                ntt_fwd(col, n0, c->roots0, fault);
                for (uint32_t r = 0; r < n0; r++) {
                    uint64_t idx = cc + (uint64_t)n1 * r;
                    x[idx] = lo_mul(col[r], c->tw[idx]);
                }
            }
            free(col);
        }
        /* kernels/fftr_sqr_ifftr__r2_8to4k.cu:20-56: row: NTT_n1, square, INTT_n1 */
#pragma omp parallel for schedule(static) if (n >= (1u << 16))
        for (int64_t r = 0; r < (int64_t)n0; r++) {
            uint64_t *row = x + (uint64_t)n1 * r;
            ntt_fwd(row, n1, c->roots1, fault);
            for (uint32_t i = 0; i < n1; i++) row[i] = lo_mul(row[i], row[i]);
            ntt_inv(row, n1, c->iroots1);
        }
        /* kernels/tw_ifftc_uw__r2_8to4k.cu:20-46: inverse twiddle, INTT_n0, unweight */
#pragma omp parallel if (n >= (1u << 16))
        {
            uint64_t *col = malloc(8ull * n0);
#pragma omp for schedule(static)
            for (int64_t cc = 0; cc < (int64_t)n1; cc++) {
                for (uint32_t r = 0; r < n0; r++) {
                    uint64_t idx = cc + (uint64_t)n1 * r;
                    col[r] = lo_mul(x[idx], c->itw[idx]);
                }
                ntt_inv(col, n0, c->iroots0);
                for (uint32_t r = 0; r < n0; r++) {
                    uint64_t idx = cc + (uint64_t)n1 * r;
                    x[idx] = lo_mul(col[r], c->unweights[idx]);
                }
            }
            free(col);
        }
    }
    carry(c);                                      /* r2b (kernel_pool.cu:36,46) */
    if (!c->prp && fault != F_NO_MINUS2) minus_two(c); /* m2, LL only (kernel_pool.cu:37,47) */
}

void lo_iterate(lo_ctx *c, uint32_t k) {
    for (uint32_t i = 0; i < k; i++) one_iteration(c);
}

/* Low 64 bits of the residue value sum d_i 2^e_i, e_i = ceil(p*i/n).  The all-ones
 * digit vector is M_p = 0 (mod M_p) and reports 0. */
uint64_t lo_res64_digits(uint32_t p, uint32_t n, const uint64_t *x, const uint8_t *w) {
    int all_ones = 1;
    for (uint32_t i = 0; i < n && all_ones; i++)
        if (x[i] != (1ull << w[i]) - 1) all_ones = 0;
    if (all_ones) return 0;
    uint64_t r = 0;
    uint32_t e = 0;
    for (uint32_t i = 0; i < n && e < 64; i++) {
        r += x[i] << e;
        e += w[i];
    }
    (void)p;
    return r;
}

uint64_t lo_res64(const lo_ctx *c) { return lo_res64_digits(c->p, c->n, c->x, c->widths); }

int lo_threads(void) {
#ifdef _OPENMP
    return omp_get_max_threads();
#else
    return 1;
#endif
}

/*
 * TEST INFRASTRUCTURE ONLY -- the big-integer oracle for the LL/PRP sequence.
 *
 * Computes s_k mod M_p with GMP's mpz layer, independently of any NTT:
 *   LL : s_0 = 4, s_k = s_{k-1}^2 - 2 mod (2^p - 1)   (lucas_prp.cu:129, kernels/m2.cu:23-41)
 *   PRP: s_0 = 3, s_k = s_{k-1}^2     mod (2^p - 1)   (lucas_prp.cu:129)
 * Iteration k (1-based) leaves s_k, matching LucasPRP::MainLoop (lucas_prp.cu:205-224).
 * The residue is kept canonical in [0, M_p); res64 is its low 64 bits.
 *
 * There is no gmp.h in the image, so the handful of documented mpz entry points
 * used here are declared by hand (GMP manual, "Integer Functions"); mpz_t is the
 * documented {int alloc; int size; limb*} struct.  Link against libgmp.so.10.
 *
 * usage: gmp_ll <p> <LL|PRP> <iterations> <every> [dump_at] [block]
 *   prints "k res64hex" every <every> iterations and at <iterations>;
 *   with dump_at (0 = none), writes the full residue at that k as little-endian u64
 *   words to the file "gmp_p<p>_k<k>.bin";
 *   with block B, also keeps the Gerbicz product D = s_0 * s_B * s_2B * ... (mod M_p) and
 *   prints "k res64hex dres64hex" (D after folding in s_k when B divides k).
 */
#include <stdint.h>
#include <stdio.h>
#include <stdlib.h>
#include <string.h>
#include <time.h>

typedef struct { int alloc; int size; uint64_t *d; } mpz_struct;
typedef mpz_struct mpz_t[1];

extern void __gmpz_init(mpz_struct *);
extern void __gmpz_clear(mpz_struct *);
extern void __gmpz_set_ui(mpz_struct *, unsigned long);
extern void __gmpz_set(mpz_struct *, const mpz_struct *);
extern void __gmpz_mul(mpz_struct *, const mpz_struct *, const mpz_struct *);
extern void __gmpz_add(mpz_struct *, const mpz_struct *, const mpz_struct *);
extern void __gmpz_sub(mpz_struct *, const mpz_struct *, const mpz_struct *);
extern void __gmpz_sub_ui(mpz_struct *, const mpz_struct *, unsigned long);
extern void __gmpz_tdiv_r_2exp(mpz_struct *, const mpz_struct *, unsigned long);
extern void __gmpz_fdiv_q_2exp(mpz_struct *, const mpz_struct *, unsigned long);
extern void __gmpz_setbit(mpz_struct *, unsigned long);
extern int __gmpz_cmp(const mpz_struct *, const mpz_struct *);
extern unsigned long __gmpz_get_ui(const mpz_struct *);
extern void *__gmpz_export(void *, size_t *, int, size_t, int, size_t, const mpz_struct *);

/* x <- x mod (2^p - 1), x >= 0, result canonical in [0, M). */
static void mod_mersenne(mpz_struct *x, mpz_struct *t, const mpz_struct *M, unsigned long p) {
    while (x->size > 0 && __gmpz_cmp(x, M) >= 0) {
        __gmpz_fdiv_q_2exp(t, x, p);
        __gmpz_tdiv_r_2exp(x, x, p);
        __gmpz_add(x, x, t);
        if (__gmpz_cmp(x, M) == 0) __gmpz_set_ui(x, 0);
    }
}

int main(int argc, char **argv) {
    if (argc < 5) {
        fprintf(stderr, "usage: %s <p> <LL|PRP> <iterations> <every> [dump_at]\n", argv[0]);
        return 2;
    }
    unsigned long p = strtoul(argv[1], NULL, 10);
    int prp = strcmp(argv[2], "PRP") == 0;
    unsigned long iters = strtoul(argv[3], NULL, 10);
    unsigned long every = strtoul(argv[4], NULL, 10);
    unsigned long dump_at = argc > 5 ? strtoul(argv[5], NULL, 10) : 0;
    unsigned long block = argc > 6 ? strtoul(argv[6], NULL, 10) : 0;
    mpz_t s, t, M, two, D;
    __gmpz_init(s); __gmpz_init(t); __gmpz_init(M); __gmpz_init(two); __gmpz_init(D);
    __gmpz_set_ui(M, 0); __gmpz_setbit(M, p); __gmpz_sub_ui(M, M, 1);
    __gmpz_set_ui(s, prp ? 3 : 4);
    __gmpz_set_ui(two, 2);
    __gmpz_set(D, s);
    time_t t0 = time(NULL);
    for (unsigned long k = 1; k <= iters; k++) {
        __gmpz_mul(s, s, s);
        mod_mersenne(s, t, M, p);
        if (!prp) {
            if (__gmpz_cmp(s, two) >= 0) __gmpz_sub_ui(s, s, 2);
            else { __gmpz_add(s, s, M); __gmpz_sub_ui(s, s, 2); }
        }
        if (block && k % block == 0) {
            __gmpz_mul(D, D, s);
            mod_mersenne(D, t, M, p);
        }
        if ((every && k % every == 0) || k == iters) {
            unsigned long lo = s->size == 0 ? 0 : s->d[0];
            if (block)
                printf("%lu %016lX %016lX\n", k, lo, D->size == 0 ? 0UL : D->d[0]);
            else
                printf("%lu %016lX\n", k, lo);
            fflush(stdout);
        }
        if (dump_at && k == dump_at) {
            char name[128];
            snprintf(name, sizeof name, "gmp_p%lu_k%lu.bin", p, k);
            size_t words = (p + 63) / 64, cnt = 0;
            uint64_t *buf = calloc(words, 8);
            __gmpz_export(buf, &cnt, -1, 8, -1, 0, s);
            FILE *f = fopen(name, "wb");
            fwrite(buf, 8, words, f);
            fclose(f);
            free(buf);
        }
    }
    fprintf(stderr, "done p=%lu %s k=%lu in %ld s\n", p, prp ? "PRP" : "LL", iters, (long)(time(NULL) - t0));
    return 0;
}

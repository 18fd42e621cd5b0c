/*
 * liblnt -- C ABI of the MI355X (gfx950) Lucas-Lehmer / PRP squaring engine.
 *
 * Plain C types only: pointers, sizes, integers.  No exceptions cross this boundary;
 * every call returns an lnt_status and lnt_last_error() gives the message.  The C++
 * command line (cuda_amd/cli) and Python (cuda_amd/lnt.py, ctypes) bind these entry points.
 *
 * The reference drives its hot path through C++ classes on one CUDA stream
 * (LucasPRP, KernelPool, Kernel: lucas_prp.cuh:15-39, kernel.cuh:15-73); each entry
 * point below names the reference call site whose role it takes.
 *
 * Residues cross the ABI as "natural digits": n unsigned words, digit i holding bits
 * [ceil(p*i/n), ceil(p*(i+1)/n)) of the residue, exactly the order of the reference's
 * .dat dump (lucas_prp.cu:254-263).
 */
#ifndef LNT_API_H
#define LNT_API_H

#include <stddef.h>
#include <stdint.h>

#ifdef __cplusplus
extern "C" {
#endif

typedef struct lnt_ctx lnt_ctx;

typedef enum {
    LNT_OK = 0,
    LNT_UNSUPPORTED = 1, /* shape not supported (reference: Initialize() == false, process.cu:53) */
    LNT_DEVICE_ERROR = 2,
    LNT_OUT_OF_MEMORY = 3,
    LNT_INVALID_ARGUMENT = 4
} lnt_status;

typedef enum { LNT_MODE_LL = 0, LNT_MODE_PRP = 1 } lnt_mode;

/* Create a test context on `device` for M_p = 2^p - 1 with NTT length
 * n = factors[0] * ... * factors[nfactors-1] (a power of two, 8 <= n <= 2^26).
 * Allocates the residue and all tables and seeds the residue with 4 (LL) or 3 (PRP).
 * Replaces LucasPRP::LucasPRP + LucasPRP::Initialize (lucas_prp.cu:24-111) together
 * with the pool/kernel Initialize() table set-up (kernel.cu:91-153, 199-255) and
 * InitializeTest (lucas_prp.cu:127-134).  The factor split is a hint only: the engine
 * chooses its own row/column split (the role of KernelPool, kernel_pool.cu:17-141).
 * Digit i has width ceil(p(i+1)/n) - ceil(p i/n), computed exactly; the reference's
 * double-precision form (common/pre_calculation.cu:35) agrees whenever p * n <= 2^53,
 * which covers every exponent up to n = 2^24. */
int lnt_create(int device, uint32_t p, const uint32_t *factors, int nfactors, int mode,
               lnt_ctx **out);

/* Replace the residue with n natural digits (LucasPRP::LoadResidue, lucas_prp.cu:326-346). */
int lnt_set_digits(lnt_ctx *ctx, const uint64_t *natural_digits, size_t n);

/* Enqueue k iterations (square, and -2 for LL) on the context's stream; asynchronous.
 * Replaces k calls of LucasPRP::OneIteration (lucas_prp.cu:227-229). */
int lnt_iterate(lnt_ctx *ctx, uint32_t k);

/* Wait for the stream and copy out the canonical residue as n natural digits
 * (the cudaMemcpy + DumpResidue order of lucas_prp.cu:164, 254-263). */
int lnt_get_digits(lnt_ctx *ctx, uint64_t *natural_digits, size_t n);

/* Low 64 bits of the canonical residue (new: the reference prints no res64). */
int lnt_res64(lnt_ctx *ctx, uint64_t *out);

/* Wait for the stream; *elapsed_ms = GPU time since the previous lnt_sync (or create).
 * Replaces the event pair of LucasPRP::Run / DisplayPitStop (lucas_prp.cu:151-161, 232-234). */
int lnt_sync(lnt_ctx *ctx, float *elapsed_ms);

/* Time `loops` iterations on scratch state for every valid engine configuration of this
 * length, validate each against the current configuration first, keep the fastest (the
 * current configuration unless another is faster by more than 2%).
 * The residue is preserved.  Replaces Process::GetBestPerformers / EvaluatePerformance
 * (process.cu:45-149), which timed candidates without checking their results. */
int lnt_tune(lnt_ctx *ctx, int loops);

/* Every candidate timed by the last lnt_tune on ctx, one "name = ms" line each (ms for
 * `loops` iterations), the names being the lnt_describe strings with '/' between the
 * parts: the record the reference keeps in its performance map and writes to the stats
 * file as "Candidate Performance (in ms for N loops): name = t" (common/common.cu:83-85,
 * 108-112).  Empty before the first lnt_tune. */
int lnt_tune_results(const lnt_ctx *ctx, char *buf, size_t len);

/* Free everything (LucasPRP::Finalize, lucas_prp.cu:113-125). */
void lnt_destroy(lnt_ctx *ctx);

/* Message of the last failure on ctx (or of the last failed lnt_create when ctx is NULL). */
const char *lnt_last_error(const lnt_ctx *ctx);

/* ---- host-side helpers (no device needed) ---- */

/* log2 of the NTT length the reference selects for exponent p
 * (NttLengths::GetNttSecuredLength, ntt_lengths.cu:88-98), in exact integer arithmetic. */
uint32_t lnt_secured_log2_length(uint32_t p);

/* The tuner's choices as text, one line per device and length: "device n lc K row3 col3"
 * (new: the reference re-tunes at every start, process.cu:45-127).  Export fills buf
 * (LNT_INVALID_ARGUMENT if it is too small); import merges lines into the choices used
 * by later lnt_create calls.  Both are host-only. */
int lnt_tune_export(char *buf, size_t len);
int lnt_tune_import(const char *text);

/* Self-test of the device field arithmetic (new): add, sub, mul, square and the shift
 * multiplies computed on the GPU for `count` operand pairs (edge values, then splitmix64
 * draws from `seed`) and compared with the host forms; *mismatches = differing results. */
int lnt_selftest(int device, uint32_t count, uint64_t seed, uint64_t *mismatches);

/* VALU throughput of the device field arithmetic (new: a measurement aid): each of
 * gl_mul, gl_sqr, gl_add, gl_sub, the three shift-multiply forms and one v_mad_u64_u32 runs
 * as eight independent chains per thread over a grid filling every CU.  ops_per_clk[i] =
 * operations per CU per measured shader clock; names gets "name;name;..."; *ghz the mean
 * measured clock.  Returns the count (up to cap), or a negative lnt_status. */
int lnt_valu_rates(int device, double *ops_per_clk, char *names, size_t names_len, double *ghz, int cap);

/* Number of digits n of a context, and its description: "name;kernel0;kernel1;..." used
 * for the stats file's Pool / Kernel<i> keys (lucas_prp.cu:311-313). */
uint32_t lnt_length(const lnt_ctx *ctx);
int lnt_describe(const lnt_ctx *ctx, char *buf, size_t buflen);

/* Per-kernel average GPU durations (ms): each kernel of the iteration is launched `iters`
 * times back to back as one captured graph (as the iteration graph runs it), timed by two
 * events around the graph launch on the context's stream; the residue is saved before and
 * restored after, so the test's state is unchanged.  names/ms receive up to `cap`
 * entries; returns the count. */
int lnt_profile(lnt_ctx *ctx, uint32_t iters, char *names, size_t names_len, float *ms, int cap);

/* ---- products, copies and the Gerbicz-checked PRP (new: the reference has no error
 * check and no general multiply) ---- */

/* dst <- dst * src mod M_p.  Both contexts must share p, n, mode and row/column split. */
int lnt_mul(lnt_ctx *dst, lnt_ctx *src);

/* dst <- src (residue and its bookkeeping). */
int lnt_copy(lnt_ctx *dst, lnt_ctx *src);

/* Test hook: flip bits (xor_mask, low 46 bits) of the signed value held in residue word
 * `word` on the device, i.e. a single-word memory corruption. */
int lnt_inject_error(lnt_ctx *ctx, uint32_t word, uint64_t xor_mask);

/* PRP with the Gerbicz check: from the current residue u_0, every `block` iterations
 * D <- D * u; every `blocks_per_check` blocks verify D_m == u_0 * D_(m-1)^(2^block), then
 * save (u, D) as the verified checkpoint on the device, or restore the last verified
 * checkpoint and redo the iterations since. */
int lnt_gerbicz_begin(lnt_ctx *ctx, uint32_t block, uint32_t blocks_per_check);
int lnt_gerbicz_iterate(lnt_ctx *ctx, uint32_t k);
/* out[0] iterations since begin, [1] last verified iteration, [2] checks passed,
 * [3] rollbacks, [4] res64 of D at the last verified checkpoint.  cap >= 5. */
int lnt_gerbicz_stats(lnt_ctx *ctx, uint64_t *out, int cap);

/* ---- iterations without device allocation ----
 * No counterpart in the reference.  For callers that cannot allocate device memory (code
 * under HIP graph capture, tight serving loops, harnesses that forbid hipMalloc):
 * k iterations (square, and -2 for LL) of a residue held in the caller's device memory,
 * enqueued on `stream` (a hipStream_t; NULL = the null stream) on the current device.
 * x holds n words, each a field element of GF(2^64 - 2^32 + 1) encoding a signed digit
 * value v_i (the word itself if <= (P-1)/2, else word - P); the residue is
 * sum v_i 2^ceil(p i/n) mod 2^p - 1.  Canonical natural digits (the .dat order) are such
 * a vector, and so is the output, which holds the k-th residue after the call; it is a
 * valid input of the next call.  n is a power of two the engine supports with
 * n <= LNT_STATIC_MAX_N; larger n fails with LNT_INVALID_ARGUMENT.
 * The work buffer and the tables live in a fixed device array of the library
 * (5 * LNT_STATIC_MAX_N + 65536 words, 40.5 MiB, one copy per device, held until the
 * process ends); it keeps the tables of the last (p, n) and rebuilds them when they
 * change.  Calls on one device share that array: calls that overlap in time on a device
 * (other streams, other threads) race, so order them on one stream.  Asynchronous like
 * lnt_iterate; errors are reported by lnt_last_error(NULL). */
#define LNT_STATIC_MAX_N (1u << 20)
int lnt_iterate_static(uint32_t p, int mode, uint64_t *x, size_t n, uint32_t k, void *stream);

/* ---- batches of independent tests (new: the reference runs one test per process) ----
 * Members keep their own residues; members of the same shape run as one batched launch
 * per kernel and one iteration of all members is a single hipGraph launch.  Work issued
 * on a member after lnt_batch_iterate runs after the batch's work.  A context belongs to
 * at most one batch; lnt_destroy takes it out of its batch. */
typedef struct lnt_batch lnt_batch;
int lnt_batch_create(int device, lnt_batch **out);
int lnt_batch_add(lnt_batch *batch, lnt_ctx *ctx);
int lnt_batch_remove(lnt_batch *batch, lnt_ctx *ctx);
int lnt_batch_size(const lnt_batch *batch);
/* out[0] members, [1] kernel-shape groups, [2] iteration graphs captured so far (members
 * joining or leaving a group with free slots only rewrite its parameters).  cap >= 3. */
int lnt_batch_stats(const lnt_batch *batch, uint64_t *out, int cap);
/* Every member advances k iterations (asynchronous). */
int lnt_batch_iterate(lnt_batch *batch, uint32_t k);
/* Wait; *elapsed_ms = GPU time since the previous lnt_batch_sync (or create). */
int lnt_batch_sync(lnt_batch *batch, float *elapsed_ms);
void lnt_batch_destroy(lnt_batch *batch);
const char *lnt_batch_last_error(const lnt_batch *batch);

/* ---- serving loop (new: the reference tests one exponent per process) ----
 * A scheduler thread admits submitted tests while their device footprint fits the
 * budget, advances all admitted tests as one batch, lets arrivals join and finished tests
 * leave between scheduling steps (continuous batching), and queues each result with its
 * timings.  Residues and Gerbicz checkpoints of admitted tests stay on the device. */
typedef struct lnt_server lnt_server;
typedef struct lnt_server_config {
    int device;
    uint64_t hbm_budget;     /* device bytes the admitted tests may hold */
    uint32_t max_tests;      /* admitted tests at most (0: no limit) */
    float step_ms;           /* target GPU time of one scheduling step (0: 5 ms) */
    uint32_t gerbicz_block;  /* PRP tests: Gerbicz block (0: unchecked) */
    uint32_t gerbicz_check;  /* blocks per check (0: 80, or the request's block count if fewer) */
    uint32_t tune_loops;     /* > 0: the first test of each NTT length is tuned (lnt_tune with this
                              * many loops per candidate) before it joins, and later tests of that
                              * length get the tuned plan; 0: default plans (the reference re-tunes
                              * at every start, process.cu:45-127) */
} lnt_server_config;
typedef struct lnt_result {
    uint64_t tag;            /* the caller's id given to lnt_server_submit */
    uint32_t p;
    int32_t prp;
    uint64_t iterations;
    uint64_t res64;
    int32_t is_zero;         /* LL: the whole residue is 0 (M_p prime when iterations = p - 2) */
    int32_t status;          /* lnt_status of this test */
    double submit_s, admit_s, done_s;  /* seconds since lnt_server_create */
    uint64_t checks, rollbacks;        /* Gerbicz checks passed / rollbacks (PRP) */
    uint64_t check_res64;              /* res64 of the Gerbicz product at the last verified check */
} lnt_result;
int lnt_server_create(const lnt_server_config *cfg, lnt_server **out);
/* Queue a test: `iterations` squarings from the start residue (LL: 4, PRP: 3). */
int lnt_server_submit(lnt_server *server, uint32_t p, int prp, uint64_t iterations, uint64_t tag);
/* Move up to cap finished results to out, waiting up to wait_ms for the first; returns
 * the count (negative lnt_status on a bad argument). */
int lnt_server_poll(lnt_server *server, lnt_result *out, int cap, int wait_ms);
/* out[0] queued, [1] admitted and running, [2] finished, [3] device bytes admitted,
 * [4] scheduling steps, [5] iterations per member summed over steps, [6] admitted,
 * [7] (cap >= 8) iteration graphs captured by the server's batches, [8] (cap >= 9) NTT
 * lengths tuned.  cap >= 7. */
int lnt_server_stats(lnt_server *server, uint64_t *out, int cap);
double lnt_server_now(lnt_server *server);
/* Stop the scheduler; tests still running end with status LNT_INVALID_ARGUMENT. */
void lnt_server_destroy(lnt_server *server);
const char *lnt_server_last_error(const lnt_server *server);

#ifdef __cplusplus
}
#endif
#endif /* LNT_API_H */

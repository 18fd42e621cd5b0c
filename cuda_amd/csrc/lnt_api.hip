// C ABI of the engine (include/lnt_api.h): context lifetime, tables, the iteration
// launcher, readout, timing and the validating tuner.
#include <hip/hip_runtime.h>

#include <algorithm>
#include <chrono>
#include <cstdio>
#include <cstdlib>
#include <cstring>
#include <future>
#include <map>
#include <mutex>
#include <string>
#include <system_error>
#include <vector>

#include "../../include/lnt_api.h"
#include "lnt_launch.hpp"
#include "lnt_tables.hpp"

using namespace lnt;

// ------------------------------------------------------------------ context

struct lnt_ctx {
    struct lnt_batch *batch = nullptr;  // the batch this context belongs to, if any
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_mark = nullptr;
    hipEvent_t ev_gz = nullptr;  // end of a Gerbicz block step (a batch waits on it)
    Plan pl;
    uint32_t p = 0;
    bool prp = false;
    bool pending = false;     // x holds pre-carry values whose LL "-2" is still due
    bool pristine = true;     // no iteration since the residue was set
    bool zero_ones = false;   // residue is 0 mod M_p but was given as a nonzero vector
    std::vector<uint64_t> set_digits;  // digits as given (returned while pristine)
    std::vector<uint8_t> wid;
    u64 *x = nullptr, *y = nullptr, *w = nullptr, *uw = nullptr;
    u64 *tw4 = nullptr, *itw4 = nullptr, *twr = nullptr, *itwr = nullptr, *twc = nullptr, *itwc = nullptr;
    u64 *twc2 = nullptr, *itwc2 = nullptr;
    u64 *twr2 = nullptr, *itwr2 = nullptr;
    u64 *f = nullptr;  // transform of the second operand of a product (allocated on first use)
    // GRAPH_ITERS iterations of the large path on one residue buffer (each owing its "-2"
    // when LL), captured once per buffer and replayed: consecutive launches from a graph
    // start sooner than from the stream.  Slot 0 serves the residue x, slot 1 the Gerbicz
    // check's squarings of its copy of D.
    struct IterGraph {
        u64 *buf = nullptr;
        hipGraph_t graph = nullptr;
        hipGraphExec_t exec = nullptr;
    };
    IterGraph graphs[2];
    LntParams P{};
    std::string err;
    std::string tune_report;  // "name = ms" per candidate of the last lnt_tune
    // Gerbicz-checked PRP (lnt_gerbicz_*): residue slots on the device
    struct Slot {
        u64 *x = nullptr;
        bool pending = false;
    };
    struct Gerbicz {
        bool on = false;
        uint32_t B = 0, L = 0;
        uint64_t it = 0, it_v = 0, checks = 0, rollbacks = 0, fails_here = 0;
        uint64_t ck_res64 = 0;
        Slot U0, D, Dprev, E, Uv, Dv;
    } g;
};

static std::mutex g_err_mu;
static std::string g_err;

static int fail(lnt_ctx *c, int code, const std::string &msg) {
    if (c) c->err = msg;
    else {
        std::lock_guard<std::mutex> g(g_err_mu);
        g_err = msg;
    }
    return code;
}

#define HIPCHK(c, call, what)                                                                  \
    do {                                                                                       \
        hipError_t e_ = (call);                                                                \
        if (e_ != hipSuccess)                                                                  \
            return fail((c), e_ == hipErrorOutOfMemory ? LNT_OUT_OF_MEMORY : LNT_DEVICE_ERROR, \
                        std::string(what) + " (" + hipGetErrorString(e_) + " at " __FILE__    \
                        " line " + std::to_string(__LINE__) + ")");                            \
    } while (0)

static void free_buffers(lnt_ctx *c) {
    u64 **bufs[] = {&c->x,      &c->y,      &c->w,      &c->uw,      &c->tw4,     &c->itw4,   &c->twr,
                    &c->itwr,   &c->twc,    &c->itwc,   &c->f,       &c->g.U0.x,  &c->g.D.x,  &c->g.Dprev.x,
                    &c->g.E.x,  &c->g.Uv.x, &c->g.Dv.x, &c->twr2, &c->itwr2, &c->twc2, &c->itwc2};
    for (u64 **b : bufs)
        if (*b) {
            (void)hipFree(*b);
            *b = nullptr;
        }
}

// Allocate and fill every table of plan c->pl on c->stream.
static int build_tables(lnt_ctx *c) {
    const Plan &pl = c->pl;
    const TableSizes z = table_sizes(pl);
    const size_t nb = (size_t)pl.n * 8;
    HIPCHK(c, hipMalloc(&c->x, nb), "hipMalloc (x)");
    HIPCHK(c, hipMalloc(&c->y, nb), "hipMalloc (work buffer)");
    HIPCHK(c, hipMalloc(&c->w, (size_t)z.W * 8), "hipMalloc (weights)");
    HIPCHK(c, hipMalloc(&c->uw, (size_t)z.Wu * 8), "hipMalloc (unweights)");
    HIPCHK(c, hipMalloc(&c->twr, (size_t)z.Lr * 8), "hipMalloc (row twiddles)");
    HIPCHK(c, hipMalloc(&c->itwr, (size_t)z.Lr * 8), "hipMalloc (row twiddles)");
    if (z.large) {
        HIPCHK(c, hipMalloc(&c->tw4, nb), "hipMalloc (4-step twiddles)");
        HIPCHK(c, hipMalloc(&c->itw4, nb), "hipMalloc (4-step twiddles)");
        HIPCHK(c, hipMalloc(&c->twc, (size_t)z.R * 8), "hipMalloc (column twiddles)");
        HIPCHK(c, hipMalloc(&c->itwc, (size_t)z.R * 8), "hipMalloc (column twiddles)");
    }
    if (z.A) {
        HIPCHK(c, hipMalloc(&c->twc2, (size_t)z.A * 8), "hipMalloc (column twiddles)");
        HIPCHK(c, hipMalloc(&c->itwc2, (size_t)z.A * 8), "hipMalloc (column twiddles)");
    }
    if (z.L2) {
        HIPCHK(c, hipMalloc(&c->twr2, (size_t)z.L2 * 8), "hipMalloc (row twiddles)");
        HIPCHK(c, hipMalloc(&c->itwr2, (size_t)z.L2 * 8), "hipMalloc (row twiddles)");
    }
    TablePtrs t;
    t.y = c->y;
    t.w = c->w;
    t.uw = c->uw;
    t.tw4 = c->tw4;
    t.itw4 = c->itw4;
    t.twr = c->twr;
    t.itwr = c->itwr;
    t.twr2 = c->twr2;
    t.itwr2 = c->itwr2;
    t.twc = c->twc;
    t.itwc = c->itwc;
    t.twc2 = c->twc2;
    t.itwc2 = c->itwc2;
    fill_tables(pl, c->p, t, c->stream);
    HIPCHK(c, hipGetLastError(), "table kernels launch");
    c->P = params_of(pl, c->p, c->x, t);
    return LNT_OK;
}

static int upload_digits(lnt_ctx *c, const uint64_t *d) {
    const u32 n = c->pl.n;
    std::vector<uint64_t> xd(n);
    to_device_digits(c->wid, d, xd.data());
    HIPCHK(c, hipMemcpyAsync(c->x, xd.data(), (size_t)n * 8, hipMemcpyHostToDevice, c->stream), "hipMemcpy (x)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->pending = false;
    c->pristine = true;
    c->set_digits.assign(d, d + n);
    // value 0 (mod M_p) given as a nonzero vector: the reference keeps the all-ones form
    std::vector<uint64_t> canon(n);
    bool zero = to_canonical_digits(c->wid, xd.data(), false, canon.data());
    bool all_zero = true;
    for (u32 i = 0; i < n; i++)
        if (d[i]) all_zero = false;
    c->zero_ones = zero && !all_zero;
    return LNT_OK;
}

// Parameters with the residue buffer x, work buffer y and product operand f swapped in.
static LntParams params_for(const lnt_ctx *c, u64 *x, u64 *y, const u64 *f) {
    LntParams P = c->P;
    P.x = view<u64>(x, c->pl.n);
    P.y = view<u64>(y, c->pl.n);
    P.f = view<const u64>(f, f ? c->pl.n : 0);
    return P;
}

static void drop_graph(lnt_ctx::IterGraph &ig) {
    if (ig.exec) (void)hipGraphExecDestroy(ig.exec);
    if (ig.graph) (void)hipGraphDestroy(ig.graph);
    ig = lnt_ctx::IterGraph{};
}

// Drops every captured iteration graph (each names the buffers and kernels of its capture).
static void drop_graph(lnt_ctx *c) {
    for (auto &ig : c->graphs) drop_graph(ig);
}

// Iterations per replay of the captured iteration graph.
static constexpr uint32_t GRAPH_ITERS = 16;

static bool launch_iteration(const Plan &pl, const LntParams &P, bool sub2, hipStream_t s) {
    return lnt_launch_col_fwd(pl, P, sub2, s) && lnt_launch_row(pl, P, s) && lnt_launch_col_inv(pl, P, s);
}

// Capture GRAPH_ITERS iterations of the residue buffer x into ig (carry with "-2" for LL).
static int build_graph(lnt_ctx *c, lnt_ctx::IterGraph &ig, u64 *x) {
    drop_graph(ig);
    const LntParams P = params_for(c, x, c->y, nullptr);
    const bool ll = !c->prp;
    HIPCHK(c, hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    bool ok = true;
    for (uint32_t i = 0; i < GRAPH_ITERS && ok; i++) ok = launch_iteration(c->pl, P, ll, c->stream);
    hipError_t ec = hipStreamEndCapture(c->stream, &ig.graph);
    if (!ok || ec != hipSuccess) {
        drop_graph(ig);
        return fail(c, ok ? LNT_DEVICE_ERROR : LNT_UNSUPPORTED, "iteration graph capture failed");
    }
    if (hipGraphInstantiate(&ig.exec, ig.graph, nullptr, nullptr, 0) != hipSuccess) {
        drop_graph(ig);
        return fail(c, LNT_DEVICE_ERROR, "hipGraphInstantiate failed");
    }
    ig.buf = x;
    return LNT_OK;
}

// k squarings (and -2 for LL) of the residue in buffer x; large path: three launches each
// (on the context's own residue and on the Gerbicz check's copy, whole groups of
// GRAPH_ITERS run as one graph launch).
static int square_buf(lnt_ctx *c, u64 *x, bool &pending, uint32_t k) {
    const Plan &pl = c->pl;
    const bool ll = !c->prp;
    const LntParams P = params_for(c, x, c->y, nullptr);
    if (pl.small) {
        // bounded launches keep each kernel short
        const uint32_t chunk = pl.n >= 2048 ? 256 : 1024;
        while (k) {
            uint32_t s = k < chunk ? k : chunk;
            if (!lnt_launch_small(pl, P, s, ll && pending, ll, c->stream))
                return fail(c, LNT_UNSUPPORTED, "no small-path kernel for this length");
            pending = true;
            k -= s;
        }
    } else {
        // the first iteration owes no "-2" when the residue was just set (or multiplied)
        if (k && ll && !pending) {
            if (!launch_iteration(pl, P, false, c->stream))
                return fail(c, LNT_UNSUPPORTED, "no kernel for this row/column split");
            pending = true;
            k--;
        }
        lnt_ctx::IterGraph *ig = x == c->x ? &c->graphs[0] : x == c->g.E.x ? &c->graphs[1] : nullptr;
        if (ig && k >= GRAPH_ITERS) {
            if (!ig->exec || ig->buf != x) {
                int rc = build_graph(c, *ig, x);
                if (rc != LNT_OK) return rc;
            }
            for (; k >= GRAPH_ITERS; k -= GRAPH_ITERS)
                HIPCHK(c, hipGraphLaunch(ig->exec, c->stream), "hipGraphLaunch");
            pending = true;
        }
        for (uint32_t i = 0; i < k; i++) {
            if (!launch_iteration(pl, P, ll && pending, c->stream))
                return fail(c, LNT_UNSUPPORTED, "no kernel for this row/column split");
            pending = true;
        }
    }
    HIPCHK(c, hipGetLastError(), "kernel launch");
    return LNT_OK;
}

static int run_iterations(lnt_ctx *c, uint32_t k) {
    int rc = square_buf(c, c->x, c->pending, k);
    if (rc == LNT_OK) c->pristine = false;
    return rc;
}

// dst <- dst * src mod M_p (both buffers hold residues of this context's shape).
// The transform of src goes to f, dst is transformed, multiplied and inverted in y.
static int mul_buf(lnt_ctx *c, u64 *dst, bool &dst_pending, const u64 *src, bool src_pending) {
    const Plan &pl = c->pl;
    const bool ll = !c->prp;
    if (!c->f) HIPCHK(c, hipMalloc(&c->f, (size_t)pl.n * 8), "hipMalloc (product operand)");
    const LntParams Ps = params_for(c, const_cast<u64 *>(src), c->f, nullptr);
    const LntParams Pd = params_for(c, dst, c->y, c->f);
    bool ok;
    if (pl.small) {
        ok = lnt_launch_small(pl, Ps, 1, ll && src_pending, ll, c->stream, LNT_FORWARD) &&
             lnt_launch_small(pl, Pd, 1, ll && dst_pending, ll, c->stream, LNT_MULTIPLY);
    } else {
        ok = lnt_launch_col_fwd(pl, Ps, ll && src_pending, c->stream) &&
             lnt_launch_row(pl, Ps, c->stream, LNT_FORWARD) &&
             lnt_launch_col_fwd(pl, Pd, ll && dst_pending, c->stream) &&
             lnt_launch_row(pl, Pd, c->stream, LNT_MULTIPLY) && lnt_launch_col_inv(pl, Pd, c->stream);
    }
    if (!ok) return fail(c, LNT_UNSUPPORTED, "no kernel for this length");
    HIPCHK(c, hipGetLastError(), "kernel launch");
    dst_pending = false;  // a product owes no "-2"
    return LNT_OK;
}

static int copy_buf(lnt_ctx *c, u64 *dst, const u64 *src) {
    HIPCHK(c, hipMemcpyAsync(dst, src, (size_t)c->pl.n * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpy (copy)");
    return LNT_OK;
}

// Canonical digits of the residue in a device buffer (waits for the stream).
static int canonical_of(lnt_ctx *c, const u64 *buf, bool pending, std::vector<uint64_t> &d) {
    const u32 n = c->pl.n;
    std::vector<uint64_t> xd(n);
    d.resize(n);
    HIPCHK(c, hipMemcpyAsync(xd.data(), buf, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy (D2H)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    to_canonical_digits(c->wid, xd.data(), !c->prp && pending, d.data());
    return LNT_OK;
}

// Plans the tuner picked, per (device, n); new contexts of that length start from them.
static std::mutex g_tuned_mu;
static std::map<std::pair<int, uint32_t>, PlanOpts> g_tuned;

// LNT_PLAN="lc=12,K=8,row3=1" overrides the plan of new contexts (experiments, tests).
static bool env_plan(PlanOpts &o) {
    const char *e = std::getenv("LNT_PLAN");
    if (!e || !*e) return false;
    std::string s = e;
    size_t pos = 0;
    while (pos < s.size()) {
        size_t q = s.find(',', pos);
        std::string kv = s.substr(pos, q == std::string::npos ? std::string::npos : q - pos);
        size_t eq = kv.find('=');
        if (eq != std::string::npos) {
            std::string k = kv.substr(0, eq);
            int v = std::atoi(kv.c_str() + eq + 1);
            if (k == "lc") o.lc = (uint32_t)v;
            if (k == "K") o.K = (uint32_t)v;
            if (k == "row3") o.row3 = v;
            if (k == "col3") o.col3 = v;
        }
        if (q == std::string::npos) break;
        pos = q + 1;
    }
    return true;
}

static int make_context(int device, uint32_t p, uint64_t n, bool prp, const PlanOpts *opts, lnt_ctx **out) {
    lnt_ctx *c = new lnt_ctx;
    c->device = device;
    c->p = p;
    c->prp = prp;
    PlanOpts o;
    if (opts) o = *opts;
    else if (!env_plan(o)) {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        auto it = g_tuned.find({device, (uint32_t)n});
        if (it != g_tuned.end()) o = it->second;
    }
    std::string why = make_plan(p, n, c->pl, o);
    if (!why.empty() && !opts) why = make_plan(p, n, c->pl, PlanOpts());  // stale tuned entry
    if (!why.empty()) {
        delete c;
        return fail(nullptr, LNT_UNSUPPORTED, why);
    }
    auto bail = [&](int code) {
        std::string m = c->err;
        free_buffers(c);
        if (c->ev_mark) (void)hipEventDestroy(c->ev_mark);
        if (c->ev_gz) (void)hipEventDestroy(c->ev_gz);
        if (c->stream) (void)hipStreamDestroy(c->stream);
        delete c;
        return fail(nullptr, code, m);
    };
    hipError_t e = hipSetDevice(device);
    if (e != hipSuccess) {
        c->err = std::string("hipSetDevice failed (") + hipGetErrorString(e) + ")";
        return bail(LNT_DEVICE_ERROR);
    }
    if (hipStreamCreateWithFlags(&c->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&c->ev_mark) != hipSuccess || hipEventCreateWithFlags(&c->ev_gz, hipEventDisableTiming) != hipSuccess) {
        c->err = "hipStreamCreate / hipEventCreate failed";
        return bail(LNT_DEVICE_ERROR);
    }
    widths(p, (uint32_t)n, c->wid);
    int rc = build_tables(c);
    if (rc != LNT_OK) return bail(rc);
    std::vector<uint64_t> d(n, 0);
    d[0] = prp ? 3 : 4;  // lucas_prp.cu:129
    rc = upload_digits(c, d.data());
    if (rc != LNT_OK) return bail(rc);
    if (hipEventRecord(c->ev_mark, c->stream) != hipSuccess) {
        c->err = "hipEventRecord failed";
        return bail(LNT_DEVICE_ERROR);
    }
    *out = c;
    return LNT_OK;
}

static int read_digits(lnt_ctx *c, uint64_t *d) {
    const u32 n = c->pl.n;
    if (c->pristine) {
        std::memcpy(d, c->set_digits.data(), (size_t)n * 8);
        return LNT_OK;
    }
    std::vector<uint64_t> xd(n);
    HIPCHK(c, hipMemcpyAsync(xd.data(), c->x, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy (x)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    bool zero = to_canonical_digits(c->wid, xd.data(), !c->prp && c->pending, d);
    if (zero && c->prp && c->zero_ones)
        for (u32 i = 0; i < n; i++) d[i] = (1ull << c->wid[i]) - 1;
    return LNT_OK;
}

// ------------------------------------------------------------------ C ABI

extern "C" {

int lnt_create(int device, uint32_t p, const uint32_t *factors, int nfactors, int mode, lnt_ctx **out) {
    if (!out) return fail(nullptr, LNT_INVALID_ARGUMENT, "out is NULL");
    *out = nullptr;
    if (mode != LNT_MODE_LL && mode != LNT_MODE_PRP) return fail(nullptr, LNT_INVALID_ARGUMENT, "bad mode");
    uint64_t n = 1;
    if (nfactors <= 0 || !factors) n = 1ull << secured_log2_length(p);
    else
        for (int i = 0; i < nfactors; i++) {
            if (factors[i] == 0) return fail(nullptr, LNT_INVALID_ARGUMENT, "zero NTT factor");
            n *= factors[i];
            if (n > (1ull << 32)) return fail(nullptr, LNT_UNSUPPORTED, "NTT length too large");
        }
    return make_context(device, p, n, mode == LNT_MODE_PRP, nullptr, out);
}

int lnt_set_digits(lnt_ctx *c, const uint64_t *d, size_t n) {
    if (!c || !d) return fail(c, LNT_INVALID_ARGUMENT, "NULL argument");
    if (n != c->pl.n) return fail(c, LNT_INVALID_ARGUMENT, "digit count does not match the NTT length");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    return upload_digits(c, d);
}

int lnt_iterate(lnt_ctx *c, uint32_t k) {
    if (!c) return fail(c, LNT_INVALID_ARGUMENT, "NULL context");
    if (k == 0) return LNT_OK;
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    return run_iterations(c, k);
}

int lnt_get_digits(lnt_ctx *c, uint64_t *d, size_t n) {
    if (!c || !d) return fail(c, LNT_INVALID_ARGUMENT, "NULL argument");
    if (n != c->pl.n) return fail(c, LNT_INVALID_ARGUMENT, "digit count does not match the NTT length");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    return read_digits(c, d);
}

int lnt_res64(lnt_ctx *c, uint64_t *out) {
    if (!c || !out) return fail(c, LNT_INVALID_ARGUMENT, "NULL argument");
    std::vector<uint64_t> d(c->pl.n);
    int rc = lnt_get_digits(c, d.data(), d.size());
    if (rc) return rc;
    bool ones = true;
    for (u32 i = 0; i < c->pl.n && ones; i++)
        if (d[i] != (1ull << c->wid[i]) - 1) ones = false;
    *out = ones ? 0 : res64(c->wid, d.data());
    return LNT_OK;
}

int lnt_sync(lnt_ctx *c, float *elapsed_ms) {
    if (!c) return fail(c, LNT_INVALID_ARGUMENT, "NULL context");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    hipEvent_t now;
    HIPCHK(c, hipEventCreate(&now), "hipEventCreate");
    HIPCHK(c, hipEventRecord(now, c->stream), "hipEventRecord");
    HIPCHK(c, hipEventSynchronize(now), "hipEventSynchronize");
    float ms = 0;
    HIPCHK(c, hipEventElapsedTime(&ms, c->ev_mark, now), "hipEventElapsedTime");
    (void)hipEventDestroy(c->ev_mark);
    c->ev_mark = now;
    if (elapsed_ms) *elapsed_ms = ms;
    return LNT_OK;
}

void lnt_destroy(lnt_ctx *c) {
    if (!c) return;
    if (c->batch) (void)lnt_batch_remove(c->batch, c);
    (void)hipSetDevice(c->device);
    if (c->stream) (void)hipStreamSynchronize(c->stream);
    drop_graph(c);
    free_buffers(c);
    if (c->ev_mark) (void)hipEventDestroy(c->ev_mark);
    if (c->ev_gz) (void)hipEventDestroy(c->ev_gz);
    if (c->stream) (void)hipStreamDestroy(c->stream);
    delete c;
}

const char *lnt_last_error(const lnt_ctx *c) {
    if (c) return c->err.c_str();
    std::lock_guard<std::mutex> g(g_err_mu);
    return g_err.c_str();
}

uint32_t lnt_secured_log2_length(uint32_t p) { return secured_log2_length(p); }

// Tuned plans as text, one line per (device, length): "device n lc K row3 col3".
int lnt_tune_export(char *buf, size_t len) {
    if (!buf || !len) return fail(nullptr, LNT_INVALID_ARGUMENT, "bad argument");
    std::string out;
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        for (const auto &kv : g_tuned) {
            char line[96];
            std::snprintf(line, sizeof line, "%d %u %u %u %d %d\n", kv.first.first, kv.first.second, kv.second.lc,
                          kv.second.K, kv.second.row3, kv.second.col3);
            out += line;
        }
    }
    std::snprintf(buf, len, "%s", out.c_str());
    return out.size() < len ? LNT_OK : fail(nullptr, LNT_INVALID_ARGUMENT, "lnt_tune_export: buffer too small");
}

int lnt_tune_import(const char *text) {
    if (!text) return fail(nullptr, LNT_INVALID_ARGUMENT, "bad argument");
    std::vector<std::pair<std::pair<int, uint32_t>, PlanOpts>> rows;
    const char *s = text;
    while (*s) {
        const char *e = std::strchr(s, '\n');
        std::string line(s, e ? (size_t)(e - s) : std::strlen(s));
        s = e ? e + 1 : s + line.size();
        if (line.find_first_not_of(" \t\r") == std::string::npos || line[0] == '#') continue;
        int dev = 0, r3 = -1, c3 = -1;
        unsigned n = 0, lc = 0, K = 0;
        if (std::sscanf(line.c_str(), "%d %u %u %u %d %d", &dev, &n, &lc, &K, &r3, &c3) != 6 || !n || (n & (n - 1)))
            return fail(nullptr, LNT_INVALID_ARGUMENT, "lnt_tune_import: bad line '" + line + "'");
        PlanOpts o;
        o.lc = lc;
        o.K = K;
        o.row3 = r3;
        o.col3 = c3;
        rows.push_back({{dev, n}, o});
    }
    std::lock_guard<std::mutex> lk(g_tuned_mu);
    for (auto &r : rows) g_tuned[r.first] = r.second;  // a stale entry falls back to the default plan
    return LNT_OK;
}

}  // extern "C"

// Field-op self-test: the device forms of gl64.hpp (carry chains) against the host forms.
__global__ void lnt_k_field_ops(const u64 *a, const u64 *b, u32 count, u64 *out) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= count) return;
    const u64 x = a[i], y = b[i];
    u64 *o = out + (u64)i * 8;
    o[0] = gl_add(x, y);
    o[1] = gl_sub(x, y);
    o[2] = gl_mul(x, y);
    o[3] = gl_sqr(x);
    o[4] = gl_mul_pow2(x, 12);
    o[5] = gl_mul_pow2(x, 48);
    o[6] = gl_sub_pow2(x, y, 84);
    o[7] = gl_sub_pow2(x, y, 132);
}

extern "C" {

int lnt_selftest(int device, uint32_t count, uint64_t seed, uint64_t *mismatches) {
    if (!mismatches || count == 0) return fail(nullptr, LNT_INVALID_ARGUMENT, "bad argument");
    if (hipSetDevice(device) != hipSuccess) return fail(nullptr, LNT_DEVICE_ERROR, "hipSetDevice failed");
    // operands: canonical values with the edge cases first (0, 1, P-1, 2^32, EPS, ...)
    std::vector<u64> a(count), b(count), out((size_t)count * 8);
    const u64 edge[] = {0, 1, 2, GL_P - 1, GL_P - 2, 1ull << 32, GL_EPS, (1ull << 63), GL_P >> 1, (GL_P >> 1) + 1};
    const u32 ne = sizeof edge / sizeof edge[0];
    u64 st = seed;
    auto next = [&]() {  // splitmix64
        st += 0x9E3779B97F4A7C15ull;
        u64 z = st;
        z = (z ^ (z >> 30)) * 0xBF58476D1CE4E5B9ull;
        z = (z ^ (z >> 27)) * 0x94D049BB133111EBull;
        return z ^ (z >> 31);
    };
    for (u32 i = 0; i < count; i++) {
        a[i] = i < ne * ne ? edge[i / ne] : next() % GL_P;
        b[i] = i < ne * ne ? edge[i % ne] : next() % GL_P;
    }
    u64 *da = nullptr, *db = nullptr, *dout = nullptr;
    bool ok = hipMalloc(&da, (size_t)count * 8) == hipSuccess && hipMalloc(&db, (size_t)count * 8) == hipSuccess &&
              hipMalloc(&dout, (size_t)count * 64) == hipSuccess &&
              hipMemcpy(da, a.data(), (size_t)count * 8, hipMemcpyHostToDevice) == hipSuccess &&
              hipMemcpy(db, b.data(), (size_t)count * 8, hipMemcpyHostToDevice) == hipSuccess;
    if (ok) {
        hipLaunchKernelGGL(lnt_k_field_ops, dim3((count + 255) / 256), dim3(256), 0, 0, da, db, count, dout);
        ok = hipGetLastError() == hipSuccess &&
             hipMemcpy(out.data(), dout, (size_t)count * 64, hipMemcpyDeviceToHost) == hipSuccess;
    }
    if (da) (void)hipFree(da);
    if (db) (void)hipFree(db);
    if (dout) (void)hipFree(dout);
    if (!ok) return fail(nullptr, LNT_DEVICE_ERROR, "lnt_selftest: device error");
    u64 bad = 0;
    for (u32 i = 0; i < count; i++) {
        const u64 x = a[i], y = b[i], *o = &out[(size_t)i * 8];
        const u64 want[8] = {gl_add(x, y),          gl_sub(x, y),          gl_mul(x, y),
                             gl_sqr(x),             gl_mul_pow2(x, 12),    gl_mul_pow2(x, 48),
                             gl_sub_pow2(x, y, 84), gl_sub_pow2(x, y, 132)};
        for (int k = 0; k < 8; k++) bad += o[k] != want[k];
    }
    *mismatches = bad;
    return LNT_OK;
}

uint32_t lnt_length(const lnt_ctx *c) { return c ? c->pl.n : 0; }

static std::string describe(const Plan &pl) {
    char b[320];
    if (pl.small && pl.row3) {
        std::snprintf(b, sizeof b, "LNT_1WG;lnt_k_small3<%u>", pl.row3);
    } else if (pl.small) {
        std::snprintf(b, sizeof b, "LNT_1WG;lnt_k_small<%u,%u>", pl.row.B1, pl.row.B2);
    } else {
        char row[64];
        if (pl.row3m) std::snprintf(row, sizeof row, "lnt_k_row3m<%u,%u,%u>", pl.row3m, pl.rBB, pl.rBC);
        else if (pl.row3) std::snprintf(row, sizeof row, "lnt_k_row3<%u,%u>", pl.row3, pl.G);
        else std::snprintf(row, sizeof row, "lnt_k_row<%u,%u>", pl.row.B1, pl.row.B2);
        if (pl.col3)
            std::snprintf(b, sizeof b, "LNT_2PASS_%ux%u;lnt_k_col3_fwd<%u,%u,%u,%u>;%s;lnt_k_col3_inv<%u,%u,%u,%u>", pl.R,
                          pl.C, pl.col3, pl.cBB, pl.cBC, pl.K, row, pl.col3, pl.cBB, pl.cBC, pl.K);
        else
            std::snprintf(b, sizeof b, "LNT_2PASS_%ux%u;lnt_k_col_fwd<%u,%u,%u>;%s;lnt_k_col_inv<%u,%u,%u>", pl.R,
                          pl.C, pl.col.B1, pl.col.B2, pl.K, row, pl.col.B1, pl.col.B2, pl.K);
    }
    return b;
}

int lnt_describe(const lnt_ctx *c, char *buf, size_t len) {
    if (!c || !buf || !len) return LNT_INVALID_ARGUMENT;
    std::string s = describe(c->pl);
    std::snprintf(buf, len, "%s", s.c_str());
    return s.size() < len ? LNT_OK : LNT_INVALID_ARGUMENT;
}

int lnt_tune_results(const lnt_ctx *c, char *buf, size_t len) {
    if (!c || !buf || !len) return LNT_INVALID_ARGUMENT;
    std::snprintf(buf, len, "%s", c->tune_report.c_str());
    return c->tune_report.size() < len ? LNT_OK : LNT_INVALID_ARGUMENT;
}

int lnt_profile(lnt_ctx *c, uint32_t iters, char *names, size_t names_len, float *ms, int cap) {
    if (!c || !iters) return fail(c, LNT_INVALID_ARGUMENT, "bad argument");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    const Plan &pl = c->pl;
    const bool ll = !c->prp;
    const int nk = pl.small ? 1 : 3;
    // Each kernel runs `iters` times back to back as one captured graph, the way the
    // iteration graph runs it (no host launch gap between dispatches and no event packet
    // between kernels), timed by two events around the graph launch.  The residue is saved
    // first and restored after: repeated launches of one kernel are timing only.
    u64 *save = nullptr;
    HIPCHK(c, hipMalloc(&save, (size_t)pl.n * 8), "hipMalloc (profile)");
    HIPCHK(c, hipMemcpyAsync(save, c->x, (size_t)pl.n * 8, hipMemcpyDeviceToDevice, c->stream), "hipMemcpy");
    const LntParams P = params_for(c, c->x, c->y, nullptr);
    const bool sub2 = ll && c->pending;
    hipEvent_t e0 = nullptr, e1 = nullptr;
    std::vector<double> tot(nk, 0.0);
    int rc = LNT_OK;
    std::string why;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) {
        rc = LNT_DEVICE_ERROR;
        why = "hipEventCreate";
    }
    for (int k = 0; k < nk && rc == LNT_OK; k++) {
        hipGraph_t g = nullptr;
        hipGraphExec_t ge = nullptr;
        bool ok = hipStreamBeginCapture(c->stream, hipStreamCaptureModeThreadLocal) == hipSuccess;
        for (uint32_t i = 0; i < iters && ok; i++) {
            if (pl.small) ok = lnt_launch_small(pl, P, 1, sub2, ll, c->stream);
            else if (k == 0) ok = lnt_launch_col_fwd(pl, P, sub2, c->stream);
            else if (k == 1) ok = lnt_launch_row(pl, P, c->stream);
            else ok = lnt_launch_col_inv(pl, P, c->stream);
        }
        const hipError_t ec = hipStreamEndCapture(c->stream, &g);
        float t = 0;
        if (!ok || ec != hipSuccess || hipGraphInstantiate(&ge, g, nullptr, nullptr, 0) != hipSuccess) {
            rc = ok ? LNT_DEVICE_ERROR : LNT_UNSUPPORTED;
            why = "profile graph capture";
        } else if (hipGraphLaunch(ge, c->stream) != hipSuccess ||  // warm-up
                   hipEventRecord(e0, c->stream) != hipSuccess || hipGraphLaunch(ge, c->stream) != hipSuccess ||
                   hipEventRecord(e1, c->stream) != hipSuccess || hipEventSynchronize(e1) != hipSuccess ||
                   hipEventElapsedTime(&t, e0, e1) != hipSuccess) {
            rc = LNT_DEVICE_ERROR;
            why = "profile graph launch";
        }
        tot[k] = t;
        if (ge) (void)hipGraphExecDestroy(ge);
        if (g) (void)hipGraphDestroy(g);
    }
    if (e0) (void)hipEventDestroy(e0);
    if (e1) (void)hipEventDestroy(e1);
    const hipError_t ce = hipMemcpyAsync(c->x, save, (size_t)pl.n * 8, hipMemcpyDeviceToDevice, c->stream);
    const hipError_t se = hipStreamSynchronize(c->stream);
    (void)hipFree(save);
    if (rc != LNT_OK) return fail(c, rc, why);
    HIPCHK(c, ce, "hipMemcpy");
    HIPCHK(c, se, "hipStreamSynchronize");
    std::string d = describe(pl);
    std::vector<std::string> parts;
    size_t pos = 0;
    while (true) {
        size_t q = d.find(';', pos);
        parts.push_back(d.substr(pos, q == std::string::npos ? std::string::npos : q - pos));
        if (q == std::string::npos) break;
        pos = q + 1;
    }
    std::string nm;
    int cnt = 0;
    for (int k = 0; k < nk && k < cap; k++) {
        if (ms) ms[k] = (float)(tot[k] / iters);
        if (k) nm += ";";
        nm += parts[k + 1];
        cnt++;
    }
    if (names && names_len) std::snprintf(names, names_len, "%s", nm.c_str());
    return cnt;
}

int lnt_tune(lnt_ctx *c, int loops) {
    if (!c || loops <= 0) return fail(c, LNT_INVALID_ARGUMENT, "bad argument");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    if (c->pl.small) {  // one configuration only: time it
        float ms = 0;
        if (lnt_sync(c, nullptr) != LNT_OK) return LNT_DEVICE_ERROR;
        std::vector<uint64_t> save(c->pl.n);
        HIPCHK(c, hipMemcpyAsync(save.data(), c->x, (size_t)c->pl.n * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
        HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        const bool pend = c->pending, prist = c->pristine;
        int rc = run_iterations(c, (uint32_t)loops);
        if (rc == LNT_OK) rc = lnt_sync(c, &ms);
        HIPCHK(c, hipMemcpyAsync(c->x, save.data(), (size_t)c->pl.n * 8, hipMemcpyHostToDevice, c->stream), "hipMemcpy");
        HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
        c->pending = pend;
        c->pristine = prist;
        if (rc != LNT_OK) return rc;
        std::string nm = describe(c->pl);
        std::replace(nm.begin(), nm.end(), ';', '/');
        char line[512];
        std::snprintf(line, sizeof line, "%s = %g\n", nm.c_str(), ms);
        c->tune_report = line;
        return LNT_OK;
    }
    const u32 n = c->pl.n;
    // current residue, exactly as the next iteration will see it
    std::vector<uint64_t> start(n);
    HIPCHK(c, hipMemcpyAsync(start.data(), c->x, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const bool start_pending = c->pending;
    struct Cand {
        PlanOpts o;
        float ms;
        std::vector<uint64_t> out;
    };
    // candidates: every row/column split, column tile and row variant this build has
    std::vector<PlanOpts> opts;
    for (uint32_t lc = 7; lc <= 14; lc++)
        for (uint32_t K : {16u, 8u, 4u})
            for (int r3 : {0, 1, 2})
                for (int c3 : {0, 1}) {
                    PlanOpts o;
                    o.lc = lc;
                    o.K = K;
                    o.row3 = r3;
                    o.col3 = c3;
                    Plan tmp;
                    if (make_plan(c->p, n, tmp, o).empty()) opts.push_back(o);
                }
    std::vector<Cand> cands;
    // warm-up long enough to capture the candidate's iteration graph, so the timed loops
    // replay it as the plain path does
    const uint32_t warm = std::max<uint32_t>(GRAPH_ITERS, (uint32_t)loops / 10);
    for (const PlanOpts &o : opts) {
        lnt_ctx *t = nullptr;
        if (make_context(c->device, c->p, n, c->prp, &o, &t) != LNT_OK) continue;
        HIPCHK(c, hipMemcpyAsync(t->x, start.data(), (size_t)n * 8, hipMemcpyHostToDevice, t->stream), "hipMemcpy");
        t->pending = start_pending;
        t->pristine = false;
        float ms = 0;
        // warm-up on a copy of the state, then time `loops` iterations from the start state
        if (run_iterations(t, warm) != LNT_OK || lnt_sync(t, nullptr) != LNT_OK) {
            lnt_destroy(t);
            continue;
        }
        HIPCHK(c, hipMemcpyAsync(t->x, start.data(), (size_t)n * 8, hipMemcpyHostToDevice, t->stream), "hipMemcpy");
        t->pending = start_pending;
        lnt_sync(t, nullptr);
        if (run_iterations(t, (uint32_t)loops) != LNT_OK || lnt_sync(t, &ms) != LNT_OK) {
            lnt_destroy(t);
            continue;
        }
        Cand cd{o, ms, std::vector<uint64_t>(n)};
        read_digits(t, cd.out.data());
        lnt_destroy(t);
        cands.push_back(std::move(cd));
    }
    if (cands.empty()) return fail(c, LNT_DEVICE_ERROR, "tuner: no candidate ran");
    c->tune_report.clear();
    for (const auto &cd : cands) {
        Plan cp;
        make_plan(c->p, n, cp, cd.o);
        std::string nm = describe(cp);
        std::replace(nm.begin(), nm.end(), ';', '/');
        char line[512];
        std::snprintf(line, sizeof line, "%s = %g\n", nm.c_str(), cd.ms);
        c->tune_report += line;
    }
    // every candidate must agree with the others before its time counts
    for (auto &cd : cands)
        if (cd.out != cands[0].out) return fail(c, LNT_DEVICE_ERROR, "tuner: candidates disagree on the residue");
    // the fastest candidate; the context's current plan stays unless another is faster by
    // more than 2% (timings of near-equal plans differ by noise, and a switch rebuilds tables)
    const Cand *best = &cands[0];
    for (auto &cd : cands)
        if (cd.ms < best->ms) best = &cd;
    for (auto &cd : cands) {
        Plan cp;
        make_plan(c->p, n, cp, cd.o);
        if (same_kernels(cp, c->pl) && best->ms > 0.98f * cd.ms) best = &cd;
    }
    {
        std::lock_guard<std::mutex> lk(g_tuned_mu);
        g_tuned[{c->device, n}] = best->o;
    }
    Plan bp;
    make_plan(c->p, n, bp, best->o);
    if (same_kernels(bp, c->pl)) return LNT_OK;
    // rebuild this context on the winning plan, keeping the residue
    lnt_ctx *nc = nullptr;
    int rc = make_context(c->device, c->p, n, c->prp, &best->o, &nc);
    if (rc != LNT_OK) return fail(c, rc, lnt_last_error(nullptr));
    HIPCHK(c, hipMemcpyAsync(nc->x, start.data(), (size_t)n * 8, hipMemcpyHostToDevice, nc->stream), "hipMemcpy");
    HIPCHK(c, hipStreamSynchronize(nc->stream), "hipStreamSynchronize");
    std::swap(c->pl, nc->pl);
    std::swap(c->x, nc->x);
    std::swap(c->y, nc->y);
    std::swap(c->w, nc->w);
    std::swap(c->uw, nc->uw);
    std::swap(c->tw4, nc->tw4);
    std::swap(c->itw4, nc->itw4);
    std::swap(c->twr, nc->twr);
    std::swap(c->itwr, nc->itwr);
    std::swap(c->twr2, nc->twr2);
    std::swap(c->itwr2, nc->itwr2);
    std::swap(c->twc, nc->twc);
    std::swap(c->itwc, nc->itwc);
    std::swap(c->twc2, nc->twc2);
    std::swap(c->itwc2, nc->itwc2);
    std::swap(c->P, nc->P);
    drop_graph(c);  // captured over the old buffers and kernels
    lnt_destroy(nc);
    return LNT_OK;
}

// ------------------------------------------------------------------ products, copies, faults

static bool same_shape(const lnt_ctx *a, const lnt_ctx *b) {
    // the transform-domain slot order depends on the split and on the row/column variants
    return a->device == b->device && a->p == b->p && a->prp == b->prp && same_slots(a->pl, b->pl);
}

int lnt_mul(lnt_ctx *dst, lnt_ctx *src) {
    if (!dst || !src) return fail(dst, LNT_INVALID_ARGUMENT, "NULL context");
    if (!same_shape(dst, src)) return fail(dst, LNT_INVALID_ARGUMENT, "lnt_mul: contexts differ in p, n, mode or split");
    HIPCHK(dst, hipSetDevice(dst->device), "hipSetDevice");
    if (src != dst) HIPCHK(dst, hipStreamSynchronize(src->stream), "hipStreamSynchronize");
    int rc = mul_buf(dst, dst->x, dst->pending, src->x, src->pending);
    if (rc == LNT_OK) dst->pristine = false;
    return rc;
}

int lnt_copy(lnt_ctx *dst, lnt_ctx *src) {
    if (!dst || !src) return fail(dst, LNT_INVALID_ARGUMENT, "NULL context");
    if (!same_shape(dst, src)) return fail(dst, LNT_INVALID_ARGUMENT, "lnt_copy: contexts differ in p, n, mode or split");
    HIPCHK(dst, hipSetDevice(dst->device), "hipSetDevice");
    if (src != dst) HIPCHK(dst, hipStreamSynchronize(src->stream), "hipStreamSynchronize");
    int rc = copy_buf(dst, dst->x, src->x);
    if (rc) return rc;
    dst->pending = src->pending;
    dst->pristine = src->pristine;
    dst->zero_ones = src->zero_ones;
    dst->set_digits = src->set_digits;
    return LNT_OK;
}

int lnt_inject_error(lnt_ctx *c, uint32_t word, uint64_t xor_mask) {
    if (!c) return fail(c, LNT_INVALID_ARGUMENT, "NULL context");
    if (word >= c->pl.n) return fail(c, LNT_INVALID_ARGUMENT, "lnt_inject_error: word index out of range");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    uint64_t v = 0;
    HIPCHK(c, hipMemcpyAsync(&v, c->x + word, 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy (D2H)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    // flip bits of the word's signed value, keeping it a valid field encoding
    int64_t sv = gl_to_i64(v);
    sv ^= (int64_t)(xor_mask & 0x3FFFFFFFFFFFull);
    v = gl_from_i64(sv);
    HIPCHK(c, hipMemcpyAsync(c->x + word, &v, 8, hipMemcpyHostToDevice, c->stream), "hipMemcpy (H2D)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    c->pristine = false;
    return LNT_OK;
}

// ------------------------------------------------------------------ Gerbicz-checked PRP
// Residues u_k = 3^(2^k) (k squarings since lnt_gerbicz_begin, u_0 = the residue then).
// Every B iterations D <- D * u (D = prod u_(iB), i = 0..m); every L blocks the check
// D_m == u_0 * D_(m-1)^(2^B) (B extra squarings of a copy of D_(m-1)).  A pass saves
// (u, D, k) as the verified checkpoint on the device; a failure restores it and redoes the
// iterations since.

static int gz_alloc(lnt_ctx *c, lnt_ctx::Slot &s) {
    if (!s.x) HIPCHK(c, hipMalloc(&s.x, (size_t)c->pl.n * 8), "hipMalloc (Gerbicz slot)");
    return LNT_OK;
}

static int gz_check(lnt_ctx *c, bool &passed) {
    auto &g = c->g;
    const u32 n = c->pl.n;
    int rc;
    // D is final for this check: read it back first and put it in canonical digits on a
    // host thread while the device runs the B squarings of E (the future's destructor
    // waits for the thread on every return path, before xd and dd go away)
    std::vector<uint64_t> xd(n), dd(n), de;
    HIPCHK(c, hipMemcpyAsync(xd.data(), g.D.x, (size_t)n * 8, hipMemcpyDeviceToHost, c->stream), "hipMemcpy (D2H)");
    HIPCHK(c, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    const bool dsub2 = !c->prp && g.D.pending;
    std::future<bool> dconv;
    try {
        dconv = std::async(std::launch::async, [c, &xd, &dd, dsub2] {
            return to_canonical_digits(c->wid, xd.data(), dsub2, dd.data());
        });
    } catch (const std::system_error &) {  // no thread available: convert here
        to_canonical_digits(c->wid, xd.data(), dsub2, dd.data());
    }
    if ((rc = copy_buf(c, g.E.x, g.Dprev.x))) return rc;
    g.E.pending = g.Dprev.pending;
    if ((rc = square_buf(c, g.E.x, g.E.pending, g.B))) return rc;
    if ((rc = mul_buf(c, g.E.x, g.E.pending, g.U0.x, g.U0.pending))) return rc;
    if ((rc = canonical_of(c, g.E.x, g.E.pending, de))) return rc;
    if (dconv.valid()) dconv.wait();
    passed = de == dd;
    if (passed) g.ck_res64 = res64(c->wid, dd.data());
    return LNT_OK;
}

// The step after the squarings that end a block (g.it % B == 0): D_prev <- D, D <- D * u;
// every L blocks the check, then save the verified checkpoint or restore it (g.it falls
// back to the checkpoint's iteration).
static int gz_block_end(lnt_ctx *c) {
    auto &g = c->g;
    int rc;
    if ((rc = copy_buf(c, g.Dprev.x, g.D.x))) return rc;
    g.Dprev.pending = g.D.pending;
    if ((rc = mul_buf(c, g.D.x, g.D.pending, c->x, c->pending))) return rc;
    if ((g.it / g.B) % g.L) return LNT_OK;
    bool passed = false;
    if ((rc = gz_check(c, passed))) return rc;
    if (passed) {
        g.checks++;
        g.fails_here = 0;
        g.it_v = g.it;
        if ((rc = copy_buf(c, g.Uv.x, c->x))) return rc;
        g.Uv.pending = c->pending;
        if ((rc = copy_buf(c, g.Dv.x, g.D.x))) return rc;
        g.Dv.pending = g.D.pending;
    } else {
        g.rollbacks++;
        if (++g.fails_here > 3) return fail(c, LNT_DEVICE_ERROR, "Gerbicz check failed 4 times from the same checkpoint");
        if ((rc = copy_buf(c, c->x, g.Uv.x))) return rc;
        c->pending = g.Uv.pending;
        if ((rc = copy_buf(c, g.D.x, g.Dv.x))) return rc;
        g.D.pending = g.Dv.pending;
        g.it = g.it_v;
    }
    return LNT_OK;
}

int lnt_gerbicz_begin(lnt_ctx *c, uint32_t block, uint32_t blocks_per_check) {
    if (!c) return fail(c, LNT_INVALID_ARGUMENT, "NULL context");
    if (!c->prp) return fail(c, LNT_INVALID_ARGUMENT, "the Gerbicz check applies to PRP tests");
    if (block == 0 || blocks_per_check == 0) return fail(c, LNT_INVALID_ARGUMENT, "block sizes must be positive");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    auto &g = c->g;
    int rc;
    for (auto *s : {&g.U0, &g.D, &g.Dprev, &g.E, &g.Uv, &g.Dv})
        if ((rc = gz_alloc(c, *s))) return rc;
    g.on = true;
    g.B = block;
    g.L = blocks_per_check;
    g.it = g.it_v = g.checks = g.rollbacks = g.fails_here = 0;
    for (auto *s : {&g.U0, &g.D, &g.Dprev, &g.Uv, &g.Dv}) {
        if ((rc = copy_buf(c, s->x, c->x))) return rc;
        s->pending = c->pending;
    }
    std::vector<uint64_t> d;
    if ((rc = canonical_of(c, g.D.x, g.D.pending, d))) return rc;
    g.ck_res64 = res64(c->wid, d.data());
    return LNT_OK;
}

int lnt_gerbicz_iterate(lnt_ctx *c, uint32_t k) {
    if (!c) return fail(c, LNT_INVALID_ARGUMENT, "NULL context");
    auto &g = c->g;
    if (!g.on) return fail(c, LNT_INVALID_ARGUMENT, "lnt_gerbicz_begin was not called");
    HIPCHK(c, hipSetDevice(c->device), "hipSetDevice");
    const uint64_t target = g.it + k;
    int rc;
    while (g.it < target) {
        uint64_t step = std::min<uint64_t>(target - g.it, g.B - g.it % g.B);
        if ((rc = square_buf(c, c->x, c->pending, (uint32_t)step))) return rc;
        c->pristine = false;
        g.it += step;
        if (g.it % g.B == 0 && (rc = gz_block_end(c))) return rc;
    }
    return LNT_OK;
}

int lnt_gerbicz_stats(lnt_ctx *c, uint64_t *out, int cap) {
    if (!c || !out || cap < 5) return fail(c, LNT_INVALID_ARGUMENT, "bad argument");
    const auto &g = c->g;
    out[0] = g.it;
    out[1] = g.it_v;
    out[2] = g.checks;
    out[3] = g.rollbacks;
    out[4] = g.ck_res64;
    return LNT_OK;
}

}  // extern "C"

// ------------------------------------------------------------------ batches
// Many independent tests advanced together: members with the same plan shape form a
// group that runs as one batched launch per kernel (per-test parameters in a device
// array); one iteration of every group is captured as a hipGraph whose groups are
// parallel branches, and each lnt_batch_iterate step replays it.

struct lnt_batch {
    int device = 0;
    hipStream_t stream = nullptr;
    hipEvent_t ev_mark = nullptr;
    std::vector<lnt_ctx *> members;
    // Members of one kernel shape share one batched launch per kernel over `cap` parameter
    // slots; an empty slot (n = 0) launches workgroups that return at once, so members can
    // join or leave a group with free slots by rewriting its parameters, without a new graph.
    struct Group {
        Plan pl;
        bool ll = false;
        u32 cap = 0;
        std::vector<lnt_ctx *> ctxs;  // cap slots, nullptr = empty
        std::vector<LntParams> h;     // cap slots, n = 0 = empty
        LntParams *d = nullptr;
        hipStream_t side = nullptr;
        hipEvent_t join = nullptr;
    };
    std::vector<Group> groups;
    hipEvent_t fork = nullptr;
    hipEvent_t done = nullptr;  // end of the last batch_iterate: members' streams wait on it
    hipGraph_t graph = nullptr;
    hipGraphExec_t exec = nullptr;
    bool dirty = true;      // the groups or their capacities changed: capture a new graph
    bool upload = false;    // parameters changed: copy them to the device before the next step
    uint64_t captures = 0;  // graphs captured so far
    std::string err;
};

static int bfail(lnt_batch *b, int code, const std::string &m) {
    if (b) b->err = m;
    return fail(nullptr, code, m);
}

#define BHIPCHK(b, call, what)                                                                           \
    do {                                                                                                  \
        hipError_t e_ = (call);                                                                           \
        if (e_ != hipSuccess)                                                                             \
            return bfail((b), e_ == hipErrorOutOfMemory ? LNT_OUT_OF_MEMORY : LNT_DEVICE_ERROR,           \
                         std::string(what) + " (" + hipGetErrorString(e_) + ")");                         \
    } while (0)

// Members share a batched launch only when every compile-time shape of their kernels and
// every table layout agree (the carry level count is per member).
static bool same_group(const Plan &a, const Plan &b) { return same_kernels(a, b); }

static void batch_release(lnt_batch *b) {
    if (b->exec) (void)hipGraphExecDestroy(b->exec);
    if (b->graph) (void)hipGraphDestroy(b->graph);
    b->exec = nullptr;
    b->graph = nullptr;
    for (auto &g : b->groups) {
        if (g.d) (void)hipFree(g.d);
        if (g.side) (void)hipStreamDestroy(g.side);
        if (g.join) (void)hipEventDestroy(g.join);
    }
    b->groups.clear();
}

// Launch one iteration of group g on stream s (all cap slots; empty ones return at once).
static bool batch_group_step(lnt_batch::Group &g, hipStream_t s) {
    if (g.pl.small) return lnt_launch_batch_small(g.pl, g.d, g.cap, 1, g.ll, s);
    return lnt_launch_batch_col_fwd(g.pl, g.d, g.cap, s) && lnt_launch_batch_row(g.pl, g.d, g.cap, s) &&
           lnt_launch_batch_col_inv(g.pl, g.d, g.cap, s);
}

// The parameter block of member c in a batch launch (its "-2" flag from its own state).
static LntParams member_params(const lnt_ctx *c) {
    LntParams P = c->P;
    P.sub2 = (!c->prp && c->pending) ? 1u : 0u;
    return P;
}

static lnt_batch::Group *group_of(lnt_batch *b, const lnt_ctx *c) {
    for (auto &g : b->groups)
        if (same_group(g.pl, c->pl) && g.ll == !c->prp) return &g;
    return nullptr;
}

// Groups for the current members, each with room to grow (at least 8 slots, a power of
// two, at least twice its members), their parameters, streams and the captured graph of
// one iteration (every group a parallel branch forked from and joined to the batch stream).
static int batch_build(lnt_batch *b) {
    batch_release(b);
    for (lnt_ctx *c : b->members) {
        lnt_batch::Group *grp = group_of(b, c);
        if (!grp) {
            b->groups.emplace_back();
            grp = &b->groups.back();
            grp->pl = c->pl;
            grp->ll = !c->prp;
        }
        grp->ctxs.push_back(c);
    }
    for (auto &g : b->groups) {
        u32 cap = 8;
        while (cap < 2 * g.ctxs.size()) cap *= 2;
        g.cap = cap;
        g.ctxs.resize(cap, nullptr);
        g.h.assign(cap, LntParams{});
        for (u32 i = 0; i < cap; i++)
            if (g.ctxs[i]) g.h[i] = member_params(g.ctxs[i]);
        BHIPCHK(b, hipMalloc(&g.d, g.h.size() * sizeof(LntParams)), "hipMalloc (batch parameters)");
        BHIPCHK(b, hipMemcpyAsync(g.d, g.h.data(), g.h.size() * sizeof(LntParams), hipMemcpyHostToDevice, b->stream),
                "hipMemcpy (batch parameters)");
        BHIPCHK(b, hipStreamCreateWithFlags(&g.side, hipStreamNonBlocking), "hipStreamCreate");
        BHIPCHK(b, hipEventCreateWithFlags(&g.join, hipEventDisableTiming), "hipEventCreate");
    }
    BHIPCHK(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    BHIPCHK(b, hipStreamBeginCapture(b->stream, hipStreamCaptureModeThreadLocal), "hipStreamBeginCapture");
    bool ok = true;
    if (hipEventRecord(b->fork, b->stream) != hipSuccess) ok = false;
    for (auto &g : b->groups) {
        if (!ok) break;
        ok = hipStreamWaitEvent(g.side, b->fork, 0) == hipSuccess && batch_group_step(g, g.side) &&
             hipEventRecord(g.join, g.side) == hipSuccess && hipStreamWaitEvent(b->stream, g.join, 0) == hipSuccess;
    }
    hipError_t ec = hipStreamEndCapture(b->stream, &b->graph);
    if (!ok || ec != hipSuccess) return bfail(b, LNT_DEVICE_ERROR, "batch graph capture failed");
    BHIPCHK(b, hipGraphInstantiate(&b->exec, b->graph, nullptr, nullptr, 0), "hipGraphInstantiate");
    b->dirty = false;
    b->upload = false;
    b->captures++;
    return LNT_OK;
}

// Copy the groups' parameter slots to the device (after members joined, left or changed).
static int batch_upload(lnt_batch *b) {
    for (auto &g : b->groups)
        BHIPCHK(b, hipMemcpyAsync(g.d, g.h.data(), g.h.size() * sizeof(LntParams), hipMemcpyHostToDevice, b->stream),
                "hipMemcpy (batch parameters)");
    b->upload = false;
    return LNT_OK;
}

// Each member's parameters from its own state: its "-2" flag (set_digits, a product or a
// copy on a member since the last batch step clears it) and its buffers and tables (the
// tuner may have re-planned it: a new kernel shape needs a new graph).
static void batch_refresh(lnt_batch *b) {
    for (auto &g : b->groups)
        for (u32 i = 0; i < g.cap; i++) {
            lnt_ctx *c = g.ctxs[i];
            if (!c) continue;
            if (!same_group(g.pl, c->pl)) {
                b->dirty = true;
                continue;
            }
            const LntParams P = member_params(c);
            if (std::memcmp(&P, &g.h[i], sizeof P) != 0) {
                g.h[i] = P;
                b->upload = true;
            }
        }
}

// After an iteration every LL member owes its "-2": refresh the flags where they changed.
static int batch_mark_pending(lnt_batch *b) {
    for (auto &g : b->groups) {
        bool changed = false;
        for (u32 i = 0; i < g.cap; i++) {
            lnt_ctx *c = g.ctxs[i];
            if (!c) continue;
            c->pending = true;
            c->pristine = false;
            u32 want = c->prp ? 0u : 1u;
            if (g.h[i].sub2 != want) {
                g.h[i].sub2 = want;
                changed = true;
            }
        }
        if (changed)
            BHIPCHK(b, hipMemcpyAsync(g.d, g.h.data(), g.h.size() * sizeof(LntParams), hipMemcpyHostToDevice, b->stream),
                    "hipMemcpy (batch parameters)");
    }
    return LNT_OK;
}

extern "C" {

int lnt_batch_create(int device, lnt_batch **out) {
    if (!out) return fail(nullptr, LNT_INVALID_ARGUMENT, "out is NULL");
    lnt_batch *b = new lnt_batch;
    b->device = device;
    if (hipSetDevice(device) != hipSuccess || hipStreamCreateWithFlags(&b->stream, hipStreamNonBlocking) != hipSuccess ||
        hipEventCreate(&b->ev_mark) != hipSuccess || hipEventCreateWithFlags(&b->fork, hipEventDisableTiming) != hipSuccess ||
        hipEventCreateWithFlags(&b->done, hipEventDisableTiming) != hipSuccess ||
        hipEventRecord(b->ev_mark, b->stream) != hipSuccess) {
        delete b;
        return fail(nullptr, LNT_DEVICE_ERROR, "lnt_batch_create: HIP stream/event set-up failed");
    }
    *out = b;
    return LNT_OK;
}

int lnt_batch_add(lnt_batch *b, lnt_ctx *c) {
    if (!b || !c) return bfail(b, LNT_INVALID_ARGUMENT, "NULL argument");
    if (c->device != b->device) return bfail(b, LNT_INVALID_ARGUMENT, "context is on another device");
    if (c->batch) return bfail(b, LNT_INVALID_ARGUMENT, "context is already in a batch");
    BHIPCHK(b, hipSetDevice(b->device), "hipSetDevice");
    BHIPCHK(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    b->members.push_back(c);
    c->batch = b;
    lnt_batch::Group *g = b->dirty ? nullptr : group_of(b, c);
    u32 slot = 0;
    while (g && slot < g->cap && g->ctxs[slot]) slot++;
    if (g && slot < g->cap) {  // a free slot of an existing group: new parameters only
        g->ctxs[slot] = c;
        g->h[slot] = member_params(c);
        b->upload = true;
    } else {
        b->dirty = true;
    }
    return LNT_OK;
}

int lnt_batch_remove(lnt_batch *b, lnt_ctx *c) {
    if (!b || !c) return bfail(b, LNT_INVALID_ARGUMENT, "NULL argument");
    BHIPCHK(b, hipSetDevice(b->device), "hipSetDevice");
    BHIPCHK(b, hipStreamSynchronize(b->stream), "hipStreamSynchronize");
    for (size_t i = 0; i < b->members.size(); i++)
        if (b->members[i] == c) {
            b->members.erase(b->members.begin() + i);
            c->batch = nullptr;
            for (auto &g : b->groups)  // its slot becomes empty
                for (u32 j = 0; j < g.cap; j++)
                    if (g.ctxs[j] == c) {
                        g.ctxs[j] = nullptr;
                        g.h[j] = LntParams{};
                        b->upload = true;
                    }
            return LNT_OK;
        }
    return bfail(b, LNT_INVALID_ARGUMENT, "context is not in the batch");
}

int lnt_batch_size(const lnt_batch *b) { return b ? (int)b->members.size() : 0; }

int lnt_batch_stats(const lnt_batch *b, uint64_t *out, int cap) {
    if (!b || !out || cap < 3) return LNT_INVALID_ARGUMENT;
    out[0] = b->members.size();
    out[1] = b->groups.size();
    out[2] = b->captures;
    return LNT_OK;
}

int lnt_batch_iterate(lnt_batch *b, uint32_t k) {
    if (!b) return fail(nullptr, LNT_INVALID_ARGUMENT, "NULL batch");
    if (k == 0 || b->members.empty()) return LNT_OK;
    BHIPCHK(b, hipSetDevice(b->device), "hipSetDevice");
    for (lnt_ctx *c : b->members) BHIPCHK(b, hipStreamSynchronize(c->stream), "hipStreamSynchronize");
    int rc;
    if (!b->dirty) batch_refresh(b);
    if (b->dirty && (rc = batch_build(b))) return rc;
    if (b->upload && (rc = batch_upload(b))) return rc;
    // Members under a Gerbicz check count the squarings as checked iterations: the
    // replays stop at each of their block ends for the block step on the member's stream.
    // A batch of single-workgroup tests only runs `step` iterations in one launch per
    // group (the kernel loops over iterations), without the graph.
    bool all_small = true;
    for (auto &g : b->groups) all_small = all_small && g.pl.small;
    uint32_t left = k;
    bool first = true;
    while (left) {
        uint32_t step = left;
        for (lnt_ctx *c : b->members)
            if (c->g.on) step = (uint32_t)std::min<uint64_t>(step, c->g.B - c->g.it % c->g.B);
        if (all_small) {
            for (auto &g : b->groups)
                if (!lnt_launch_batch_small(g.pl, g.d, g.cap, step, g.ll, b->stream))
                    return bfail(b, LNT_UNSUPPORTED, "no batched kernel for this shape");
            BHIPCHK(b, hipGetLastError(), "kernel launch");
            if (first && (rc = batch_mark_pending(b))) return rc;
            first = false;
        }
        for (uint32_t i = 0; i < step && !all_small; i++) {
            BHIPCHK(b, hipGraphLaunch(b->exec, b->stream), "hipGraphLaunch");
            if (first && (rc = batch_mark_pending(b))) return rc;
            first = false;
        }
        left -= step;
        // later work on a member's own stream starts after the batch's
        BHIPCHK(b, hipEventRecord(b->done, b->stream), "hipEventRecord");
        for (lnt_ctx *c : b->members) BHIPCHK(b, hipStreamWaitEvent(c->stream, b->done, 0), "hipStreamWaitEvent");
        for (lnt_ctx *c : b->members) {
            if (!c->g.on) continue;
            c->g.it += step;
            if (c->g.it % c->g.B) continue;
            if ((rc = gz_block_end(c))) return bfail(b, rc, std::string("member p=") + std::to_string(c->p) + ": " + c->err);
            BHIPCHK(b, hipEventRecord(c->ev_gz, c->stream), "hipEventRecord");
            BHIPCHK(b, hipStreamWaitEvent(b->stream, c->ev_gz, 0), "hipStreamWaitEvent");
        }
    }
    return LNT_OK;
}

int lnt_batch_sync(lnt_batch *b, float *elapsed_ms) {
    if (!b) return fail(nullptr, LNT_INVALID_ARGUMENT, "NULL batch");
    BHIPCHK(b, hipSetDevice(b->device), "hipSetDevice");
    hipEvent_t now;
    BHIPCHK(b, hipEventCreate(&now), "hipEventCreate");
    BHIPCHK(b, hipEventRecord(now, b->stream), "hipEventRecord");
    BHIPCHK(b, hipEventSynchronize(now), "hipEventSynchronize");
    float ms = 0;
    BHIPCHK(b, hipEventElapsedTime(&ms, b->ev_mark, now), "hipEventElapsedTime");
    (void)hipEventDestroy(b->ev_mark);
    b->ev_mark = now;
    if (elapsed_ms) *elapsed_ms = ms;
    return LNT_OK;
}

void lnt_batch_destroy(lnt_batch *b) {
    if (!b) return;
    (void)hipSetDevice(b->device);
    if (b->stream) (void)hipStreamSynchronize(b->stream);
    for (lnt_ctx *c : b->members) c->batch = nullptr;
    batch_release(b);
    if (b->fork) (void)hipEventDestroy(b->fork);
    if (b->done) (void)hipEventDestroy(b->done);
    if (b->ev_mark) (void)hipEventDestroy(b->ev_mark);
    if (b->stream) (void)hipStreamDestroy(b->stream);
    delete b;
}

const char *lnt_batch_last_error(const lnt_batch *b) { return b ? b->err.c_str() : lnt_last_error(nullptr); }

}  // extern "C"

// ------------------------------------------------------------------ iterations without allocation
// lnt_iterate_static: the squaring of lnt_iterate on a residue in the caller's device memory,
// with the work buffer and every table in `scratch`, a fixed array of this code object
// (one copy per device, held until the process ends).  It allocates no device memory.
// Residues of up to LNT_STATIC_MAX_N words: five n-word buffers (work buffer, weights,
// unweights, 4-step twiddles and their inverses) plus the line twiddles.
static constexpr u64 kStaticMaxN = LNT_STATIC_MAX_N;
static constexpr u64 kStaticWords = 5 * kStaticMaxN + 65536;
static constexpr int kStaticDevices = 64;
static __device__ __align__(256) u64 scratch[kStaticWords];

// x[0] <- x[0] - 2 (field encoding of a signed value): the LL "-2" of the last iteration.
__global__ void lnt_k_minus2(u64 *x) { x[0] = gl_sub(x[0], 2); }

namespace {
struct StaticDevice {
    u64 *base = nullptr;  // device address of `scratch` on this device (hipGetSymbolAddress)
    bool tables = false;  // tables of (p, n) below are in scratch
    u32 p = 0, n = 0;
};
std::mutex g_static_mu;
StaticDevice g_static[kStaticDevices];

// Carve the buffers of plan pl out of scratch (each 256-byte aligned); false if they do
// not fit.
bool static_layout(const Plan &pl, u64 *base, TablePtrs &t) {
    const TableSizes z = table_sizes(pl);
    u64 off = 0;
    auto take = [&](u64 words) {
        u64 *ptr = base + off;
        off += (words + 31) & ~(u64)31;
        return ptr;
    };
    t.y = take(z.n);
    t.w = take(z.W);
    t.uw = take(z.Wu);
    t.twr = take(z.Lr);
    t.itwr = take(z.Lr);
    if (z.large) {
        t.tw4 = take(z.n);
        t.itw4 = take(z.n);
        t.twc = take(z.R);
        t.itwc = take(z.R);
    }
    if (z.A) {
        t.twc2 = take(z.A);
        t.itwc2 = take(z.A);
    }
    if (z.L2) {
        t.twr2 = take(z.L2);
        t.itwr2 = take(z.L2);
    }
    return off <= kStaticWords;
}
}  // namespace

extern "C" int lnt_iterate_static(uint32_t p, int mode, uint64_t *x, size_t n, uint32_t k, void *stream) {
    if (mode != LNT_MODE_LL && mode != LNT_MODE_PRP) return fail(nullptr, LNT_INVALID_ARGUMENT, "bad mode");
    if (!x) return fail(nullptr, LNT_INVALID_ARGUMENT, "lnt_iterate_static: x is NULL");
    if (n > kStaticMaxN)
        return fail(nullptr, LNT_INVALID_ARGUMENT,
                    "lnt_iterate_static: n exceeds the static maximum of " + std::to_string(kStaticMaxN) + " words");
    Plan pl;
    std::string why = make_plan(p, n, pl, PlanOpts());
    if (!why.empty()) return fail(nullptr, LNT_UNSUPPORTED, "lnt_iterate_static: " + why);
    int dev = 0;
    if (hipGetDevice(&dev) != hipSuccess || dev < 0 || dev >= kStaticDevices)
        return fail(nullptr, LNT_DEVICE_ERROR, "lnt_iterate_static: no current device");
    hipStream_t s = (hipStream_t)stream;
    std::lock_guard<std::mutex> lk(g_static_mu);
    StaticDevice &sd = g_static[dev];
    if (!sd.base) {
        // the array's host-side name is not a device address: take this device's copy
        u64 *ptr = nullptr;
        hipError_t e = hipGetSymbolAddress((void **)&ptr, HIP_SYMBOL(scratch));
        if (e != hipSuccess || !ptr)
            return fail(nullptr, LNT_DEVICE_ERROR,
                        std::string("lnt_iterate_static: hipGetSymbolAddress failed (") + hipGetErrorString(e) + ")");
        sd.base = ptr;
    }
    TablePtrs t;
    if (!static_layout(pl, sd.base, t))
        return fail(nullptr, LNT_UNSUPPORTED, "lnt_iterate_static: the plan's tables exceed the static scratch");
    if (!sd.tables || sd.p != p || sd.n != (u32)n) {
        fill_tables(pl, p, t, s);
        sd.tables = true;
        sd.p = p;
        sd.n = (u32)n;
    }
    if (k == 0) return LNT_OK;
    const bool ll = mode == LNT_MODE_LL;
    const LntParams P = params_of(pl, p, x, t);
    bool ok = true;
    if (pl.small) {
        const uint32_t chunk = pl.n >= 2048 ? 256 : 1024;
        bool first = true;
        for (uint32_t left = k; left && ok;) {
            const uint32_t c = left < chunk ? left : chunk;
            ok = lnt_launch_small(pl, P, c, ll && !first, ll, s);
            first = false;
            left -= c;
        }
    } else {
        for (uint32_t i = 0; i < k && ok; i++)
            ok = lnt_launch_col_fwd(pl, P, ll && i > 0, s) && lnt_launch_row(pl, P, s) && lnt_launch_col_inv(pl, P, s);
    }
    if (!ok) return fail(nullptr, LNT_UNSUPPORTED, "lnt_iterate_static: no kernel for this length");
    if (ll) hipLaunchKernelGGL(lnt_k_minus2, dim3(1), dim3(1), 0, s, x);
    hipError_t e = hipGetLastError();
    if (e != hipSuccess) {
        sd.tables = false;
        return fail(nullptr, LNT_DEVICE_ERROR,
                    std::string("lnt_iterate_static: kernel launch (") + hipGetErrorString(e) + ")");
    }
    return LNT_OK;
}

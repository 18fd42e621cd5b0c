// Serving loop over the C ABI (include/lnt_api.h): a scheduler thread admits submitted
// tests while their device footprint fits the budget, advances the admitted tests in
// batches (lnt_batch_*: one hipGraph launch per iteration of all members), lets arrivals
// join and finished tests leave between scheduling steps (continuous batching), and
// queues each result with its submit / admit / done times.  All GPU work is issued from
// the scheduler thread; lnt_server_submit and lnt_server_poll only touch the queues.
//
// Tests are batched by size class (single-workgroup lengths <= 4096, lengths up to 2^17,
// larger), each class on its own stream: a class advances as many iterations per step as
// fit the step's time target, so small tests are not held to the iteration rate of the
// largest ones, and the classes' launches overlap on the device.
//
// The reference runs one test per process (main_process.cu:98-118); this loop has no
// counterpart there.  Residues, Gerbicz checkpoints and tables of admitted tests stay on
// the device until the test finishes.
#include <algorithm>
#include <atomic>
#include <chrono>
#include <condition_variable>
#include <cstring>
#include <deque>
#include <mutex>
#include <set>
#include <string>
#include <thread>
#include <vector>

#include "../../include/lnt_api.h"

// Default Gerbicz check interval in blocks.  A check costs one block of squarings plus
// reading the checked residue back (measured at 2^23, block 1000: 256 ms against
// 214 ms per 1000 iterations), so the overhead is about 1.2 / L: 80 blocks keep it at 1.5%
// (1.8% with the block products).
constexpr uint32_t kDefaultCheckBlocks = 80;

namespace {

struct Pending {
    uint32_t p;
    int prp;
    uint64_t iterations, tag;
    double submit_s;
};

struct Active {
    Pending req;
    lnt_ctx *ctx = nullptr;
    uint64_t bytes = 0, progress = 0;
    bool gz = false;
    int cls = 0;
    double admit_s = 0;
};

constexpr int CLASSES = 3;
int size_class(uint32_t p) {
    const uint32_t l = lnt_secured_log2_length(p);
    return l <= 12 ? 0 : l <= 17 ? 1 : 2;
}

// Device bytes of one test: residue, work buffer, the two 4-step twiddle tables and the
// product operand (8 bytes per word each), six residue slots under the Gerbicz check, and
// the line and weight tables (weights take 2R words each from 2^13 digits on, n words
// each below, within the 256 KiB allowance).
uint64_t footprint(uint32_t p, bool gz) {
    const uint64_t n = 1ull << std::max<uint32_t>(3, lnt_secured_log2_length(p));
    return 8 * n * (5 + (gz ? 6 : 0)) + (256u << 10);
}

}  // namespace

struct lnt_server {
    lnt_server_config cfg{};
    std::chrono::steady_clock::time_point t0;
    mutable std::mutex mu;  // guards the queues, the counters and err
    std::condition_variable cv;
    std::deque<Pending> queue;
    std::deque<lnt_result> results;
    std::vector<Active> active;
    lnt_batch *batch[CLASSES] = {};
    uint64_t bytes_in_use = 0, steps = 0, admitted = 0, finished = 0, iterations_run = 0;
    std::set<uint32_t> tuned;                // NTT lengths tuned (scheduler thread only)
    std::atomic<uint64_t> tuned_count{0};  // its size, for lnt_server_stats
    double ms_per_iter[CLASSES] = {0.05, 0.05, 0.05};
    bool stop = false;
    std::string err;
    std::thread worker;

    void set_err(const std::string &m) {
        std::lock_guard<std::mutex> g(mu);
        err = m;
    }

    double now() const {
        return std::chrono::duration<double>(std::chrono::steady_clock::now() - t0).count();
    }

    void finish(Active &a, int status) {
        lnt_result r{};
        r.tag = a.req.tag;
        r.p = a.req.p;
        r.prp = a.req.prp;
        r.iterations = a.req.iterations;
        r.status = status;
        r.submit_s = a.req.submit_s;
        r.admit_s = a.admit_s;
        if (status == LNT_OK && a.ctx) {
            uint64_t v = 0;
            status = lnt_res64(a.ctx, &v);
            r.res64 = v;
            if (status == LNT_OK && !a.req.prp) {  // the LL verdict needs the whole residue
                const size_t n = lnt_length(a.ctx);
                std::vector<uint64_t> d(n);
                status = lnt_get_digits(a.ctx, d.data(), n);
                r.is_zero = std::all_of(d.begin(), d.end(), [](uint64_t x) { return x == 0; });
            }
            if (a.gz) {
                uint64_t st[5];
                if (lnt_gerbicz_stats(a.ctx, st, 5) == LNT_OK) {
                    r.checks = st[2];
                    r.rollbacks = st[3];
                    r.check_res64 = st[4];
                }
            }
            r.status = status;
        }
        r.done_s = now();
        if (a.ctx) lnt_destroy(a.ctx);  // also leaves the batch
        a.ctx = nullptr;
        std::lock_guard<std::mutex> g(mu);
        bytes_in_use -= a.bytes;
        results.push_back(r);
        finished++;
        cv.notify_all();
    }

    // Admit queued tests in arrival order, skipping those that do not fit yet.
    void admit() {
        std::vector<Pending> take;
        {
            std::lock_guard<std::mutex> g(mu);
            for (auto it = queue.begin(); it != queue.end();) {
                const bool gz = it->prp && cfg.gerbicz_block;
                const uint64_t b = footprint(it->p, gz);
                const bool room = cfg.max_tests == 0 || active.size() + take.size() < cfg.max_tests;
                if (room && bytes_in_use + b <= cfg.hbm_budget) {
                    bytes_in_use += b;
                    take.push_back(*it);
                    it = queue.erase(it);
                } else {
                    ++it;
                }
            }
        }
        for (const Pending &q : take) {
            Active a;
            a.req = q;
            a.gz = q.prp && cfg.gerbicz_block;
            a.bytes = footprint(q.p, a.gz);
            a.cls = size_class(q.p);
            a.admit_s = now();
            int rc = lnt_create(cfg.device, q.p, nullptr, 0, q.prp ? LNT_MODE_PRP : LNT_MODE_LL, &a.ctx);
            if (rc == LNT_OK && cfg.tune_loops && !tuned.count(lnt_length(a.ctx))) {
                // first test of this length: time every plan (validated against each other)
                // and keep the fastest for it and every later context of the length
                tuned.insert(lnt_length(a.ctx));
                tuned_count = tuned.size();
                rc = lnt_tune(a.ctx, (int)cfg.tune_loops);
                if (rc != LNT_OK) set_err(lnt_last_error(a.ctx));
            }
            if (rc == LNT_OK && a.gz) {
                // blocks per check: the configured value, or by default one check every
                // kDefaultCheckBlocks blocks and never fewer than one per request
                uint32_t L = cfg.gerbicz_check;
                if (!L) {
                    const uint64_t blocks = q.iterations / cfg.gerbicz_block;
                    L = (uint32_t)std::max<uint64_t>(1, std::min<uint64_t>(kDefaultCheckBlocks, blocks));
                }
                rc = lnt_gerbicz_begin(a.ctx, cfg.gerbicz_block, L);
            }
            if (rc == LNT_OK) rc = lnt_batch_add(batch[a.cls], a.ctx);
            admitted++;
            if (rc != LNT_OK || q.iterations == 0) {
                finish(a, rc);
                continue;
            }
            active.push_back(a);
        }
    }

    void run() {
        while (true) {
            {
                std::lock_guard<std::mutex> g(mu);
                if (stop) break;
            }
            admit();
            if (active.empty()) {
                std::unique_lock<std::mutex> g(mu);
                cv.wait_for(g, std::chrono::milliseconds(20), [&] { return stop || !queue.empty(); });
                continue;
            }
            // per class: iterations that fit the step's time target, capped by the nearest finish
            uint64_t left[CLASSES], kc[CLASSES] = {};
            for (int c = 0; c < CLASSES; c++) left[c] = UINT64_MAX;
            for (auto &a : active) left[a.cls] = std::min(left[a.cls], a.req.iterations - a.progress);
            int rc = LNT_OK;
            for (int c = 0; c < CLASSES && rc == LNT_OK; c++) {
                if (left[c] == UINT64_MAX) continue;
                uint64_t k = (uint64_t)std::max(1.0, cfg.step_ms / std::max(ms_per_iter[c], 1e-4));
                kc[c] = std::min<uint64_t>({k, left[c], 1u << 16});
                rc = lnt_batch_iterate(batch[c], (uint32_t)kc[c]);
                if (rc != LNT_OK) set_err(lnt_batch_last_error(batch[c]));
            }
            for (int c = 0; c < CLASSES && rc == LNT_OK; c++) {
                if (!kc[c]) continue;
                float ms = 0;
                rc = lnt_batch_sync(batch[c], &ms);
                if (rc != LNT_OK) set_err(lnt_batch_last_error(batch[c]));
                ms_per_iter[c] = 0.5 * ms_per_iter[c] + 0.5 * (ms / (double)kc[c]);
                iterations_run += kc[c];
            }
            if (rc != LNT_OK) {  // a device error ends every admitted test
                for (auto &a : active) finish(a, rc);
                active.clear();
                continue;
            }
            steps++;
            std::vector<Active> keep;
            for (auto &a : active) {
                if (a.gz) {
                    uint64_t st[5];
                    rc = lnt_gerbicz_stats(a.ctx, st, 5);
                    a.progress = st[0];  // a rollback moves it back
                } else {
                    a.progress += kc[a.cls];
                }
                if (rc != LNT_OK || a.progress >= a.req.iterations)
                    finish(a, rc);
                else
                    keep.push_back(a);
            }
            active.swap(keep);
        }
        for (auto &a : active) finish(a, LNT_INVALID_ARGUMENT);  // stopped before the end
        active.clear();
    }
};

extern "C" {

int lnt_server_create(const lnt_server_config *cfg, lnt_server **out) {
    if (!cfg || !out) return LNT_INVALID_ARGUMENT;
    lnt_server *s = new lnt_server;
    s->cfg = *cfg;
    if (s->cfg.step_ms <= 0) s->cfg.step_ms = 5.0f;

    for (int c = 0; c < CLASSES; c++) {
        int rc = lnt_batch_create(cfg->device, &s->batch[c]);
        if (rc != LNT_OK) {
            for (int j = 0; j < c; j++) lnt_batch_destroy(s->batch[j]);
            delete s;
            return rc;
        }
    }
    s->t0 = std::chrono::steady_clock::now();
    s->worker = std::thread([s] { s->run(); });
    *out = s;
    return LNT_OK;
}

int lnt_server_submit(lnt_server *s, uint32_t p, int prp, uint64_t iterations, uint64_t tag) {
    if (!s) return LNT_INVALID_ARGUMENT;
    if (lnt_secured_log2_length(p) > 26 || p < 67) {
        std::lock_guard<std::mutex> g(s->mu);
        s->err = "exponent outside the supported range";
        return LNT_UNSUPPORTED;
    }
    std::lock_guard<std::mutex> g(s->mu);
    s->queue.push_back(Pending{p, prp ? 1 : 0, iterations, tag, s->now()});
    s->cv.notify_all();
    return LNT_OK;
}

int lnt_server_poll(lnt_server *s, lnt_result *out, int cap, int wait_ms) {
    if (!s || !out || cap <= 0) return -LNT_INVALID_ARGUMENT;
    std::unique_lock<std::mutex> g(s->mu);
    if (s->results.empty() && wait_ms > 0)
        s->cv.wait_for(g, std::chrono::milliseconds(wait_ms), [&] { return !s->results.empty(); });
    int k = 0;
    while (k < cap && !s->results.empty()) {
        out[k++] = s->results.front();
        s->results.pop_front();
    }
    return k;
}

int lnt_server_stats(lnt_server *s, uint64_t *out, int cap) {
    if (!s || !out || cap < 7) return LNT_INVALID_ARGUMENT;
    std::lock_guard<std::mutex> g(s->mu);
    out[0] = s->queue.size();
    out[1] = s->admitted - s->finished;
    out[2] = s->finished;
    out[3] = s->bytes_in_use;
    out[4] = s->steps;
    out[5] = s->iterations_run;
    out[6] = s->admitted;
    if (cap >= 8) {
        uint64_t caps = 0, st[3];
        for (int c = 0; c < CLASSES; c++)
            if (lnt_batch_stats(s->batch[c], st, 3) == LNT_OK) caps += st[2];
        out[7] = caps;
    }
    if (cap >= 9) out[8] = s->tuned_count.load();
    return LNT_OK;
}

double lnt_server_now(lnt_server *s) { return s ? s->now() : 0.0; }

void lnt_server_destroy(lnt_server *s) {
    if (!s) return;
    {
        std::lock_guard<std::mutex> g(s->mu);
        s->stop = true;
        s->cv.notify_all();
    }
    if (s->worker.joinable()) s->worker.join();
    for (int c = 0; c < CLASSES; c++) lnt_batch_destroy(s->batch[c]);
    delete s;
}

// The message is copied under the lock into storage of the calling thread: the pointer
// stays valid until this thread's next call, whatever the scheduler thread does meanwhile.
const char *lnt_server_last_error(const lnt_server *s) {
    thread_local std::string copy;
    if (!s) return "";
    std::lock_guard<std::mutex> g(s->mu);
    copy = s->err;
    return copy.c_str();
}

}  // extern "C"

// CPU emulation of the engine's kernels (test infrastructure).
//
// Compiles cuda_amd/csrc/lnt_kernels.hpp with LNT_EMU: each workgroup's threads are
// OpenMP threads, the workgroup barrier is an OpenMP barrier, and every global / LDS
// index is bounds-checked.  It runs the same kernel bodies and the same host logic
// (lnt_host.hpp) as the HIP build, so tests can check the decomposition, the carry and
// the indexing against the oracle before the code runs on a GPU.
#define LNT_EMU 1
#include <omp.h>

#include <cstdio>
#include <cstdlib>
#include <vector>

#include "../../cuda_amd/csrc/lnt_host.hpp"
#include "../../cuda_amd/csrc/lnt_kernels.hpp"

using namespace lnt;

struct Emu {
    Plan pl;
    uint32_t p;
    bool prp, pending;
    std::vector<uint64_t> x, y, f, w, uw, tw4, itw4, twr, itwr, twr2, itwr2, twc, itwc, twc2, itwc2;
    std::vector<uint8_t> wid;
    LntParams P;
};

template <class T>
static GView<T> gv(std::vector<uint64_t> &v) {
    return GView<T>{(T *)v.data(), (u64)v.size()};
}

static void line_tables(uint32_t L, Line ln, std::vector<uint64_t> &tw, std::vector<uint64_t> &itw) {
    tw.resize(L);
    itw.resize(L);
    u64 om = gl_pow(554, (GL_P - 1) / L), iom = gl_inv(om);
    for (uint32_t i = 0; i < L; i++) {
        tw[i] = lnt_line_tw(om, ln.B1, ln.B2, i);
        itw[i] = lnt_line_tw(iom, ln.B1, ln.B2, i);
    }
}

extern "C" void *emu_create3(uint32_t p, uint32_t n, int prp, uint32_t lc, uint32_t K, int row3, int col3) {
    Emu *e = new Emu;
    e->p = p;
    e->prp = prp;
    PlanOpts o;
    o.lc = lc;
    o.K = K;
    o.row3 = row3;
    o.col3 = col3;
    std::string err = make_plan(p, n, e->pl, o);
    if (!err.empty()) {
        std::fprintf(stderr, "emu_create: %s\n", err.c_str());
        delete e;
        return nullptr;
    }
    const Plan &pl = e->pl;
    e->x.assign(n, 0);
    e->y.assign(n, 0);
    e->f.assign(n, 0);
    u64 Bw = gl_pow(7, (GL_P - 1) / 192 / n * 5);
    u64 Binv = gl_inv(Bw);
    u64 ninv = gl_inv(n);
    if (pl.small) {  // weights as (w, -2^31 w) pairs
        e->w.resize(2 * n);
        e->uw.resize(n);
        for (uint32_t i = 0; i < n; i++) {
            const u64 wi = lnt_weight(Bw, p, n, i);
            e->w[2 * i] = wi;
            e->w[2 * i + 1] = gl_mul(wi, GL_P - (1ull << 31));
            e->uw[i] = gl_mul(gl_inv(wi), ninv);
        }
    } else {  // row factors (weights as pairs); the column factors go into tw4 / itw4 below
        e->w.resize(4 * pl.R);
        e->uw.resize(2 * pl.R);
        for (uint32_t i = 0; i < 4 * pl.R; i++) e->w[i] = lnt_weight_pair(Bw, p, n, pl.C, i);
        for (uint32_t i = 0; i < 2 * pl.R; i++) e->uw[i] = lnt_row_factor(Binv, 2, p, n, pl.C, i);
    }
    widths(p, n, e->wid);
    if (pl.row3m) {
        Line l1, l2;
        l1.B1 = pl.row3m;
        l1.B2 = pl.rBB * pl.rBC;
        l2.B1 = pl.rBB;
        l2.B2 = pl.rBC;
        line_tables(pl.C, l1, e->twr, e->itwr);
        line_tables(pl.rBB * pl.rBC, l2, e->twr2, e->itwr2);
    } else if (pl.row3) {
        Line l1, l2;
        l1.B1 = pl.row3;
        l1.B2 = pl.row3 * pl.row3;
        l2.B1 = l2.B2 = pl.row3;
        line_tables(pl.C, l1, e->twr, e->itwr);
        line_tables(pl.row3 * pl.row3, l2, e->twr2, e->itwr2);
    } else {
        line_tables(pl.small ? n : pl.C, pl.row, e->twr, e->itwr);
        e->twr2.assign(1, 1);
        e->itwr2.assign(1, 1);
    }
    e->twc2.assign(1, 1);
    e->itwc2.assign(1, 1);
    if (!pl.small && pl.col3) {
        Line l1, l2;
        l1.B1 = pl.col3;
        l1.B2 = pl.R / pl.col3;
        l2.B1 = pl.cBB;
        l2.B2 = pl.cBC;
        line_tables(pl.R, l1, e->twc, e->itwc);
        line_tables(pl.R / pl.col3, l2, e->twc2, e->itwc2);
        e->tw4.resize(n);
        e->itw4.resize(n);
        u64 om = gl_pow(554, (GL_P - 1) / n), iom = gl_inv(om);
        for (uint64_t i = 0; i < n; i++) {
            e->tw4[i] = gl_mul(lnt_tw4_3(om, pl.C, pl.col3, pl.cBB, pl.cBC, i), lnt_col_factor(Bw, 1, p, n, pl.C, i));
            e->itw4[i] =
                gl_mul(lnt_tw4_3(iom, pl.C, pl.col3, pl.cBB, pl.cBC, i), lnt_col_factor(Binv, ninv, p, n, pl.C, i));
        }
    } else if (!pl.small) {
        line_tables(pl.R, pl.col, e->twc, e->itwc);
        e->tw4.resize(n);
        e->itw4.resize(n);
        u64 om = gl_pow(554, (GL_P - 1) / n), iom = gl_inv(om);
        for (uint64_t i = 0; i < n; i++) {
            e->tw4[i] = gl_mul(lnt_tw4(om, pl.C, pl.col.B1, pl.col.B2, i), lnt_col_factor(Bw, 1, p, n, pl.C, i));
            e->itw4[i] = gl_mul(lnt_tw4(iom, pl.C, pl.col.B1, pl.col.B2, i), lnt_col_factor(Binv, ninv, p, n, pl.C, i));
        }
    } else {
        e->twc.assign(1, 1);
        e->itwc.assign(1, 1);
        e->tw4.assign(1, 1);
        e->itw4.assign(1, 1);
    }
    LntParams &P = e->P;
    P.x = gv<u64>(e->x);
    P.y = gv<u64>(e->y);
    P.w = gv<const u64>(e->w);
    P.uw = gv<const u64>(e->uw);
    P.tw4 = gv<const u64>(e->tw4);
    P.itw4 = gv<const u64>(e->itw4);
    P.twr = gv<const u64>(e->twr);
    P.itwr = gv<const u64>(e->itwr);
    P.twr2 = gv<const u64>(e->twr2);
    P.itwr2 = gv<const u64>(e->itwr2);
    P.twc = gv<const u64>(e->twc);
    P.itwc = gv<const u64>(e->itwc);
    P.twc2 = gv<const u64>(e->twc2);
    P.itwc2 = gv<const u64>(e->itwc2);
    P.p = p;
    P.n = n;
    P.log_n = pl.log_n;
    P.C = pl.C;
    P.R = pl.R;
    P.levels = pl.levels;
    P.bias = pl.bias ? 1u : 0u;
    std::vector<uint64_t> d(n, 0);
    d[0] = prp ? 3 : 4;
    to_device_digits(e->wid, d.data(), e->x.data());
    e->pending = false;
    return e;
}

extern "C" void *emu_create2(uint32_t p, uint32_t n, int prp, uint32_t lc, uint32_t K, int row3) {
    return emu_create3(p, n, prp, lc, K, row3, -1);
}

extern "C" void *emu_create(uint32_t p, uint32_t n, int prp, uint32_t force_lc) {
    return emu_create2(p, n, prp, force_lc, 0, -1);
}

extern "C" void emu_destroy(void *h) { delete (Emu *)h; }

template <class F>
static void run_wg(int threads, uint32_t lds_words, F body) {
    std::vector<u64> lds(lds_words, 0xDEADBEEFDEADBEEFull);
    LView L{lds.data(), lds_words};
#pragma omp parallel num_threads(threads)
    {
        int tid = omp_get_thread_num();
        body(tid, L);
    }
}

template <int B1, int B2, int M, int MODE = LNT_SQUARE>
static void small_iter(Emu *e, uint32_t k, const LntParams &P, bool sub2) {
    uint32_t lds = std::max<uint32_t>(B1 * B2, B1 * (B2 + LNT_PAD));
    run_wg(B2, lds, [&](int tid, LView L) { lnt_small_body<B1, B2, M, MODE>(tid, L, P, k, sub2, !e->prp); });
}

template <int B, int M, int MODE = LNT_SQUARE>
static void small3_iter(Emu *e, uint32_t k, const LntParams &P, bool sub2) {
    run_wg(B * B, lnt_small3_lds<B>(), [&](int tid, LView L) { lnt_small3_body<B, M, MODE>(tid, L, P, k, sub2, !e->prp); });
}

template <int B>
static void small3_any(Emu *e, uint32_t k, const LntParams &P, bool sub2, int mode) {
    if (mode == LNT_FORWARD) {
        if (e->pl.M == 4) small3_iter<B, 4, LNT_FORWARD>(e, 1, P, sub2);
        else small3_iter<B, 8, LNT_FORWARD>(e, 1, P, sub2);
    } else if (mode == LNT_MULTIPLY) {
        if (e->pl.M == 4) small3_iter<B, 4, LNT_MULTIPLY>(e, 1, P, sub2);
        else small3_iter<B, 8, LNT_MULTIPLY>(e, 1, P, sub2);
    } else if (e->pl.M == 4) small3_iter<B, 4>(e, k, P, sub2);
    else small3_iter<B, 8>(e, k, P, sub2);
}

template <int B1, int B2, int K, int M>
static void col_fwd(Emu *e, bool sub2, const LntParams &P) {
    const Plan &pl = e->pl;
    uint32_t lds = lnt_col_fwd_lds<B1, B2, K>();
    for (uint32_t b = 0; b < pl.C / K; b++)
        run_wg(K * B2, lds, [&](int tid, LView L) { lnt_col_fwd_body<B1, B2, K, M>(tid, b, L, P, sub2); });
}

template <int B1, int B2, int K>
static void col_inv(Emu *e, const LntParams &P) {
    const Plan &pl = e->pl;
    uint32_t lds = B1 * (B2 + LNT_PAD) * K;
    for (uint32_t b = 0; b < pl.C / K; b++)
        run_wg(K * B2, lds, [&](int tid, LView L) { lnt_col_inv_body<B1, B2, K>(tid, b, L, P); });
}

template <int BA, int BB, int BC, int K, int M>
static void col3_fwd(Emu *e, bool sub2, const LntParams &P) {
    const Plan &pl = e->pl;
    for (uint32_t b = 0; b < pl.C / K; b++)
        run_wg(K * BB * BC, LntCol3<BA, BB, BC, K>::LDS,
               [&](int tid, LView L) { lnt_col3_fwd_body<BA, BB, BC, K, M>(tid, b, L, P, sub2); });
}

template <int BA, int BB, int BC, int K>
static void col3_inv(Emu *e, const LntParams &P) {
    const Plan &pl = e->pl;
    for (uint32_t b = 0; b < pl.C / K; b++)
        run_wg(K * BB * BC, LntCol3<BA, BB, BC, K>::LDS,
               [&](int tid, LView L) { lnt_col3_inv_body<BA, BB, BC, K>(tid, b, L, P); });
}

template <int B, int G>
static void row3_sq(Emu *e, const LntParams &P, int mode) {
    const Plan &pl = e->pl;
    uint32_t lds = G * B * B * B;
    for (uint32_t b = 0; b < pl.R / G; b++)
        run_wg(G * B * B, lds, [&](int tid, LView L) {
            if (mode == LNT_FORWARD) lnt_row3_body<B, G, LNT_FORWARD>(tid, b, L, P);
            else if (mode == LNT_MULTIPLY) lnt_row3_body<B, G, LNT_MULTIPLY>(tid, b, L, P);
            else lnt_row3_body<B, G>(tid, b, L, P);
        });
}

template <int BA, int BB, int BC>
static void row3m_sq(Emu *e, const LntParams &P, int mode) {
    const Plan &pl = e->pl;
    for (uint32_t b = 0; b < pl.R; b++)
        run_wg(BB * BC, LntRow3m<BA, BB, BC>::LDS, [&](int tid, LView L) {
            if (mode == LNT_FORWARD) lnt_row3m_body<BA, BB, BC, LNT_FORWARD>(tid, b, L, P);
            else if (mode == LNT_MULTIPLY) lnt_row3m_body<BA, BB, BC, LNT_MULTIPLY>(tid, b, L, P);
            else lnt_row3m_body<BA, BB, BC>(tid, b, L, P);
        });
}

template <int B1, int B2, int G>
static void row_sq(Emu *e, const LntParams &P, int mode) {
    const Plan &pl = e->pl;
    uint32_t lds = G * B1 * (B2 + LNT_PAD);
    for (uint32_t b = 0; b < pl.R / G; b++)
        run_wg(G * B2, lds, [&](int tid, LView L) {
            if (mode == LNT_FORWARD) lnt_row_body<B1, B2, G, LNT_FORWARD>(tid, b, L, P);
            else if (mode == LNT_MULTIPLY) lnt_row_body<B1, B2, G, LNT_MULTIPLY>(tid, b, L, P);
            else lnt_row_body<B1, B2, G>(tid, b, L, P);
        });
}

// Dispatch tables for the emulated sizes (a subset of the HIP build's).
static bool run_small(Emu *e, uint32_t k, const LntParams &P, bool sub2, int mode = LNT_SQUARE) {
    const Plan &pl = e->pl;
    if (pl.row3 == 16) return small3_any<16>(e, k, P, sub2, mode), true;
    if (pl.row3 == 8) return small3_any<8>(e, k, P, sub2, mode), true;
#define SM(b1, b2)                                                                          \
    if (pl.row.B1 == b1 && pl.row.B2 == b2) {                                               \
        if (mode == LNT_FORWARD) {                                                          \
            if (pl.M == 4) small_iter<b1, b2, 4, LNT_FORWARD>(e, 1, P, sub2);               \
            else small_iter<b1, b2, 8, LNT_FORWARD>(e, 1, P, sub2);                         \
        } else if (mode == LNT_MULTIPLY) {                                                  \
            if (pl.M == 4) small_iter<b1, b2, 4, LNT_MULTIPLY>(e, 1, P, sub2);              \
            else small_iter<b1, b2, 8, LNT_MULTIPLY>(e, 1, P, sub2);                        \
        } else if (pl.M == 4) small_iter<b1, b2, 4>(e, k, P, sub2);                         \
        else small_iter<b1, b2, 8>(e, k, P, sub2);                                          \
        return true;                                                                        \
    }
    SM(4, 2) SM(4, 4) SM(8, 4) SM(8, 8) SM(16, 8) SM(16, 16) SM(32, 16) SM(32, 32) SM(64, 32) SM(64, 64)
#undef SM
    return false;
}

// one pass: mode LNT_SQUARE = col_fwd, row square, col_inv; LNT_FORWARD = col_fwd, row
// forward (transform left in P.y); LNT_MULTIPLY = col_fwd, row multiply by P.f, col_inv
static bool run_large_once(Emu *e, bool sub2, const LntParams &P, int mode = LNT_SQUARE) {
    const Plan &pl = e->pl;
    bool ok = false;
#define CF(b1, b2, k)                                                            \
    if (!ok && pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) {                \
        if (pl.M == 4) col_fwd<b1, b2, k, 4>(e, sub2, P); else col_fwd<b1, b2, k, 8>(e, sub2, P); \
        ok = true;                                                               \
    }
    if (!pl.col3) {
        CF(4, 4, 16) CF(8, 4, 16) CF(8, 8, 16) CF(16, 8, 16) CF(16, 16, 16) CF(32, 16, 16) CF(32, 32, 16)
        CF(32, 32, 8) CF(64, 32, 8) CF(64, 32, 4) CF(64, 64, 4)
    }
#undef CF
#define C3F(ba, bb, bc, k)                                                                    \
    if (!ok && pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) {                  \
        if (pl.M == 4) col3_fwd<ba, bb, bc, k, 4>(e, sub2, P); else col3_fwd<ba, bb, bc, k, 8>(e, sub2, P); \
        ok = true;                                                                            \
    }
    if (pl.col3) {
        C3F(8, 8, 8, 16) C3F(8, 8, 8, 8) C3F(16, 8, 8, 16) C3F(16, 8, 8, 8) C3F(16, 16, 8, 8) C3F(16, 16, 8, 4)
        C3F(16, 16, 16, 4)
    }
#undef C3F
    if (!ok) return false;
    ok = false;
    if (pl.row3m == 32 && pl.rBB == 16 && pl.rBC == 16) row3m_sq<32, 16, 16>(e, P, mode), ok = true;
    if (pl.row3m == 32 && pl.rBB == 32 && pl.rBC == 16) row3m_sq<32, 32, 16>(e, P, mode), ok = true;
    if (pl.row3m == 16 && pl.rBB == 32 && pl.rBC == 16) row3m_sq<16, 32, 16>(e, P, mode), ok = true;
    if (pl.row3m == 16 && pl.rBB == 32 && pl.rBC == 32) row3m_sq<16, 32, 32>(e, P, mode), ok = true;
    if (pl.row3 == 16) row3_sq<16, 1>(e, P, mode), ok = true;
    if (pl.row3 == 8) row3_sq<8, 4>(e, P, mode), ok = true;
#define RW(b1, b2)                                                        \
    if (!ok && !pl.row3m && pl.row.B1 == b1 && pl.row.B2 == b2) {                      \
        row_sq<b1, b2, 64 / b2>(e, P, mode);                              \
        ok = true;                                                        \
    }
    RW(16, 8) RW(16, 16) RW(32, 16) RW(32, 32) RW(64, 32) RW(64, 64)
#undef RW
    if (!ok) return false;
    if (mode == LNT_FORWARD) return true;
    ok = false;
#define CI(b1, b2, k)                                                           \
    if (!ok && pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) {               \
        col_inv<b1, b2, k>(e, P);                                               \
        ok = true;                                                              \
    }
    if (!pl.col3) {
        CI(4, 4, 16) CI(8, 4, 16) CI(8, 8, 16) CI(16, 8, 16) CI(16, 16, 16) CI(32, 16, 16) CI(32, 32, 16)
        CI(32, 32, 8) CI(64, 32, 8) CI(64, 32, 4) CI(64, 64, 4)
    }
#undef CI
#define C3I(ba, bb, bc, k)                                                    \
    if (!ok && pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) {  \
        col3_inv<ba, bb, bc, k>(e, P);                                        \
        ok = true;                                                            \
    }
    if (pl.col3) {
        C3I(8, 8, 8, 16) C3I(8, 8, 8, 8) C3I(16, 8, 8, 16) C3I(16, 8, 8, 8) C3I(16, 16, 8, 8) C3I(16, 16, 8, 4)
        C3I(16, 16, 16, 4)
    }
#undef C3I
    return ok;
}

extern "C" int emu_iterate(void *h, uint32_t k) {
    Emu *e = (Emu *)h;
    if (k == 0) return 0;
    if (e->pl.small) {
        if (!run_small(e, k, e->P, !e->prp && e->pending)) return -1;
    } else {
        for (uint32_t i = 0; i < k; i++) {
            bool sub2 = !e->prp && e->pending;
            if (!run_large_once(e, sub2, e->P)) return -1;
            e->pending = true;
        }
    }
    e->pending = true;
    return 0;
}

extern "C" void emu_get_digits(void *h, uint64_t *d) {
    Emu *e = (Emu *)h;
    to_canonical_digits(e->wid, e->x.data(), !e->prp && e->pending, d);
}

extern "C" uint64_t emu_res64(void *h) {
    Emu *e = (Emu *)h;
    std::vector<uint64_t> d(e->pl.n);
    emu_get_digits(h, d.data());
    return res64(e->wid, d.data());
}

extern "C" void emu_set_digits(void *h, const uint64_t *d) {
    Emu *e = (Emu *)h;
    to_device_digits(e->wid, d, e->x.data());
    e->pending = false;
}

extern "C" uint32_t emu_levels(void *h) { return ((Emu *)h)->pl.levels; }
extern "C" uint32_t emu_secured(uint32_t p) { return secured_log2_length(p); }
extern "C" uint32_t emu_carry_levels(uint32_t p, uint32_t n) { return carry_levels(p, n); }
extern "C" uint64_t emu_gl_mul_biased(uint32_t db, uint64_t w, uint64_t wb) { return gl_mul_biased(db, w, wb); }
extern "C" int emu_digits_fit_31(uint32_t p, uint32_t n) {
    const uint32_t levels = carry_levels(p, n);
    return levels ? (int)digits_fit_31(p, n, levels) : -1;
}

// The plan make_plan picks (no tables): C, R, row3, row3m, rBB, rBC, col3, K; 0 or -1 if rejected.
extern "C" int emu_plan(uint32_t p, uint32_t n, int lc, int row3, uint32_t *out) {
    PlanOpts o;
    o.lc = lc;
    o.row3 = row3;
    Plan pl;
    if (!make_plan(p, n, pl, o).empty()) return -1;
    const uint32_t v[8] = {pl.C, pl.R, pl.row3, pl.row3m, pl.rBB, pl.rBC, pl.col3, pl.K};
    for (int i = 0; i < 8; i++) out[i] = v[i];
    return 0;
}

// make_context's plan flow: the given options, and on failure the default plan built into
// the same Plan object.  out: n, C, R, col3, cBB, cBC, K, row3, row3m, G; returns 1 when the
// options were used, 0 for the fallback, -1 if both fail; *consistent = plan_consistent.
extern "C" int emu_plan_retry(uint32_t p, uint32_t n, uint32_t lc, uint32_t K, int row3, int col3, uint32_t *out,
                              int *consistent) {
    PlanOpts o;
    o.lc = lc;
    o.K = K;
    o.row3 = row3;
    o.col3 = col3;
    Plan pl;
    int used = 1;
    if (!make_plan(p, n, pl, o).empty()) {
        used = 0;
        if (!make_plan(p, n, pl, PlanOpts()).empty()) return -1;
    }
    const uint32_t v[10] = {pl.n, pl.C, pl.R, pl.col3, pl.cBB, pl.cBC, pl.K, pl.row3, pl.row3m, pl.G};
    for (int i = 0; i < 10; i++) out[i] = v[i];
    *consistent = plan_consistent(pl) ? 1 : 0;
    return used;
}

// dst <- dst * src (same shape), through the forward / multiply kernel modes.
extern "C" int emu_mul(void *hd, void *hs) {
    Emu *d = (Emu *)hd, *s = (Emu *)hs;
    LntParams Ps = s->P;
    Ps.y = GView<u64>{d->f.data(), (u64)d->f.size()};
    LntParams Pd = d->P;
    Pd.f = GView<const u64>{d->f.data(), (u64)d->f.size()};
    const bool ll = !d->prp;
    if (d->pl.small) {
        if (!run_small(s, 1, Ps, ll && s->pending, LNT_FORWARD)) return -1;
        if (!run_small(d, 1, Pd, ll && d->pending, LNT_MULTIPLY)) return -1;
    } else {
        if (!run_large_once(s, ll && s->pending, Ps, LNT_FORWARD)) return -1;
        if (!run_large_once(d, ll && d->pending, Pd, LNT_MULTIPLY)) return -1;
    }
    d->pending = false;
    return 0;
}

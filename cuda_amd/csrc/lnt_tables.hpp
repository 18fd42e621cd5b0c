// Tables of a plan (weights, unweights, line and 4-step twiddles) and the kernel
// parameters over them: the kernels that fill the tables, their sizes, and the
// LntParams a plan's kernels take.  Shared by the context path and the allocation-free
// path of the C ABI (lnt_api.hip).
#pragma once
#include <hip/hip_runtime.h>

#include "lnt_host.hpp"
#include "lnt_kernels.hpp"

namespace lnt {

// ------------------------------------------------------------------ table kernels

// one-workgroup path: digit weights as (w_i, -2^31 w_i) pairs (2n words), unweights (n words)
static __global__ void lnt_k_weights(u64 Bw, u64 Binv, u64 ninv, u32 p, u32 n, u64 *w, u64 *uw) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    const u64 wi = lnt_weight(Bw, p, n, (u32)i);
    w[2 * i] = wi;
    w[2 * i + 1] = gl_mul(wi, GL_P - (1ull << 31));
    uw[i] = i ? gl_mul(gl_pow(Binv, (u64)(n - ((u32)((u64)p * i) & (n - 1)))), ninv) : ninv;
}

// large path: L = 2R row slots; weights as (w, -2^31 w) pairs (2L words), unweights (L words)
static __global__ void lnt_k_row_weights(u64 Bw, u64 Binv, u32 p, u32 n, u32 C, u32 L, u64 *w, u64 *uw) {
    const u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= 2 * L) return;
    w[i] = lnt_weight_pair(Bw, p, n, C, i);
    if (i < L) uw[i] = lnt_row_factor(Binv, 2, p, n, C, i);
}

static __global__ void lnt_k_line_tw(u64 om, u32 B1, u32 B2, u32 L, u64 *tw) {
    u32 i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= L) return;
    tw[i] = lnt_line_tw(om, B1, B2, i);
}

// 4-step twiddles times the column weight factor B^e_c * s (lnt_col_factor)
static __global__ void lnt_k_tw4(u64 om, u32 C, u32 B1, u32 B2, u32 n, u64 B, u64 s, u32 p, u64 *tw) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tw[i] = gl_mul(lnt_tw4(om, C, B1, B2, i), lnt_col_factor(B, s, p, n, C, i));
}

static __global__ void lnt_k_tw4_3(u64 om, u32 C, u32 BA, u32 BB, u32 BC, u32 n, u64 B, u64 s, u32 p, u64 *tw) {
    u64 i = (u64)blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= n) return;
    tw[i] = gl_mul(lnt_tw4_3(om, C, BA, BB, BC, i), lnt_col_factor(B, s, p, n, C, i));
}

template <class T>
static GView<T> view(T *p, u64 n) {
    return GView<T>{p, n};
}

// Word counts of the tables of a plan (0 where the plan has no such table).
struct TableSizes {
    u32 n, Lr, L2, R, A;  // residue, row line, second-level row line, column line, second-level column line
    u32 W, Wu;            // weight / unweight table words: 2n / n (one workgroup) or 4R / 2R
    bool large;
};
static TableSizes table_sizes(const Plan &pl) {
    TableSizes t;
    t.n = pl.n;
    t.large = !pl.small;
    t.Lr = pl.small ? pl.n : pl.C;
    t.L2 = pl.row3m ? pl.rBB * pl.rBC : pl.row3 ? pl.row3 * pl.row3 : 0;
    t.R = pl.small ? 0 : pl.R;
    t.A = (!pl.small && pl.col3) ? pl.R / pl.col3 : 0;
    t.W = pl.small ? 2 * pl.n : 4 * pl.R;
    t.Wu = pl.small ? pl.n : 2 * pl.R;
    return t;
}

// Device pointers of one set of tables (and the work buffer y).
struct TablePtrs {
    u64 *y = nullptr, *w = nullptr, *uw = nullptr, *tw4 = nullptr, *itw4 = nullptr;
    u64 *twr = nullptr, *itwr = nullptr, *twr2 = nullptr, *itwr2 = nullptr;
    u64 *twc = nullptr, *itwc = nullptr, *twc2 = nullptr, *itwc2 = nullptr;
};

// Launch the kernels that fill every table of plan pl (exponent p) on stream s.
static void fill_tables(const Plan &pl, u32 p, const TablePtrs &t, hipStream_t s) {
    const TableSizes z = table_sizes(pl);
    const u32 n = pl.n;
    const u64 Bw = gl_pow(7, (GL_P - 1) / 192 / n * 5);
    const u64 ninv = gl_inv(n), Binv = gl_inv(Bw);
    if (z.large)
        hipLaunchKernelGGL(lnt_k_row_weights, dim3((z.W + 255) / 256), dim3(256), 0, s, Bw, Binv, p, n, pl.C, z.Wu,
                           t.w, t.uw);
    else
        hipLaunchKernelGGL(lnt_k_weights, dim3((n + 255) / 256), dim3(256), 0, s, Bw, Binv, ninv, p, n, t.w, t.uw);
    const u32 Lr = z.Lr;
    u64 omr = gl_pow(554, (GL_P - 1) / Lr);
    // row line twiddles: first phase B1 x B2 (3-phase rows: row3 x row3^2, row3m x rBB rBC)
    const u32 rB1 = pl.row3m ? pl.row3m : pl.row3 ? pl.row3 : pl.row.B1;
    const u32 rB2 = pl.row3m ? pl.rBB * pl.rBC : pl.row3 ? pl.row3 * pl.row3 : pl.row.B2;
    hipLaunchKernelGGL(lnt_k_line_tw, dim3((Lr + 255) / 256), dim3(256), 0, s, omr, rB1, rB2, Lr, t.twr);
    hipLaunchKernelGGL(lnt_k_line_tw, dim3((Lr + 255) / 256), dim3(256), 0, s, gl_inv(omr), rB1, rB2, Lr, t.itwr);
    if (z.large) {
        u64 omc = gl_pow(554, (GL_P - 1) / pl.R);
        u64 omn = gl_pow(554, (GL_P - 1) / n);
        if (pl.col3) {
            // phase 1 (BA x A line) and phase 2 (BB x BC line) twiddles, 3-phase slot order
            const u32 A = z.A;
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((pl.R + 255) / 256), dim3(256), 0, s, omc, pl.col3, A, pl.R, t.twc);
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((pl.R + 255) / 256), dim3(256), 0, s, gl_inv(omc), pl.col3, A, pl.R,
                               t.itwc);
            u64 oma = gl_pow(554, (GL_P - 1) / A);
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((A + 255) / 256), dim3(256), 0, s, oma, pl.cBB, pl.cBC, A, t.twc2);
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((A + 255) / 256), dim3(256), 0, s, gl_inv(oma), pl.cBB, pl.cBC, A,
                               t.itwc2);
            hipLaunchKernelGGL(lnt_k_tw4_3, dim3((n + 255) / 256), dim3(256), 0, s, omn, pl.C, pl.col3, pl.cBB, pl.cBC, n,
                               Bw, (u64)1, p, t.tw4);
            hipLaunchKernelGGL(lnt_k_tw4_3, dim3((n + 255) / 256), dim3(256), 0, s, gl_inv(omn), pl.C, pl.col3, pl.cBB,
                               pl.cBC, n, Binv, ninv, p, t.itw4);
        } else {
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((pl.R + 255) / 256), dim3(256), 0, s, omc, pl.col.B1, pl.col.B2, pl.R,
                               t.twc);
            hipLaunchKernelGGL(lnt_k_line_tw, dim3((pl.R + 255) / 256), dim3(256), 0, s, gl_inv(omc), pl.col.B1,
                               pl.col.B2, pl.R, t.itwc);
            hipLaunchKernelGGL(lnt_k_tw4, dim3((n + 255) / 256), dim3(256), 0, s, omn, pl.C, pl.col.B1, pl.col.B2, n,
                               Bw, (u64)1, p, t.tw4);
            hipLaunchKernelGGL(lnt_k_tw4, dim3((n + 255) / 256), dim3(256), 0, s, gl_inv(omn), pl.C, pl.col.B1,
                               pl.col.B2, n, Binv, ninv, p, t.itw4);
        }
    }
    // second-phase row twiddles: a line of L2 = B2 words split (row3 x row3 or rBB x rBC)
    if (z.L2) {
        const u32 s1 = pl.row3m ? pl.rBB : pl.row3, s2 = pl.row3m ? pl.rBC : pl.row3;
        u64 om2 = gl_pow(554, (GL_P - 1) / z.L2);
        hipLaunchKernelGGL(lnt_k_line_tw, dim3((z.L2 + 255) / 256), dim3(256), 0, s, om2, s1, s2, z.L2, t.twr2);
        hipLaunchKernelGGL(lnt_k_line_tw, dim3((z.L2 + 255) / 256), dim3(256), 0, s, gl_inv(om2), s1, s2, z.L2,
                           t.itwr2);
    }
}

// Kernel parameters of plan pl over residue x and the tables t.
static LntParams params_of(const Plan &pl, u32 p, u64 *x, const TablePtrs &t) {
    const TableSizes z = table_sizes(pl);
    const u32 n = pl.n;
    LntParams P{};
    P.x = view<u64>(x, n);
    P.y = view<u64>(t.y, n);
    P.w = view<const u64>(t.w, z.W);
    P.uw = view<const u64>(t.uw, z.Wu);
    P.tw4 = view<const u64>(t.tw4, z.large ? n : 0);
    P.itw4 = view<const u64>(t.itw4, z.large ? n : 0);
    P.twr = view<const u64>(t.twr, z.Lr);
    P.itwr = view<const u64>(t.itwr, z.Lr);
    P.twr2 = view<const u64>(t.twr2, z.L2);
    P.itwr2 = view<const u64>(t.itwr2, z.L2);
    P.twc = view<const u64>(t.twc, z.R);
    P.itwc = view<const u64>(t.itwc, z.R);
    P.twc2 = view<const u64>(t.twc2, z.A);
    P.itwc2 = view<const u64>(t.itwc2, z.A);
    P.p = p;
    P.n = n;
    P.log_n = pl.log_n;
    P.C = pl.C;
    P.R = pl.R;
    P.levels = pl.levels;
    P.bias = pl.bias ? 1u : 0u;
    return P;
}

}  // namespace lnt

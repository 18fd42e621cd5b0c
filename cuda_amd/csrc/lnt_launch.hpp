// HIP kernel wrappers and launchers for the bodies in lnt_kernels.hpp.
// Each kernel sizes its LDS statically from its template shape (gfx950 allows up to
// 160 KiB per workgroup).  The launchers map a Plan to the matching instantiation and
// return false when the shape has none.
#pragma once
#include <hip/hip_runtime.h>

#include "lnt_host.hpp"
#include "lnt_kernels.hpp"

template <int B1, int B2, int M, int MODE>
__global__ __launch_bounds__(B2 < 64 ? 64 : B2) void lnt_k_small(LntParams P, u32 iters, int sub2_first, int ll) {
    constexpr u32 W = (B1 * B2 > B1 * (B2 + LNT_PAD)) ? B1 * B2 : B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    lnt_small_body<B1, B2, M, MODE>(threadIdx.x, LView{smem, W}, P, iters, sub2_first != 0, ll != 0);
}

template <int B1, int B2, int K, int M>
__global__ __launch_bounds__(K * B2) void lnt_k_col_fwd(LntParams P, int sub2) {
    constexpr u32 A = B1 * B2 * K, B = B1 * (B2 + LNT_PAD) * K;
    constexpr u32 W = A > B ? A : B;
    __shared__ u64 smem[W];
    lnt_col_fwd_body<B1, B2, K, M>(threadIdx.x, blockIdx.x, LView{smem, W}, P, sub2 != 0);
}

template <int B1, int B2, int K>
__global__ __launch_bounds__(K * B2) void lnt_k_col_inv(LntParams P) {
    constexpr u32 W = B1 * (B2 + LNT_PAD) * K;
    __shared__ u64 smem[W];
    lnt_col_inv_body<B1, B2, K>(threadIdx.x, blockIdx.x, LView{smem, W}, P);
}

template <int B1, int B2, int G, int MODE>
__global__ __launch_bounds__(G * B2) void lnt_k_row(LntParams P) {
    constexpr u32 W = G * B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    lnt_row_body<B1, B2, G, MODE>(threadIdx.x, blockIdx.x, LView{smem, W}, P);
}

// Launchers (defined in lnt_k_small.hip / lnt_k_col.hip / lnt_k_row.hip).
// mode: LNT_SQUARE (iters squarings), LNT_FORWARD / LNT_MULTIPLY (one pass, iters = 1)
bool lnt_launch_small(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll,
                      hipStream_t s, int mode = LNT_SQUARE);
bool lnt_launch_col_fwd(const lnt::Plan &pl, const LntParams &P, bool sub2, hipStream_t s);
bool lnt_launch_col_inv(const lnt::Plan &pl, const LntParams &P, hipStream_t s);
bool lnt_launch_row(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode = LNT_SQUARE);

// HIP kernel wrappers and launchers for the bodies in lnt_kernels.hpp.
// Each kernel sizes its LDS statically from its template shape (gfx950 allows up to
// 160 KiB per workgroup).  The launchers map a Plan to the matching instantiation and
// return false when the shape has none.
#pragma once
#include <hip/hip_runtime.h>

#include "lnt_host.hpp"
#include "lnt_kernels.hpp"

// Workgroups are dispatched to the 8 XCDs round-robin (workgroup i to XCD i % 8), each
// XCD with its own L2.  Renumbering them so that each XCD gets a contiguous range of
// tiles lets a column tile find its left neighbour's boundary lines (its carry halo) in
// the same L2.
__device__ __forceinline__ u32 lnt_xcd_bid(u32 b, u32 grid) {
    return (grid & 7) ? b : (b & 7) * (grid >> 3) + (b >> 3);
}

template <int B1, int B2, int M, int MODE>
__global__ __launch_bounds__(B2 < 64 ? 64 : B2) void lnt_k_small(LntParams P, u32 iters, int sub2_first, int ll) {
    constexpr u32 W = (B1 * B2 > B1 * (B2 + LNT_PAD)) ? B1 * B2 : B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    lnt_small_body<B1, B2, M, MODE>(threadIdx.x, LView{smem, W}, P, iters, sub2_first != 0, ll != 0);
}

template <int B, int M, int MODE>
__global__ __launch_bounds__(B * B) void lnt_k_small3(LntParams P, u32 iters, int sub2_first, int ll) {
    constexpr u32 W = lnt_small3_lds<B>();
    __shared__ u64 smem[W];
    lnt_small3_body<B, M, MODE>(threadIdx.x, LView{smem, W}, P, iters, sub2_first != 0, ll != 0);
}

template <int B1, int B2, int K, int M>
__global__ __launch_bounds__(K * B2) void lnt_k_col_fwd(LntParams P, int sub2) {
    constexpr u32 W = lnt_col_fwd_lds<B1, B2, K>();
    __shared__ u64 smem[W];
    lnt_col_fwd_body<B1, B2, K, M>(threadIdx.x, lnt_xcd_bid(blockIdx.x, gridDim.x), LView{smem, W}, P, sub2 != 0);
}

template <int B1, int B2, int K>
__global__ __launch_bounds__(K * B2) void lnt_k_col_inv(LntParams P) {
    constexpr u32 W = B1 * (B2 + LNT_PAD) * K;
    __shared__ u64 smem[W];
    lnt_col_inv_body<B1, B2, K>(threadIdx.x, lnt_xcd_bid(blockIdx.x, gridDim.x), LView{smem, W}, P);
}

template <int BA, int BB, int BC, int K, int M>
__global__ __launch_bounds__(K * BB * BC) void lnt_k_col3_fwd(LntParams P, int sub2) {
    constexpr u32 W = LntCol3<BA, BB, BC, K>::LDS;
    __shared__ u64 smem[W];
    lnt_col3_fwd_body<BA, BB, BC, K, M>(threadIdx.x, lnt_xcd_bid(blockIdx.x, gridDim.x), LView{smem, W}, P,
                                        sub2 != 0);
}

template <int BA, int BB, int BC, int K>
__global__ __launch_bounds__(K * BB * BC) void lnt_k_col3_inv(LntParams P) {
    constexpr u32 W = LntCol3<BA, BB, BC, K>::LDS;
    __shared__ u64 smem[W];
    lnt_col3_inv_body<BA, BB, BC, K>(threadIdx.x, lnt_xcd_bid(blockIdx.x, gridDim.x), LView{smem, W}, P);
}

template <int B1, int B2, int G, int MODE>
__global__ __launch_bounds__(G * B2) void lnt_k_row(LntParams P) {
    constexpr u32 W = G * B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    lnt_row_body<B1, B2, G, MODE>(threadIdx.x, blockIdx.x, LView{smem, W}, P);
}

// 256-thread rows of 16^3 (or 4 rows of 8^3) words in 32 KiB of LDS; 128 registers per
// lane keep 4 waves per SIMD
template <int B, int G, int MODE>
__global__ __launch_bounds__(G * B * B) __attribute__((amdgpu_waves_per_eu(4))) void lnt_k_row3(LntParams P) {
    constexpr u32 W = G * B * B * B;
    __shared__ u64 smem[W];
    lnt_row3_body<B, G, MODE>(threadIdx.x, blockIdx.x, LView{smem, W}, P);
}

// 16 words per thread: two workgroups per CU fit the LDS when each wave stays within 128
// registers (4 waves per SIMD).
template <int BA, int BB, int BC, int MODE>
__global__ __launch_bounds__(BB * BC) __attribute__((amdgpu_waves_per_eu(BA == 16 ? 4 : 1))) void lnt_k_row3m(
    LntParams P) {
    constexpr u32 W = LntRow3m<BA, BB, BC>::LDS;
    __shared__ u64 smem[W];
    lnt_row3m_body<BA, BB, BC, MODE>(threadIdx.x, blockIdx.x, LView{smem, W}, P);
}

// ---- batched launches: many tests of one shape in one grid of wgs workgroups per slot.
// Row kernels: workgroup b serves slot b / wgs (parameters Ps[b / wgs], uniform per
// workgroup) as its workgroup b % wgs; column kernels interleave the slots (slot b % slots,
// tile b / slots of the XCD-renumbered b).  The "-2" flag is per test (Ps[].sub2).  Each kernel takes a private copy of
// its parameter block: read through a reference into global memory, every field would
// have to be re-read after each global store (possible aliasing), which costs registers.
template <int B1, int B2, int M>
__global__ __launch_bounds__(B2 < 64 ? 64 : B2) void lnt_kb_small(const LntParams *__restrict__ Ps, u32 iters,
                                                                   int ll) {
    constexpr u32 W = (B1 * B2 > B1 * (B2 + LNT_PAD)) ? B1 * B2 : B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    const LntParams P = Ps[blockIdx.x];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_small_body<B1, B2, M>(threadIdx.x, LView{smem, W}, P, iters, P.sub2 != 0, ll != 0);
}

template <int B, int M>
__global__ __launch_bounds__(B * B) void lnt_kb_small3(const LntParams *__restrict__ Ps, u32 iters, int ll) {
    constexpr u32 W = lnt_small3_lds<B>();
    __shared__ u64 smem[W];
    const LntParams P = Ps[blockIdx.x];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_small3_body<B, M>(threadIdx.x, LView{smem, W}, P, iters, P.sub2 != 0, ll != 0);
}

template <int B1, int B2, int K, int M>
__global__ __launch_bounds__(K * B2) void lnt_kb_col_fwd(const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = lnt_col_fwd_lds<B1, B2, K>();
    __shared__ u64 smem[W];
    // member-minor order: the tiles of each XCD's contiguous range cover every slot, so
    // empty slots (and the members) spread over all XCDs while neighbouring tiles of one
    // member stay in one XCD's range
    const u32 b = lnt_xcd_bid(blockIdx.x, gridDim.x), slots = gridDim.x / wgs;
    const LntParams P = Ps[b % slots];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_col_fwd_body<B1, B2, K, M>(threadIdx.x, b / slots, LView{smem, W}, P, P.sub2 != 0);
}

template <int B1, int B2, int K>
__global__ __launch_bounds__(K * B2) void lnt_kb_col_inv(const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = B1 * (B2 + LNT_PAD) * K;
    __shared__ u64 smem[W];
    // member-minor order: the tiles of each XCD's contiguous range cover every slot, so
    // empty slots (and the members) spread over all XCDs while neighbouring tiles of one
    // member stay in one XCD's range
    const u32 b = lnt_xcd_bid(blockIdx.x, gridDim.x), slots = gridDim.x / wgs;
    const LntParams P = Ps[b % slots];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_col_inv_body<B1, B2, K>(threadIdx.x, b / slots, LView{smem, W}, P);
}

template <int BA, int BB, int BC, int K, int M>
__global__ __launch_bounds__(K * BB * BC) void lnt_kb_col3_fwd(const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = LntCol3<BA, BB, BC, K>::LDS;
    __shared__ u64 smem[W];
    // member-minor order: the tiles of each XCD's contiguous range cover every slot, so
    // empty slots (and the members) spread over all XCDs while neighbouring tiles of one
    // member stay in one XCD's range
    const u32 b = lnt_xcd_bid(blockIdx.x, gridDim.x), slots = gridDim.x / wgs;
    const LntParams P = Ps[b % slots];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_col3_fwd_body<BA, BB, BC, K, M>(threadIdx.x, b / slots, LView{smem, W}, P, P.sub2 != 0);
}

template <int BA, int BB, int BC, int K>
__global__ __launch_bounds__(K * BB * BC) void lnt_kb_col3_inv(const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = LntCol3<BA, BB, BC, K>::LDS;
    __shared__ u64 smem[W];
    // member-minor order: the tiles of each XCD's contiguous range cover every slot, so
    // empty slots (and the members) spread over all XCDs while neighbouring tiles of one
    // member stay in one XCD's range
    const u32 b = lnt_xcd_bid(blockIdx.x, gridDim.x), slots = gridDim.x / wgs;
    const LntParams P = Ps[b % slots];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_col3_inv_body<BA, BB, BC, K>(threadIdx.x, b / slots, LView{smem, W}, P);
}

template <int B1, int B2, int G>
__global__ __launch_bounds__(G * B2) void lnt_kb_row(const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = G * B1 * (B2 + LNT_PAD);
    __shared__ u64 smem[W];
    const LntParams P = Ps[blockIdx.x / wgs];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_row_body<B1, B2, G>(threadIdx.x, blockIdx.x % wgs, LView{smem, W}, P);
}

template <int B, int G>
__global__ __launch_bounds__(G * B * B) __attribute__((amdgpu_waves_per_eu(4))) void lnt_kb_row3(
    const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = G * B * B * B;
    __shared__ u64 smem[W];
    const LntParams P = Ps[blockIdx.x / wgs];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_row3_body<B, G>(threadIdx.x, blockIdx.x % wgs, LView{smem, W}, P);
}

template <int BA, int BB, int BC>
__global__ __launch_bounds__(BB * BC) __attribute__((amdgpu_waves_per_eu(BA == 16 ? 4 : 1))) void lnt_kb_row3m(
    const LntParams *__restrict__ Ps, u32 wgs) {
    constexpr u32 W = LntRow3m<BA, BB, BC>::LDS;
    __shared__ u64 smem[W];
    const LntParams P = Ps[blockIdx.x / wgs];
    if (P.n == 0) return;  // an empty slot of the batch
    lnt_row3m_body<BA, BB, BC>(threadIdx.x, blockIdx.x % wgs, LView{smem, W}, P);
}

// Batched launchers: `count` tests of plan pl, parameters in device memory.
bool lnt_launch_batch_small(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll, hipStream_t s);
bool lnt_launch_batch_col_fwd(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s);
bool lnt_launch_batch_col_inv(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s);
bool lnt_launch_batch_row(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s);

// Launchers (defined in lnt_k_small.hip / lnt_k_col.hip / lnt_k_row.hip).
// mode: LNT_SQUARE (iters squarings), LNT_FORWARD / LNT_MULTIPLY (one pass, iters = 1)
bool lnt_launch_small(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll,
                      hipStream_t s, int mode = LNT_SQUARE);
bool lnt_launch_col_fwd(const lnt::Plan &pl, const LntParams &P, bool sub2, hipStream_t s);
bool lnt_launch_col_inv(const lnt::Plan &pl, const LntParams &P, hipStream_t s);
bool lnt_launch_row(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode = LNT_SQUARE);

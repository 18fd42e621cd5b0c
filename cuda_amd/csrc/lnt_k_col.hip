// Instantiations of the column kernels (column lengths 64 .. 4096).
#include "lnt_launch.hpp"

template <int B1, int B2, int K>
static bool go_col_fwd(const lnt::Plan &pl, const LntParams &P, bool sub2, hipStream_t s) {
    dim3 grid(pl.C / K), block(K * B2);
    if (pl.M == 4)
        hipLaunchKernelGGL((lnt_k_col_fwd<B1, B2, K, 4>), grid, block, 0, s, P, (int)sub2);
    else
        hipLaunchKernelGGL((lnt_k_col_fwd<B1, B2, K, 8>), grid, block, 0, s, P, (int)sub2);
    return true;
}

template <int B1, int B2, int K>
static bool go_col_inv(const lnt::Plan &pl, const LntParams &P, hipStream_t s) {
    dim3 grid(pl.C / K), block(K * B2);
    hipLaunchKernelGGL((lnt_k_col_inv<B1, B2, K>), grid, block, 0, s, P);
    return true;
}

#define COL_SHAPES(X) \
    X(4, 4, 16) X(8, 4, 16) X(8, 8, 16) X(16, 8, 16) X(16, 16, 16) X(32, 16, 16) X(32, 32, 16) X(32, 32, 8) X(64, 32, 8) X(64, 32, 4) X(64, 64, 4)

// 3-phase column shapes (BA, BB, BC, K)
#define COL3_SHAPES(X) \
    X(8, 8, 8, 16) X(8, 8, 8, 8) X(16, 8, 8, 16) X(16, 8, 8, 8) X(16, 16, 8, 8) X(16, 16, 8, 4) X(16, 16, 16, 4)

template <int BA, int BB, int BC, int K>
static bool go_col3_fwd(const lnt::Plan &pl, const LntParams &P, bool sub2, hipStream_t s) {
    dim3 grid(pl.C / K), block(K * BB * BC);
    if (pl.M == 4)
        hipLaunchKernelGGL((lnt_k_col3_fwd<BA, BB, BC, K, 4>), grid, block, 0, s, P, (int)sub2);
    else
        hipLaunchKernelGGL((lnt_k_col3_fwd<BA, BB, BC, K, 8>), grid, block, 0, s, P, (int)sub2);
    return true;
}

bool lnt_launch_col_fwd(const lnt::Plan &pl, const LntParams &P, bool sub2, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.col3) {
#define C3F(ba, bb, bc, k) \
    if (pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) return go_col3_fwd<ba, bb, bc, k>(pl, P, sub2, s);
        COL3_SHAPES(C3F)
#undef C3F
        return false;
    }
#define CF(b1, b2, k) if (pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) return go_col_fwd<b1, b2, k>(pl, P, sub2, s);
    COL_SHAPES(CF)
#undef CF
    return false;
}

// lnt_k_col_ilp.hip (built with another machine scheduler)
bool lnt_launch_col3_inv_16x8x8_k16(const lnt::Plan &pl, const LntParams &P, hipStream_t s);

template <int BA, int BB, int BC, int K>
static bool go_col3_inv(const lnt::Plan &pl, const LntParams &P, hipStream_t s) {
    if constexpr (BA == 16 && BB == 8 && BC == 8 && K == 16) {
        return lnt_launch_col3_inv_16x8x8_k16(pl, P, s);  // not instantiated in this file
    } else {
        hipLaunchKernelGGL((lnt_k_col3_inv<BA, BB, BC, K>), dim3(pl.C / K), dim3(K * BB * BC), 0, s, P);
        return true;
    }
}

bool lnt_launch_col_inv(const lnt::Plan &pl, const LntParams &P, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.col3) {
#define C3I(ba, bb, bc, k) \
    if (pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) return go_col3_inv<ba, bb, bc, k>(pl, P, s);
        COL3_SHAPES(C3I)
#undef C3I
        return false;
    }
#define CI(b1, b2, k) if (pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) return go_col_inv<b1, b2, k>(pl, P, s);
    COL_SHAPES(CI)
#undef CI
    return false;
}

bool lnt_launch_batch_col_fwd(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.col3) {
#define C3F(ba, bb, bc, k)                                                                                 \
    if (pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) {                                      \
        const u32 wgs = pl.C / k;                                                                          \
        if (pl.M == 4)                                                                                     \
            hipLaunchKernelGGL((lnt_kb_col3_fwd<ba, bb, bc, k, 4>), dim3(wgs * count), dim3(k * bb * bc), 0, s, Ps, wgs); \
        else                                                                                               \
            hipLaunchKernelGGL((lnt_kb_col3_fwd<ba, bb, bc, k, 8>), dim3(wgs * count), dim3(k * bb * bc), 0, s, Ps, wgs); \
        return true;                                                                                       \
    }
        COL3_SHAPES(C3F)
#undef C3F
        return false;
    }
#define CF(b1, b2, k)                                                                                     \
    if (pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) {                                                \
        const u32 wgs = pl.C / k;                                                                         \
        if (pl.M == 4)                                                                                    \
            hipLaunchKernelGGL((lnt_kb_col_fwd<b1, b2, k, 4>), dim3(wgs * count), dim3(k * b2), 0, s, Ps, wgs); \
        else                                                                                              \
            hipLaunchKernelGGL((lnt_kb_col_fwd<b1, b2, k, 8>), dim3(wgs * count), dim3(k * b2), 0, s, Ps, wgs); \
        return true;                                                                                      \
    }
    COL_SHAPES(CF)
#undef CF
    return false;
}

bool lnt_launch_batch_col_inv(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.col3) {
#define C3I(ba, bb, bc, k)                                                                                \
    if (pl.col3 == ba && pl.cBB == bb && pl.cBC == bc && pl.K == k) {                                     \
        const u32 wgs = pl.C / k;                                                                         \
        hipLaunchKernelGGL((lnt_kb_col3_inv<ba, bb, bc, k>), dim3(wgs * count), dim3(k * bb * bc), 0, s, Ps, wgs); \
        return true;                                                                                      \
    }
        COL3_SHAPES(C3I)
#undef C3I
        return false;
    }
#define CI(b1, b2, k)                                                                                   \
    if (pl.col.B1 == b1 && pl.col.B2 == b2 && pl.K == k) {                                              \
        const u32 wgs = pl.C / k;                                                                       \
        hipLaunchKernelGGL((lnt_kb_col_inv<b1, b2, k>), dim3(wgs * count), dim3(k * b2), 0, s, Ps, wgs); \
        return true;                                                                                    \
    }
    COL_SHAPES(CI)
#undef CI
    return false;
}

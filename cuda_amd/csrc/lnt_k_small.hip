// Instantiations of the single-workgroup kernels for n = 8 .. 4096 (3-phase at 512 and 4096).
#include "lnt_launch.hpp"

template <int B1, int B2, int M>
static bool go_small_m(const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s, int mode) {
    dim3 grid(1), block(B2);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_small<B1, B2, M, LNT_FORWARD>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_small<B1, B2, M, LNT_MULTIPLY>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else
        hipLaunchKernelGGL((lnt_k_small<B1, B2, M, LNT_SQUARE>), grid, block, 0, s, P, iters, (int)sub2_first, (int)ll);
    return true;
}

template <int B1, int B2>
static bool go_small(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s,
                     int mode) {
    return pl.M == 4 ? go_small_m<B1, B2, 4>(P, iters, sub2_first, ll, s, mode)
                     : go_small_m<B1, B2, 8>(P, iters, sub2_first, ll, s, mode);
}

template <int B, int M>
static bool go_small3_m(const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s, int mode) {
    dim3 grid(1), block(B * B);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_small3<B, M, LNT_FORWARD>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_small3<B, M, LNT_MULTIPLY>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else
        hipLaunchKernelGGL((lnt_k_small3<B, M, LNT_SQUARE>), grid, block, 0, s, P, iters, (int)sub2_first, (int)ll);
    return true;
}

// lnt_k_small_ilp.hip (n = 512, built with another machine scheduler)
bool lnt_launch_small3_8(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll,
                         hipStream_t s, int mode);
bool lnt_launch_batch_small3_8(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll,
                               hipStream_t s);

template <int B>
static bool go_small3(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s,
                      int mode) {
    if constexpr (B == 8) {
        return lnt_launch_small3_8(pl, P, iters, sub2_first, ll, s, mode);  // not instantiated here
    } else {
        return pl.M == 4 ? go_small3_m<B, 4>(P, iters, sub2_first, ll, s, mode)
                         : go_small3_m<B, 8>(P, iters, sub2_first, ll, s, mode);
    }
}

bool lnt_launch_small(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s,
                      int mode) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.row3 == 16) return go_small3<16>(pl, P, iters, sub2_first, ll, s, mode);
    if (pl.row3 == 8) return go_small3<8>(pl, P, iters, sub2_first, ll, s, mode);
    if (pl.row3) return false;
#define SM(b1, b2) if (pl.row.B1 == b1 && pl.row.B2 == b2) return go_small<b1, b2>(pl, P, iters, sub2_first, ll, s, mode);
    SM(4, 2) SM(4, 4) SM(8, 4) SM(8, 8) SM(16, 8) SM(16, 16) SM(32, 16) SM(32, 32) SM(64, 32) SM(64, 64)
#undef SM
    return false;
}

template <int B1, int B2>
static bool go_batch_small(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll, hipStream_t s) {
    dim3 grid(count), block(B2);
    if (pl.M == 4)
        hipLaunchKernelGGL((lnt_kb_small<B1, B2, 4>), grid, block, 0, s, Ps, iters, (int)ll);
    else
        hipLaunchKernelGGL((lnt_kb_small<B1, B2, 8>), grid, block, 0, s, Ps, iters, (int)ll);
    return true;
}

template <int B>
static bool go_batch_small3(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll, hipStream_t s) {
    if constexpr (B == 8) {
        return lnt_launch_batch_small3_8(pl, Ps, count, iters, ll, s);  // not instantiated here
    } else {
        dim3 grid(count), block(B * B);
        if (pl.M == 4)
            hipLaunchKernelGGL((lnt_kb_small3<B, 4>), grid, block, 0, s, Ps, iters, (int)ll);
        else
            hipLaunchKernelGGL((lnt_kb_small3<B, 8>), grid, block, 0, s, Ps, iters, (int)ll);
        return true;
    }
}

bool lnt_launch_batch_small(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
    if (pl.row3 == 16) return go_batch_small3<16>(pl, Ps, count, iters, ll, s);
    if (pl.row3 == 8) return go_batch_small3<8>(pl, Ps, count, iters, ll, s);
    if (pl.row3) return false;
#define SM(b1, b2) if (pl.row.B1 == b1 && pl.row.B2 == b2) return go_batch_small<b1, b2>(pl, Ps, count, iters, ll, s);
    SM(4, 2) SM(4, 4) SM(8, 4) SM(8, 8) SM(16, 8) SM(16, 16) SM(32, 16) SM(32, 32) SM(64, 32) SM(64, 64)
#undef SM
    return false;
}

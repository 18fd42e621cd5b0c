// Instantiations of the row kernels (row lengths 128 .. 16384).
#include "lnt_launch.hpp"

template <int B1, int B2>
static bool go_row(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode) {
    constexpr int G = 64 / B2;
    dim3 grid(pl.R / G), block(G * B2);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_row<B1, B2, G, LNT_FORWARD>), grid, block, 0, s, P);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_row<B1, B2, G, LNT_MULTIPLY>), grid, block, 0, s, P);
    else
        hipLaunchKernelGGL((lnt_k_row<B1, B2, G, LNT_SQUARE>), grid, block, 0, s, P);
    return true;
}

template <int B, int G>
static bool go_row3(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode) {
    dim3 grid(pl.R / G), block(G * B * B);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_row3<B, G, LNT_FORWARD>), grid, block, 0, s, P);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_row3<B, G, LNT_MULTIPLY>), grid, block, 0, s, P);
    else
        hipLaunchKernelGGL((lnt_k_row3<B, G, LNT_SQUARE>), grid, block, 0, s, P);
    return true;
}

template <int BA, int BB, int BC>
static bool go_row3m(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode) {
    dim3 grid(pl.R), block(BB * BC);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_row3m<BA, BB, BC, LNT_FORWARD>), grid, block, 0, s, P);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_row3m<BA, BB, BC, LNT_MULTIPLY>), grid, block, 0, s, P);
    else
        hipLaunchKernelGGL((lnt_k_row3m<BA, BB, BC, LNT_SQUARE>), grid, block, 0, s, P);
    return true;
}

bool lnt_launch_row(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
#define R3M(ba, bb, bc) \
    if (pl.row3m == ba && pl.rBB == bb && pl.rBC == bc) return go_row3m<ba, bb, bc>(pl, P, s, mode);
    R3M(32, 16, 16) R3M(32, 32, 16) R3M(16, 32, 16) R3M(16, 32, 32)
#undef R3M
    if (pl.row3m) return false;
    if (pl.row3 == 16) return go_row3<16, 1>(pl, P, s, mode);
    if (pl.row3 == 8) return go_row3<8, 4>(pl, P, s, mode);
#define RW(b1, b2) if (pl.row.B1 == b1 && pl.row.B2 == b2) return go_row<b1, b2>(pl, P, s, mode);
    RW(16, 8) RW(16, 16) RW(32, 16) RW(32, 32) RW(64, 32) RW(64, 64)
#undef RW
    return false;
}

bool lnt_launch_batch_row(const lnt::Plan &pl, const LntParams *Ps, u32 count, hipStream_t s) {
    if (!lnt::plan_consistent(pl)) return false;  // a kernel of another shape would run out of bounds
#define R3M(ba, bb, bc)                                                                                   \
    if (pl.row3m == ba && pl.rBB == bb && pl.rBC == bc) {                                                  \
        hipLaunchKernelGGL((lnt_kb_row3m<ba, bb, bc>), dim3(pl.R * count), dim3(bb * bc), 0, s, Ps, pl.R); \
        return true;                                                                                       \
    }
    R3M(32, 16, 16) R3M(32, 32, 16) R3M(16, 32, 16) R3M(16, 32, 32)
#undef R3M
    if (pl.row3m) return false;
    if (pl.row3 == 16) {
        const u32 wgs = pl.R;
        hipLaunchKernelGGL((lnt_kb_row3<16, 1>), dim3(wgs * count), dim3(256), 0, s, Ps, wgs);
        return true;
    }
    if (pl.row3 == 8) {
        const u32 wgs = pl.R / 4;
        hipLaunchKernelGGL((lnt_kb_row3<8, 4>), dim3(wgs * count), dim3(256), 0, s, Ps, wgs);
        return true;
    }
#define RW(b1, b2)                                                                                   \
    if (pl.row.B1 == b1 && pl.row.B2 == b2) {                                                        \
        constexpr int G = 64 / b2;                                                                   \
        const u32 wgs = pl.R / G;                                                                    \
        hipLaunchKernelGGL((lnt_kb_row<b1, b2, G>), dim3(wgs * count), dim3(G * b2), 0, s, Ps, wgs); \
        return true;                                                                                 \
    }
    RW(16, 8) RW(16, 16) RW(32, 16) RW(32, 32) RW(64, 32) RW(64, 64)
#undef RW
    return false;
}

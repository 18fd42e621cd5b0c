// Instantiations of the row kernel (row lengths 128 .. 4096).
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

bool lnt_launch_row(const lnt::Plan &pl, const LntParams &P, hipStream_t s, int mode) {
#define RW(b1, b2) if (pl.row.B1 == b1 && pl.row.B2 == b2) return go_row<b1, b2>(pl, P, s, mode);
    RW(16, 8) RW(16, 16) RW(32, 16) RW(32, 32) RW(64, 32) RW(64, 64)
#undef RW
    return false;
}

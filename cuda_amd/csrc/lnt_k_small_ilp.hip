// The 512-word single-workgroup kernels (lnt_k_small3<8>: 64 threads of 8 words, plain and
// batched), compiled with the max-ilp machine scheduler (-amdgpu-sched-strategy=max-ilp, set
// for this file in the Makefile).  With one wave per workgroup nothing hides the wait states
// between a carry-out and its carry-in, and this scheduler fills them with independent work
// (1100 -> 76 wait-state cycles per thread): 9.3 -> 7.2 us per iteration at n = 512.  The
// other single-workgroup kernels are slower under it (n = 1024: 41 -> 46 us, n = 4096:
// 24.7 -> 30.3 us) and stay in lnt_k_small.hip.
#include "lnt_launch.hpp"

template <int M>
static void go_small3_8(const LntParams &P, u32 iters, bool sub2_first, bool ll, hipStream_t s, int mode) {
    dim3 grid(1), block(64);
    if (mode == LNT_FORWARD)
        hipLaunchKernelGGL((lnt_k_small3<8, M, LNT_FORWARD>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else if (mode == LNT_MULTIPLY)
        hipLaunchKernelGGL((lnt_k_small3<8, M, LNT_MULTIPLY>), grid, block, 0, s, P, 1u, (int)sub2_first, (int)ll);
    else
        hipLaunchKernelGGL((lnt_k_small3<8, M, LNT_SQUARE>), grid, block, 0, s, P, iters, (int)sub2_first, (int)ll);
}

bool lnt_launch_small3_8(const lnt::Plan &pl, const LntParams &P, u32 iters, bool sub2_first, bool ll,
                         hipStream_t s, int mode) {
    if (pl.row3 != 8) return false;
    if (pl.M == 4)
        go_small3_8<4>(P, iters, sub2_first, ll, s, mode);
    else
        go_small3_8<8>(P, iters, sub2_first, ll, s, mode);
    return true;
}

bool lnt_launch_batch_small3_8(const lnt::Plan &pl, const LntParams *Ps, u32 count, u32 iters, bool ll,
                               hipStream_t s) {
    if (pl.row3 != 8) return false;
    if (pl.M == 4)
        hipLaunchKernelGGL((lnt_kb_small3<8, 4>), dim3(count), dim3(64), 0, s, Ps, iters, (int)ll);
    else
        hipLaunchKernelGGL((lnt_kb_small3<8, 8>), dim3(count), dim3(64), 0, s, Ps, iters, (int)ll);
    return true;
}

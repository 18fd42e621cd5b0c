// The inverse column kernel of the 2^22 plan (lnt_k_col3_inv<16, 8, 8, 16>: 1024-word
// columns, 16 columns per workgroup), compiled with the max-ilp machine scheduler
// (-amdgpu-sched-strategy=max-ilp, set for this file in the Makefile).  The scheduler
// interleaves the carry-chain pairs of independent words so that the wait states between
// a carry-out and its carry-in are filled with work: 1.0 us faster for this shape at
// n = 2^22, while other column shapes spill or slow down under it (they stay in
// lnt_k_col.hip with the default scheduler).
#include "lnt_launch.hpp"

bool lnt_launch_col3_inv_16x8x8_k16(const lnt::Plan &pl, const LntParams &P, hipStream_t s) {
    if (!(pl.col3 == 16 && pl.cBB == 8 && pl.cBC == 8 && pl.K == 16)) return false;
    hipLaunchKernelGGL((lnt_k_col3_inv<16, 8, 8, 16>), dim3(pl.C / 16), dim3(16 * 8 * 8), 0, s, P);
    return true;
}

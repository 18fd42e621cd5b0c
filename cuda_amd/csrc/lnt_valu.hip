// VALU throughput of the field arithmetic on the device (lnt_valu_rates): each probe runs
// eight independent chains of one operation per thread over a grid that fills every CU,
// and reports field operations and VALU instructions per CU per clock.  The clock is
// measured inside the probe (shader clock counter against the 100 MHz real-time counter),
// so rates are per actual cycle, not per nominal one.
#include <hip/hip_runtime.h>

#include <cstdio>
#include <cstring>
#include <string>

#include "../../include/lnt_api.h"
#include "gl64.hpp"

namespace {

constexpr int kChains = 8;

template <int OP>
__device__ __forceinline__ u64 valu_step(u64 a, u64 b) {
    if (OP == 0) return gl_mul(a, b);
    if (OP == 1) return gl_sqr(a) ^ b;
    if (OP == 2) return gl_add(a, b);
    if (OP == 3) return gl_sub(a, b);
    if (OP == 4) return gl_mul_pow2(a, 12) ^ b;   // shift multiply, s <= 32
    if (OP == 5) return gl_mul_pow2(a, 48) ^ b;   // 32 < s < 64
    if (OP == 6) return gl_mul_pow2(a, 84) ^ b;   // 64 < s < 96
    if (OP == 7) return (u64)(u32)a * (u32)b + b;  // one v_mad_u64_u32 (+ the chain's add)
    return a;
}

template <int OP>
__global__ __launch_bounds__(256) void lnt_k_valu(u64 *out, u64 seed, int iters, unsigned long long *clk) {
    const u32 tid = blockIdx.x * blockDim.x + threadIdx.x;
    u64 v[kChains];
    u64 b = seed * 0x9E3779B97F4A7C15ull + tid;
#pragma unroll
    for (int j = 0; j < kChains; j++) v[j] = (seed ^ (tid * 0x1234567ull + j * 0x55555ull)) % GL_P;
    const unsigned long long t0 = clock64(), w0 = wall_clock64();
    for (int it = 0; it < iters; it++) {
#pragma unroll
        for (int j = 0; j < kChains; j++) v[j] = valu_step<OP>(v[j], b + j);
        asm volatile("" : "+v"(b));  // keeps the operand live per iteration
    }
    const unsigned long long t1 = clock64(), w1 = wall_clock64();
    u64 r = 0;
#pragma unroll
    for (int j = 0; j < kChains; j++) r ^= v[j];
    out[tid] = r;
    if (tid == 0) {
        clk[0] = t1 - t0;
        clk[1] = w1 - w0;
    }
}

const char *const kNames[] = {"gl_mul", "gl_sqr", "gl_add", "gl_sub", "gl_mul_pow2(12)", "gl_mul_pow2(48)",
                              "gl_mul_pow2(84)", "v_mad_u64_u32"};

template <int OP>
int probe(int cus, int blocks, int iters, u64 *d, unsigned long long *dclk, double *rate) {
    hipEvent_t e0, e1;
    if (hipEventCreate(&e0) != hipSuccess || hipEventCreate(&e1) != hipSuccess) return LNT_DEVICE_ERROR;
    hipLaunchKernelGGL(lnt_k_valu<OP>, dim3(blocks), dim3(256), 0, 0, d, 1, 4, dclk);  // warm-up
    (void)hipEventRecord(e0, 0);
    hipLaunchKernelGGL(lnt_k_valu<OP>, dim3(blocks), dim3(256), 0, 0, d, 7, iters, dclk);
    (void)hipEventRecord(e1, 0);
    int rc = hipEventSynchronize(e1) == hipSuccess ? LNT_OK : LNT_DEVICE_ERROR;
    float ms = 0;
    unsigned long long clk[2] = {0, 0};
    if (rc == LNT_OK && (hipEventElapsedTime(&ms, e0, e1) != hipSuccess ||
                         hipMemcpy(clk, dclk, sizeof clk, hipMemcpyDeviceToHost) != hipSuccess))
        rc = LNT_DEVICE_ERROR;
    (void)hipEventDestroy(e0);
    (void)hipEventDestroy(e1);
    if (rc != LNT_OK || ms <= 0 || !clk[1]) return LNT_DEVICE_ERROR;
    const double ghz = (double)clk[0] / ((double)clk[1] / 100e6) / 1e9;  // wall_clock64 runs at 100 MHz
    const double ops = (double)blocks * 256.0 * iters * kChains;
    rate[0] = ops / (ms * 1e-3) / ((double)cus * ghz * 1e9);  // field ops per CU per clock
    rate[1] = ghz;
    return LNT_OK;
}

}  // namespace

extern "C" int lnt_valu_rates(int device, double *ops_per_clk, char *names, size_t names_len, double *ghz,
                              int cap) {
    if (!ops_per_clk || cap <= 0) return -LNT_INVALID_ARGUMENT;
    if (hipSetDevice(device) != hipSuccess) return -LNT_DEVICE_ERROR;
    hipDeviceProp_t pr;
    if (hipGetDeviceProperties(&pr, device) != hipSuccess) return -LNT_DEVICE_ERROR;
    const int cus = pr.multiProcessorCount;
    const int blocks = cus * 16, iters = 2000;
    u64 *d = nullptr;
    unsigned long long *dclk = nullptr;
    if (hipMalloc(&d, (size_t)blocks * 256 * 8) != hipSuccess) return -LNT_OUT_OF_MEMORY;
    if (hipMalloc(&dclk, 16) != hipSuccess) {
        (void)hipFree(d);
        return -LNT_OUT_OF_MEMORY;
    }
    const int nops = (int)(sizeof kNames / sizeof kNames[0]);
    int count = 0, rc = LNT_OK;
    double r[2] = {0, 0}, clock_sum = 0;
    std::string nm;
    for (int op = 0; op < nops && count < cap && rc == LNT_OK; op++) {
        switch (op) {
            case 0: rc = probe<0>(cus, blocks, iters, d, dclk, r); break;
            case 1: rc = probe<1>(cus, blocks, iters, d, dclk, r); break;
            case 2: rc = probe<2>(cus, blocks, iters, d, dclk, r); break;
            case 3: rc = probe<3>(cus, blocks, iters, d, dclk, r); break;
            case 4: rc = probe<4>(cus, blocks, iters, d, dclk, r); break;
            case 5: rc = probe<5>(cus, blocks, iters, d, dclk, r); break;
            case 6: rc = probe<6>(cus, blocks, iters, d, dclk, r); break;
            default: rc = probe<7>(cus, blocks, iters, d, dclk, r); break;
        }
        if (rc != LNT_OK) break;
        ops_per_clk[count++] = r[0];
        clock_sum += r[1];
        nm += (nm.empty() ? "" : ";") + std::string(kNames[op]);
    }
    (void)hipFree(d);
    (void)hipFree(dclk);
    if (rc != LNT_OK) return -rc;
    if (names && names_len) std::snprintf(names, names_len, "%s", nm.c_str());
    if (ghz) *ghz = count ? clock_sum / count : 0;
    return count;
}

// Small DFTs held entirely in one thread's registers, for B = 1..64 points.
//
// omega_B = 2^(192/B) for B | 64 (the roots of unity of order <= 64 are powers of two
// in GF(2^64 - 2^32 + 1); they coincide with 554^((P-1)/B), the root family the
// reference uses, common/pre_calculation.cu:61-66).  Every twiddle inside these
// transforms is therefore a compile-time shift (gl_mul_pow2) once the loops unroll.
//
//  dft_fwd<B>: radix-2 decimation in frequency, natural order in, bit-reversed out:
//              v[k] <- X[brev_B(k)], X[k] = sum_j x[j] omega_B^(j k).
//  dft_inv<B>: radix-2 decimation in time, bit-reversed in, natural out, unscaled:
//              dft_inv<B>(dft_fwd<B>(x)) = B * x.
#pragma once
#include "gl64.hpp"

template <int B>
GL_HD void dft_fwd(u64 *v) {
#pragma unroll
    for (int len = B; len >= 2; len >>= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int s = 0; s < B; s += len) {
#pragma unroll
            for (int j = 0; j < half; j++) {
                u64 a = v[s + j], b = v[s + j + half];
                v[s + j] = gl_add(a, b);
                v[s + j + half] = gl_sub_pow2(a, b, (192 / len) * j);
            }
        }
    }
}

template <int B>
GL_HD void dft_inv(u64 *v) {
#pragma unroll
    for (int len = 2; len <= B; len <<= 1) {
        const int half = len >> 1;
#pragma unroll
        for (int s = 0; s < B; s += len) {
#pragma unroll
            for (int j = 0; j < half; j++) {
                u64 a = v[s + j], b = v[s + j + half];
                const int e = j == 0 ? 0 : 192 - (192 / len) * j;  // omega_len^-j
                if (e >= 96) {
                    u64 bb = gl_mul_pow2_lt96(b, e - 96);           // = -(b omega^-j)
                    v[s + j] = gl_sub(a, bb);
                    v[s + j + half] = gl_add(a, bb);
                } else {
                    u64 bb = gl_mul_pow2_lt96(b, e);
                    v[s + j] = gl_add(a, bb);
                    v[s + j + half] = gl_sub(a, bb);
                }
            }
        }
    }
}

// Bit reversal of v within log2(B) bits (compile-time B).
template <int B>
GL_HD constexpr int brev_c(int v) {
    int r = 0;
    for (int b = 1; b < B; b <<= 1) {
        r = (r << 1) | (v & 1);
        v >>= 1;
    }
    return r;
}

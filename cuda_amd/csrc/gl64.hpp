// Arithmetic in the Goldilocks field GF(P), P = 2^64 - 2^32 + 1, for gfx950 VALU
// (and the host, so the same code is unit-tested on the CPU).
//
// The field and its reduction identities are the reference's (common/arithmetics.cuh:26,
// 66-87: 2^64 = 2^32 - 1 and 2^96 = -1 mod P).  The code is written for CDNA:
//  * additions select between s and s + EPS from two carry bits (no compare against P),
//  * products are four 32x32->64 multiply-adds (v_mad_u64_u32) feeding one reduction,
//  * multiplication by a power of two uses shifts only (2 is a 192nd root of unity).
// Contracts: "canonical" means in [0, P).  Every function below returns a canonical value
// when its inputs are canonical; gl_mul / gl_mul_pow2 accept any 64-bit input.
#pragma once
#include <stdint.h>

#ifdef __HIPCC__
#include <hip/hip_runtime.h>
#define GL_HD __host__ __device__ __forceinline__
#define GL_HDM __host__ __device__ __forceinline__
#else
#define GL_HD static inline __attribute__((always_inline))
#define GL_HDM inline __attribute__((always_inline))
#endif

#define GL_P 0xFFFFFFFF00000001ull
#define GL_EPS 0x00000000FFFFFFFFull  // 2^64 - P = 2^32 - 1

typedef uint64_t u64;
typedef uint32_t u32;

// a + b mod P.  Requires a + b < 2^64 + P - 2^32 (true for a < 2^64, b <= P - 2^32,
// or a, b <= P).  The sum needs correcting iff it carried out of 64 bits or s + EPS does.
GL_HD u64 gl_add(u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    u32 c0, c1, c2, c3;
    const u32 sl = __builtin_addc((u32)a, (u32)b, 0u, &c0);
    const u32 sh = __builtin_addc((u32)(a >> 32), (u32)(b >> 32), c0, &c1);
    const u32 tl = __builtin_addc(sl, 0xFFFFFFFFu, 0u, &c2);
    const u32 th = __builtin_addc(sh, 0u, c2, &c3);
    return (c1 | c3) ? (((u64)th << 32) | tl) : (((u64)sh << 32) | sl);
#else
    u64 s = a + b;
    bool c1 = s < a;
    u64 t = s + GL_EPS;
    bool c2 = t < s;
    return (c1 | c2) ? t : s;
#endif
}

// a - b mod P for a < 2^64, b <= P: on borrow, a - b + 2^64 - EPS = a - b + P.
// On the device the borrow of the 32-bit subtract chain selects the correction directly
// (v_sub_co / v_subb, then subtract 0:EPS), one instruction fewer than a separate compare.
GL_HD u64 gl_sub(u64 a, u64 b) {
#if defined(__HIP_DEVICE_COMPILE__)
    u32 b0, b1, b2, b3;
    const u32 dl = __builtin_subc((u32)a, (u32)b, 0u, &b0);
    const u32 dh = __builtin_subc((u32)(a >> 32), (u32)(b >> 32), b0, &b1);
    const u32 rl = __builtin_subc(dl, b1 ? 0xFFFFFFFFu : 0u, 0u, &b2);
    const u32 rh = __builtin_subc(dh, 0u, b2, &b3);
    return ((u64)rh << 32) | rl;
#else
    u64 d = a - b;
    return (a < b) ? d - GL_EPS : d;
#endif
}

// -a mod P, canonical.
GL_HD u64 gl_neg(u64 a) { return a ? GL_P - a : 0; }

// Reduce the 128-bit value hi:lo (x3 x2 x1 x0 in 32-bit limbs):
// x = x1:x0 + x2 * (2^32 - 1) - x3 (mod P).  Result canonical for any hi, lo.
GL_HD u64 gl_reduce128(u64 hi, u64 lo) {
    u32 x2 = (u32)hi, x3 = (u32)(hi >> 32);
    u64 t0 = gl_sub(lo, x3);         // lo - x3 + P on borrow, cannot underflow
    u64 t1 = ((u64)x2 << 32) - x2;   // x2 * EPS <= P - 2^32
    return gl_add(t0, t1);
}

// 64x64 -> 128 product from four 32x32->64 products (v_mad_u64_u32 on CDNA).
GL_HD void gl_mul_wide(u64 a, u64 b, u64 &hi, u64 &lo) {
    u32 a0 = (u32)a, a1 = (u32)(a >> 32), b0 = (u32)b, b1 = (u32)(b >> 32);
    u64 p00 = (u64)a0 * b0;
    u64 t = (u64)a0 * b1 + (p00 >> 32);        // < 2^64
    u64 u = (u64)a1 * b0 + (u32)t;             // < 2^64
    u64 v = (u64)a1 * b1 + ((t >> 32) + (u >> 32));
    lo = (u << 32) | (u32)p00;
    hi = v;
}

GL_HD u64 gl_mul(u64 a, u64 b) {
    u64 hi, lo;
    gl_mul_wide(a, b, hi, lo);
    return gl_reduce128(hi, lo);
}

// a^2 mod P.  On the device the general product is the faster square: its four
// v_mad_u64_u32 cost less than doubling the cross term through carries (measured on gfx950:
// 3.42 against 2.84 squarings per CU per clock).
GL_HD u64 gl_sqr(u64 a) {
#if defined(__HIP_DEVICE_COMPILE__)
    return gl_mul(a, a);
#else
    u32 a0 = (u32)a, a1 = (u32)(a >> 32);
    u64 p00 = (u64)a0 * a0;
    u64 m = (u64)a0 * a1;                       // cross term, counted twice
    u64 t = (m << 1);                           // low 64 bits of 2m
    u32 tc = (u32)(m >> 63);                    // bit 64 of 2m
    u64 mid = (p00 >> 32) + (u32)t;             // < 2^33
    u64 lo = (mid << 32) | (u32)p00;
    u64 hi = (u64)a1 * a1 + (t >> 32) + ((u64)tc << 32) + (mid >> 32);
    return gl_reduce128(hi, lo);
#endif
}

// (db - 2^31) w mod P for a biased small integer db < 2^32, given wb = -2^31 w mod P:
// the 96-bit sum db w + wb from two 32x64 multiply-adds, then hi:lo = lo + hi (2^32 - 1)
// with hi < 2^32.  Half the work of gl_mul for the carried digits (|d| < 2^31) that the
// weights multiply.  Canonical output for canonical w, wb.
GL_HD u64 gl_mul_biased(u32 db, u64 w, u64 wb) {
    const u64 m0 = (u64)db * (u32)w + (u32)wb;                            // < 2^64
    const u64 m1 = (u64)db * (u32)(w >> 32) + (m0 >> 32) + (wb >> 32);    // < 2^64
    const u64 lo = (m1 << 32) | (u32)m0;
    const u32 x2 = (u32)(m1 >> 32);
    return gl_add(lo, ((u64)x2 << 32) - x2);  // x2 (2^32 - 1) <= P - 2^32
}

// x * 2^s mod P for 0 <= s < 96 (s is a compile-time constant after inlining in every
// caller, so the branches fold away).  Canonical output for any 64-bit x.
GL_HD u64 gl_mul_pow2_lt96(u64 x, int s) {
    if (s == 0) return x;
    if (s <= 32) {
        // x * 2^s = lo + h * 2^64, h < 2^32:  lo + h * EPS
        u64 lo = x << s;
        u64 h = x >> (64 - s);
        return gl_add(lo, (h << 32) - h);
    }
    if (s < 64) {
        // x = a + b 2^(64-s): x 2^s = (x << s) + b 2^64, b = b0 + b1 2^32:
        // b 2^64 = b0 * EPS - b1.
        u64 lo = x << s;
        u64 b = x >> (64 - s);
        u64 b0 = (u32)b, b1 = b >> 32;
        return gl_sub(gl_add(lo, (b0 << 32) - b0), b1);
    }
    if (s == 64) {
        // x * EPS = x1:x0 * (2^32 - 1) = reduce(hi = x, lo = 0) - ... use the 128-bit path
        return gl_reduce128(x, 0);
    }
    // 64 < s < 96: x = m + q 2^(96-s), x 2^s = m 2^s + q 2^96 = (m 2^(s-64)) * EPS - q
    u64 h = (u32)(x << (s - 64));
    u64 q = x >> (96 - s);
    return gl_sub((h << 32) - h, q);
}

// x * 2^s mod P for 0 <= s < 192.
GL_HD u64 gl_mul_pow2(u64 x, int s) {
    if (s >= 96) return gl_neg(gl_mul_pow2_lt96(x, s - 96));
    return gl_mul_pow2_lt96(x, s);
}

// Multiply (a - b) by 2^s, 0 <= s < 192, without an explicit negation:
// 2^96 = -1, so (a - b) 2^s = (b - a) 2^(s - 96).
GL_HD u64 gl_sub_pow2(u64 a, u64 b, int s) {
    if (s >= 96) return gl_mul_pow2_lt96(gl_sub(b, a), s - 96);
    return gl_mul_pow2_lt96(gl_sub(a, b), s);
}

GL_HD u64 gl_pow(u64 a, u64 e) {
    u64 r = 1;
    while (e) {
        if (e & 1) r = gl_mul(r, a);
        e >>= 1;
        a = gl_mul(a, a);
    }
    return r;
}

GL_HD u64 gl_inv(u64 a) { return gl_pow(a, GL_P - 2); }

// Signed integer <-> field element.  |v| < P/2 is assumed.
GL_HD u64 gl_from_i64(int64_t v) { return v < 0 ? (u64)v + GL_P : (u64)v; }
GL_HD int64_t gl_to_i64(u64 e) { return e > (GL_P >> 1) ? (int64_t)(e - GL_P) : (int64_t)e; }

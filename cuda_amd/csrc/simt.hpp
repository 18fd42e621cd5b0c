// Thin execution-model layer so each kernel body is written once and runs either as a
// HIP kernel on gfx950 or, for index/bounds verification, under a CPU emulation where
// the threads of one workgroup are OpenMP threads and the workgroup barrier is an
// OpenMP barrier.  In the emulation every global and LDS index is bounds-checked.
#pragma once
#include "gl64.hpp"

#if defined(LNT_EMU)
#include <omp.h>
#include <cstdio>
#include <cstdlib>
#define LNT_SYNC() _Pragma("omp barrier")
#define LNT_CHECK(cond, what)                                                      \
    do {                                                                           \
        if (!(cond)) {                                                             \
            std::fprintf(stderr, "LNT_EMU bounds: %s (%s:%d)\n", what, __FILE__,   \
                         __LINE__);                                                \
            std::abort();                                                          \
        }                                                                          \
    } while (0)
#define LNT_KFN static inline
#define LNT_OPAQUE(ptr) asm volatile("" : "+r"(ptr))
#else
#define LNT_SYNC() __syncthreads()
#define LNT_CHECK(cond, what) ((void)0)
#define LNT_KFN __device__ __forceinline__
// Hide a (uniform) pointer's value from the optimiser so loads through it are not
// hoisted out of a loop (which would pin whole tables in registers).
#define LNT_OPAQUE(ptr) asm volatile("" : "+s"(ptr))
#endif

// Global-memory and LDS views carrying their length so the emulation can check.
// On the device every access goes through a global-address-space pointer: the batched
// kernels read their views from a parameter array in memory, where the compiler cannot
// infer the address space and would otherwise emit flat accesses with a 64-bit VGPR
// address per element (twice the registers of the saddr + offset global form).
template <class T>
struct GView {
    T *p;
    u64 n;
#if defined(__HIP_DEVICE_COMPILE__)
    __device__ __forceinline__ T &operator[](u64 i) const {
        typedef __attribute__((address_space(1))) T GT;
        return *(T *)((GT *)p + i);
    }
#else
    GL_HDM T &operator[](u64 i) const {
        LNT_CHECK(i < n, "global index");
        return p[i];
    }
#endif
};

struct LView {
    u64 *p;
    u32 n;
    GL_HDM u64 &operator[](u32 i) const {
        LNT_CHECK(i < n, "lds index");
        return p[i];
    }
};

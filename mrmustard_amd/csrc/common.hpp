// Shared device helpers for the Bargmann->Fock kernels (gfx950, wave64).
//
// Arithmetic follows the reference recurrence (core.py:97-104) operation by
// operation: every complex product is (ac-bd, ad+bc) with no fused
// multiply-add (the library is compiled with -ffp-contract=off), and the
// division by sqrt(k) is correctly rounded.  Only the VJP reductions, whose
// summation order cannot match a sequential CPU loop anyway, use explicit fma.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

#include "../../include/mm_hermite.h"

#define MM_MAXD 8
#define MM_WAVE 64

namespace mm {

// ---------------------------------------------------------------- complex ops

__device__ __forceinline__ double2 cadd(double2 a, double2 b) { return make_double2(a.x + b.x, a.y + b.y); }

// (a.x + i a.y)(b.x + i b.y) in the reference's operation order.
__device__ __forceinline__ double2 cmul(double2 a, double2 b) {
    return make_double2(a.x * b.x - a.y * b.y, a.x * b.y + a.y * b.x);
}

// complex * real where the real operand is first promoted to complex(s, 0.0)
// (numba's rule for complex*float): the product's zero terms are kept so that
// signed zeros and non-finite values behave like the reference.
__device__ __forceinline__ double2 cscale_promoted(double2 a, double s) {
    return make_double2(a.x * s - a.y * 0.0, a.x * 0.0 + a.y * s);
}

// Fused complex multiply-accumulate acc += a*b (VJP reductions only).
__device__ __forceinline__ double2 cfma(double2 a, double2 b, double2 acc) {
    acc.x = fma(a.x, b.x, acc.x);
    acc.x = fma(-a.y, b.y, acc.x);
    acc.y = fma(a.x, b.y, acc.y);
    acc.y = fma(a.y, b.x, acc.y);
    return acc;
}

// acc += w * z for real w (VJP reductions only).
__device__ __forceinline__ double2 caxpy(double w, double2 z, double2 acc) {
    return make_double2(fma(w, z.x, acc.x), fma(w, z.y, acc.y));
}

// Correctly rounded x / s given r = RN(1/s): Markstein's correction
// q = RN(x r); q' = RN(q + RN(x - s q) r)  (P. Markstein, IBM J. Res. Dev.
// 34(1), 1990; the theorem needs r correctly rounded and no underflow of the
// remainder, so tiny numerators take the plain IEEE division instead).
__device__ __forceinline__ double div_by(double x, double s, double r) {
    double q = x * r;
    double e = fma(-q, s, x);
    double qq = fma(e, r, q);
    if (__builtin_expect(fabs(x) < 0x1p-960, 0)) qq = x / s;
    return qq;
}

__device__ __forceinline__ double2 cdiv_by(double2 v, double s, double r) {
    return make_double2(div_by(v.x, s, r), div_by(v.y, s, r));
}

// CPython's _Py_c_quot (numba's complex division), used for dL/dc = sum/c.
__device__ __forceinline__ double2 cdiv_py(double2 a, double2 b) {
    double2 z;
    if (fabs(b.x) >= fabs(b.y)) {
        if (b.x == 0.0 && b.y == 0.0) return make_double2(__builtin_nan(""), __builtin_nan(""));
        double ratio = b.y / b.x;
        double denom = b.x + b.y * ratio;
        z.x = (a.x + a.y * ratio) / denom;
        z.y = (a.y - a.x * ratio) / denom;
    } else {
        double ratio = b.x / b.y;
        double denom = b.x * ratio + b.y;
        z.x = (a.x * ratio + a.y) / denom;
        z.y = (a.y * ratio - a.x) / denom;
    }
    return z;
}

// ------------------------------------------------------------ wave helpers

__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

__device__ __forceinline__ double shfl_xor_d(double v, int mask) {
    int lo = __double2loint(v), hi = __double2hiint(v);
    lo = __shfl_xor(lo, mask, MM_WAVE);
    hi = __shfl_xor(hi, mask, MM_WAVE);
    return __hiloint2double(hi, lo);
}

__device__ __forceinline__ double2 wave_sum(double2 v) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        v.x += shfl_xor_d(v.x, m);
        v.y += shfl_xor_d(v.y, m);
    }
    return v;
}

}  // namespace mm

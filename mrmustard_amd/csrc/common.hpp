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

// c ? a : b, one component at a time (a select of the whole struct can be
// lowered through private memory).
__device__ __forceinline__ double2 csel(bool c, double2 a, double2 b) {
    return make_double2(c ? a.x : b.x, c ? a.y : b.y);
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

// Numerators that Markstein's correction below cannot take: tiny non-zero
// ones (the remainder would underflow) and non-finite ones (x - s q is then
// inf - inf).
__device__ __forceinline__ bool div_rare(double x) {
    const double ax = fabs(x);
    return ((ax < 0x1p-960) && (x != 0.0)) || !(ax <= 0x1.fffffffffffffp+1023);
}

// Correctly rounded x / s given r = RN(1/s): Markstein's correction
// q = RN(x r); q' = RN(q + RN(x - s q) r)  (P. Markstein, IBM J. Res. Dev.
// 34(1), 1990; the theorem needs r correctly rounded and no underflow of the
// remainder).  s = sqrt(k) >= 1, so a finite numerator never overflows.
__device__ __forceinline__ double div_by(double x, double s, double r) {
    double q = x * r;
    double e = fma(-q, s, x);
    double qq = fma(e, r, q);
    if (__builtin_expect(div_rare(x), 0)) qq = x / s;
    return qq;
}

// v / s as numba divides a complex by a float: s is promoted to complex(s, 0)
// and CPython's quotient gives ((x + y*0) / s, (y - x*0) / s).  For finite
// parts that is x / s and y / s; an infinite part makes the other one NaN.
__device__ __forceinline__ double2 cdiv_promoted(double2 v, double s) {
    return make_double2((v.x + v.y * 0.0) / s, (v.y - v.x * 0.0) / s);
}

__device__ __forceinline__ double2 cdiv_by(double2 v, double s, double r) {
    double2 q = make_double2(div_by(v.x, s, r), div_by(v.y, s, r));
    if (__builtin_expect(div_rare(v.x) || div_rare(v.y), 0)) q = cdiv_promoted(v, s);
    return q;
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

// ------------------------------------------------------------ XCD placement
// Workgroups of a 1-D grid are dispatched to the 8 XCDs round-robin, so
// block b runs on XCD b % 8, and each XCD has its own L2.  This maps block b
// to a logical index such that each XCD holds a contiguous range of them: work
// split by logical index then keeps neighbouring chunks (which read the same
// data) on one L2.  Any grid size works; the mapping is a permutation and
// only placement, not results, depends on it.
constexpr int MM_XCDS = 8;

__device__ __forceinline__ int64_t xcd_block(int64_t b, int64_t n) {
    const int64_t per = n / MM_XCDS, full = per * MM_XCDS;
    if (b >= full) return b;  // the remainder keeps its own order
    return (b % MM_XCDS) * per + b / MM_XCDS;
}

// The points q0, q0 + inc, ... below hi of a slab of S points that block blk
// of nblk (256 threads each) visits.  With nblk a multiple of the XCD count,
// the blocks of one XCD (consecutive logical indices, xcd_block) share one
// contiguous eighth of the slab and walk it in order, so the points that read
// the same lines of the slabs below them do so through one L2; otherwise
// every block strides over the whole slab.
struct Span {
    int64_t q0, inc, hi;
};

__device__ __forceinline__ Span xcd_span(int64_t blk, int64_t nblk, int64_t S) {
    Span s;
    if (nblk >= MM_XCDS && nblk % MM_XCDS == 0) {
        const int64_t per = nblk / MM_XCDS, x = blk / per, j = blk - x * per;
        const int64_t chunk = ((S + MM_XCDS - 1) / MM_XCDS + 255) / 256 * 256;
        const int64_t lo = x * chunk;
        s.hi = lo + chunk < S ? lo + chunk : S;
        s.q0 = lo + j * 256 + threadIdx.x;
        s.inc = per * 256;
    } else {
        s.q0 = blk * 256 + threadIdx.x;
        s.inc = nblk * 256;
        s.hi = S;
    }
    return s;
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

// DPP move of a double (two 32-bit halves); rows outside ROW_MASK read 0.
template <int CTRL, int ROW_MASK>
__device__ __forceinline__ double dpp_d(double v) {
    const int lo = __builtin_amdgcn_update_dpp(0, __double2loint(v), CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp(0, __double2hiint(v), CTRL, ROW_MASK, 0xF, false);
    return __hiloint2double(hi, lo);
}

// Sums of N values over the wave with DPP lane moves (no LDS traffic):
// pairs, quads, half rows, rows, then rows 0+1 and all four rows through the
// row broadcasts.  Lane 63 holds the totals; the order is fixed, so the
// result is deterministic.
template <int N>
__device__ __forceinline__ void wave_sum_dpp(double2 (&v)[N]) {
#define MM_DPP_STEP(CTRL, MASK)                          \
    _Pragma("unroll") for (int e = 0; e < N; ++e) {      \
        v[e].x += dpp_d<CTRL, MASK>(v[e].x);             \
        v[e].y += dpp_d<CTRL, MASK>(v[e].y);             \
    }
    MM_DPP_STEP(0xB1, 0xF)   // quad_perm [1,0,3,2]
    MM_DPP_STEP(0x4E, 0xF)   // quad_perm [2,3,0,1]
    MM_DPP_STEP(0x141, 0xF)  // row_half_mirror
    MM_DPP_STEP(0x140, 0xF)  // row_mirror
    MM_DPP_STEP(0x142, 0xA)  // row_bcast:15 into rows 1 and 3
    MM_DPP_STEP(0x143, 0xC)  // row_bcast:31 into rows 2 and 3
#undef MM_DPP_STEP
}

// Butterfly sums of N values over the wave, stage by stage so that the N
// shuffles of a stage are in flight together.  Lane 0 holds the totals.
template <int N>
__device__ __forceinline__ void wave_sum_many(double2 (&v)[N]) {
#pragma unroll
    for (int m = 32; m >= 1; m >>= 1) {
        double tx[N], ty[N];
#pragma unroll
        for (int e = 0; e < N; ++e) {
            tx[e] = shfl_xor_d(v[e].x, m);
            ty[e] = shfl_xor_d(v[e].y, m);
        }
#pragma unroll
        for (int e = 0; e < N; ++e) {
            v[e].x += tx[e];
            v[e].y += ty[e];
        }
    }
}

// cdiv_by on N values with a single branch for the rare tiny numerators.
template <int N>
__device__ __forceinline__ void cdiv_many(double2 (&v)[N], double s, double r) {
    double2 out[N];
    bool tiny = false;
#pragma unroll
    for (int e = 0; e < N; ++e) {
        double qx = v[e].x * r, qy = v[e].y * r;
        const double ex = fma(-qx, s, v[e].x), ey = fma(-qy, s, v[e].y);
        out[e].x = fma(ex, r, qx);
        out[e].y = fma(ey, r, qy);
        // exact zeros come out right (up to the sign of zero) and need no fix-up
        tiny = tiny || div_rare(v[e].x) || div_rare(v[e].y);
    }
    if (__builtin_expect(tiny, 0)) {
#pragma unroll
        for (int e = 0; e < N; ++e)
            if (div_rare(v[e].x) || div_rare(v[e].y)) out[e] = cdiv_promoted(v[e], s);
    }
#pragma unroll
    for (int e = 0; e < N; ++e) v[e] = out[e];
}

}  // namespace mm

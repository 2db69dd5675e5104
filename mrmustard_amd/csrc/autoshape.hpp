// Fock cutoffs of Gaussian density matrices: autoshape_numba
// (mrmustard/math/lattice/autoshape.py:23-154) for a batch of states.
//
// Two launches.  One wave per (state, mode m) forms the reduced state of mode
// m, the Schur complement (autoshape.py:84-108):
//   R   = A_nn - X,  X = [[0, I], [I, 0]] over the other M-1 modes,
//   A_  = A_mm - A_mn R^-1 A_nm,  b_ = b_m - A_mn R^-1 b_n,
//   c_  = c exp(-b_n R^-1 b_n / 2) / sqrt(det R).
// The wave solves R Y = [A_nm | b_n] by Gauss-Jordan elimination with partial
// pivoting in LDS (det R from the pivots).  Then one lane per (state, mode)
// runs the reference's early-stopped diagonal recursion (autoshape.py:110-153),
// a dependent chain of up to max_shape steps, and writes the step count
// clipped to [min_shape, max_shape] (autoshape.py:154).
//
// Index layout of the 2M-dimensional A and b: entry s*M + i is (bra if s = 0,
// ket if s = 1) of mode i, as the reference's reshape (2, M, 2, M) reads it.
#pragma once

#include "common.hpp"

namespace mm {

constexpr int AS_MAXM = 8;                 // modes per state
constexpr int AS_MAXN = 2 * AS_MAXM - 2;  // order of R

struct AutoshapeArgs {
    const double2 *A, *b, *c;  // B x (2M x 2M), B x 2M, B
    int64_t B;
    int M;
    double max_prob;
    int max_shape, min_shape;
    const double *sq;  // sqrt(n), n < 100000
    int64_t *shape;    // B x M
    double2 *marg;     // B x M x 8: the single-mode (A_00, A_01, A_10, A_11, b_0, b_1, c_) and a flag
};

__device__ __forceinline__ double2 csub(double2 a, double2 b) { return make_double2(a.x - b.x, a.y - b.y); }

// principal square root (branch cut on the negative real axis, as np.sqrt)
__device__ __forceinline__ double2 csqrt_principal(double2 z) {
    if (z.x == 0.0 && z.y == 0.0) return make_double2(0.0, z.y);
    const double r = hypot(z.x, z.y);
    if (z.x >= 0.0) {
        const double t = sqrt(0.5 * (r + z.x));
        return make_double2(t, z.y / (2.0 * t));
    }
    const double t = sqrt(0.5 * (r - z.x));
    return make_double2(fabs(z.y) / (2.0 * t), copysign(t, z.y));
}

__device__ __forceinline__ double2 cexp_(double2 z) {
    const double e = exp(z.x);
    return make_double2(e * cos(z.y), e * sin(z.y));
}

// Wave per (state, mode): the Schur complement, written to marg for the recursion.
__global__ __launch_bounds__(64) void autoshape_schur_kernel(AutoshapeArgs a) {
    __shared__ double2 W[AS_MAXN][AS_MAXN + 3];
    __shared__ double2 Amn[2][AS_MAXN];
    __shared__ int piv_row;
    const int lane = threadIdx.x;
    const int M = a.M, n2 = 2 * M - 2, ncol = n2 + 3, twoM = 2 * M;
    const int64_t st = blockIdx.x / M;
    const int m = int(blockIdx.x % M);
    if (st >= a.B) return;
    const double2 *A = a.A + st * twoM * twoM;
    const double2 *b = a.b + st * twoM;
    auto other = [&](int i) { return i < m ? i : i + 1; };  // i-th mode other than m
    auto ridx = [&](int r) { return (r / (M - 1)) * M + other(r % (M - 1)); };  // R's row r in A's index

    // W = [A_nn - X | A_nm | b_n]; Amn = A_mn
    for (int e = lane; e < n2 * ncol; e += MM_WAVE) {
        const int r = e / ncol, col = e % ncol;
        double2 v;
        if (col < n2) {
            v = A[ridx(r) * twoM + ridx(col)];
            if ((r / (M - 1)) != (col / (M - 1)) && (r % (M - 1)) == (col % (M - 1))) v.x -= 1.0;
        } else if (col < n2 + 2) {
            v = A[((col - n2) * M + m) * twoM + ridx(r)];  // A_nm[r][u] = A_mn[u][r]
        } else {
            v = b[ridx(r)];
        }
        W[r][col] = v;
    }
    for (int e = lane; e < 2 * n2; e += MM_WAVE) Amn[e / n2][e % n2] = A[((e / n2) * M + m) * twoM + ridx(e % n2)];
    wave_sync();

    // Gauss-Jordan with partial pivoting; det R = product of pivots x sign
    double2 det = make_double2(1.0, 0.0);
    for (int p = 0; p < n2; ++p) {
        if (lane == 0) {
            int best = p;
            double bv = hypot(W[p][p].x, W[p][p].y);
            for (int r = p + 1; r < n2; ++r) {
                const double v = hypot(W[r][p].x, W[r][p].y);
                if (v > bv) {
                    bv = v;
                    best = r;
                }
            }
            piv_row = best;
        }
        wave_sync();
        const int pr = piv_row;
        if (pr != p) {
            for (int col = lane; col < ncol; col += MM_WAVE) {
                const double2 t = W[p][col];
                W[p][col] = W[pr][col];
                W[pr][col] = t;
            }
            det = make_double2(-det.x, -det.y);
        }
        wave_sync();
        const double2 pv = W[p][p];
        det = cmul(det, pv);
        wave_sync();
        for (int col = lane; col < ncol; col += MM_WAVE) W[p][col] = cdiv_py(W[p][col], pv);
        wave_sync();
        for (int e = lane; e < n2 * ncol; e += MM_WAVE) {
            const int r = e / ncol, col = e % ncol;
            if (r == p || col == p) continue;
            W[r][col] = csub(W[r][col], cmul(W[r][p], W[p][col]));
        }
        wave_sync();
        if (lane < n2 && lane != p) W[lane][p] = make_double2(0.0, 0.0);
        wave_sync();
    }
    if (lane != 0) return;

    // single-mode (A_, b_, c_): Y = W[:, n2 .. n2+2]
    double2 Am[2][2], bm[2];
    for (int u = 0; u < 2; ++u) {
        for (int v = 0; v < 2; ++v) {
            double2 s = make_double2(0.0, 0.0);
            for (int r = 0; r < n2; ++r) s = cadd(s, cmul(Amn[u][r], W[r][n2 + v]));
            Am[u][v] = csub(A[(u * M + m) * twoM + (v * M + m)], s);
        }
        double2 s = make_double2(0.0, 0.0);
        for (int r = 0; r < n2; ++r) s = cadd(s, cmul(Amn[u][r], W[r][n2 + 2]));
        bm[u] = csub(b[u * M + m], s);
    }
    double2 qf = make_double2(0.0, 0.0);
    for (int r = 0; r < n2; ++r) qf = cadd(qf, cmul(b[ridx(r)], W[r][n2 + 2]));
    const double2 ex = cexp_(make_double2(-0.5 * qf.x, -0.5 * qf.y));
    const double2 cm = cdiv_py(cmul(a.c[st], ex), csqrt_principal(det));
    double2 *o = a.marg + (st * M + m) * 8;
    o[0] = Am[0][0];
    o[1] = Am[0][1];
    o[2] = Am[1][0];
    o[3] = Am[1][1];
    o[4] = bm[0];
    o[5] = bm[1];
    o[6] = cm;
    // a singular R (np.linalg.inv raises there) is reported as -1
    o[7] = make_double2((isfinite(cm.x) && isfinite(cm.y)) ? 0.0 : 1.0, 0.0);
}

// Lane per (state, mode): the reference's early-stopped diagonal recursion
// (autoshape.py:110-153) on the single-mode marginal, and the step count
// clipped to [min_shape, max_shape] (autoshape.py:154).
__global__ __launch_bounds__(256) void autoshape_recursion_kernel(AutoshapeArgs a) {
    const int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (q >= a.B * a.M) return;
    const double2 *o = a.marg + q * 8;
    double2 Am[2][2], bm[2];
    Am[0][0] = o[0];
    Am[0][1] = o[1];
    Am[1][0] = o[2];
    Am[1][1] = o[3];
    bm[0] = o[4];
    bm[1] = o[5];
    const double2 cm = o[6];
    const bool bad = o[7].x != 0.0;
    // buffers transposed with respect to the reference's diagram
    const double *SQ = a.sq;
    double2 buf2[2][2], buf3[2][3];
    for (int i = 0; i < 2; ++i) {
        buf2[i][0] = buf2[i][1] = make_double2(0.0, 0.0);
        buf3[i][0] = buf3[i][1] = buf3[i][2] = make_double2(0.0, 0.0);
    }
    buf3[0][1] = cm;
    double norm = hypot(cm.x, cm.y);
    int k = 0;
    while (!bad && norm < a.max_prob && k < a.max_shape) {
        const int o2 = k & 1, n = o2 ^ 1;
        const double s0 = SQ[k], s1 = SQ[k + 1], s2 = SQ[k + 2];
        for (int u = 0; u < 2; ++u) {
            const double2 av = cadd(cmul(Am[u][0], buf2[o2][0]), cmul(Am[u][1], buf2[o2][1]));
            const double2 t = cadd(cmul(bm[u], buf3[o2][1]), make_double2(av.x * s0, av.y * s0));
            buf2[n][u] = make_double2(t.x / s1, t.y / s1);
        }
        double2 t0 = cadd(cadd(cmul(bm[0], buf2[n][0]), cmul(make_double2(Am[0][0].x * s1, Am[0][0].y * s1), buf3[o2][1])),
                          cmul(make_double2(Am[0][1].x * s0, Am[0][1].y * s0), buf3[o2][0]));
        double2 t1 = cadd(cadd(cmul(bm[1], buf2[n][0]), cmul(make_double2(Am[1][0].x * s1, Am[1][0].y * s1), buf3[o2][1])),
                          cmul(make_double2(Am[1][1].x * s0, Am[1][1].y * s0), buf3[o2][0]));
        double2 t2 = cadd(cadd(cmul(bm[1], buf2[n][1]), cmul(make_double2(Am[1][0].x * s0, Am[1][0].y * s0), buf3[o2][2])),
                          cmul(make_double2(Am[1][1].x * s1, Am[1][1].y * s1), buf3[o2][1]));
        buf3[n][0] = make_double2(t0.x / s2, t0.y / s2);
        buf3[n][1] = make_double2(t1.x / s1, t1.y / s1);
        buf3[n][2] = make_double2(t2.x / s2, t2.y / s2);
        norm += hypot(buf3[n][1].x, buf3[n][1].y);
        ++k;
    }
    const int kc = k < a.min_shape ? a.min_shape : (k > a.max_shape ? a.max_shape : k);
    a.shape[q] = bad ? -1 : kc;
}

}  // namespace mm

// The diagonal strategy: photon-number amplitudes G[a,a,b,b,c,c,...] of a
// 2M-index Bargmann triple without the full 2M-dimensional lattice
// (mrmustard/math/lattice/strategies/compactFock/diagonal_amps.py:17-248).
//
// The reference keeps five sub-lattices indexed by params p = (a, b, c, ...):
//   arr0[p]           G[p ⊗ p]                        (the output)
//   arr1[i, p]        G[p ⊗ p + e_i]                  (one index raised)
//   arr2[d, p]        G[p ⊗ p + 2 e_2d]
//   arr1010[d, k, p]  G[p ⊗ p + e_2d + e_2i]          (i = d + 1 + k)
//   arr1001[d, k, p]  G[p ⊗ p + e_2d + e_2i+1]
// and visits params in order of weight |p|, each applying a "diagonal pivot"
// (writes arr1[:, p], diagonal_amps.py:108-150) and, for every d with
// p[:d] == 0, an "off-diagonal pivot" (writes arr0[p + e_d], arr2[d, p],
// arr1010/arr1001[d, :, p], diagonal_amps.py:19-105).
//
// Every value written at weight w is read by exactly one params of weight
// w + 1 and nothing of weight w reads another params' weight-w values, so one
// launch per weight runs all params of that weight in parallel, a thread per
// (params, batch column).  A thread finds its params from its rank within the
// weight by unranking against a count table T[j][s] = #{(p_j..p_{M-1}) : p_k <
// cutoff_k, sum = s}, in np.ndindex order (helperFunctions.py:31-46).
//
// Layout: arrays are [slice][flat p][batch column], batch innermost.  Slices
// of arr2 / arr1010 / arr1001 for d are only ever touched at params with
// p[:d] == 0, whose flat index is below prod(cutoffs[d:]), so slice d is
// stored with that length: the workspace is about (2M + 2(M - 1) + 1)
// lattices instead of the reference's (2M^2 + M).
#pragma once

#include "common.hpp"

namespace mm {

constexpr int DG_MAXM = 10;

struct DiagArgs {
    const double2 *A, *B, *c;  // A: 2M x 2M; B: 2M x nb; c: 1 or nb values
    int c_batched;
    double2 *arr0, *arr1, *arr2, *a1010, *a1001;
    const int64_t *T;  // (M + 1) x (S + 1) counts
    const double *sq, *rsq;
    int64_t P, nb;
    int S;
    int cut[DG_MAXM];
    int64_t stride[DG_MAXM];
    int64_t len[DG_MAXM];   // prod(cut[d:])
    int64_t off2[DG_MAXM];  // start of arr2's slice d (entries)
    int64_t off11[DG_MAXM]; // start of arr1010/arr1001's block d (M - d - 1 slices of len[d])
};

// Count table: T[M][s] = (s == 0); T[j][s] = sum_{v < cut_j, v <= s} T[j+1][s-v]
// = pre[s] - pre[s - cut_j] with pre the inclusive prefix sum of row j+1.
// One block; the prefix sum is chunked per thread (pre: W int64 of scratch).
__global__ __launch_bounds__(1024) void diag_counts_kernel(int64_t *T, int64_t *pre, DiagArgs a, int M) {
    __shared__ int64_t part[1024];
    const int W = a.S + 1, nt = blockDim.x, tid = threadIdx.x;
    const int chunk = (W + nt - 1) / nt;
    const int lo = min(W, tid * chunk), hi = min(W, lo + chunk);
    for (int s = tid; s < W; s += nt) T[int64_t(M) * W + s] = (s == 0);
    for (int j = M - 1; j >= 0; --j) {
        __syncthreads();
        const int64_t *row = T + int64_t(j + 1) * W;
        int64_t acc = 0;
        for (int s = lo; s < hi; ++s) acc += row[s];
        part[tid] = acc;
        __syncthreads();
        if (tid == 0) {
            int64_t run = 0;
            for (int k = 0; k < nt; ++k) {
                const int64_t v = part[k];
                part[k] = run;
                run += v;
            }
        }
        __syncthreads();
        acc = part[tid];
        for (int s = lo; s < hi; ++s) {
            acc += row[s];
            pre[s] = acc;
        }
        __syncthreads();
        const int cj = a.cut[j];
        for (int s = tid; s < W; s += nt) T[int64_t(j) * W + s] = pre[s] - (s >= cj ? pre[s - cj] : 0);
    }
}

// Compact index k -> index into the caller's A/b: identity, or the reference's
// reorder_AB_bargmann (backend_numpy.py:367-375) from [modes..., modes...].
template <int M, bool REORDER>
__device__ __forceinline__ constexpr int src_index(int k) {
    return REORDER ? (k & 1) * M + (k >> 1) : k;
}

template <int M, bool REORDER>
__device__ __forceinline__ double2 a_at(const double2 *A, int i, int j) {
    return A[src_index<M, REORDER>(i) * (2 * M) + src_index<M, REORDER>(j)];
}

// (g B_i + A_i . gin) / sqrt(k), the dot over j >= j0 (gin vanishes below)
template <int M, bool REORDER>
__device__ __forceinline__ double2 pivot_out(const DiagArgs &a, int i, int j0, double2 g, const double2 *gin, int64_t col, int k) {
    double2 dot = make_double2(0.0, 0.0);
#pragma unroll
    for (int j = 0; j < 2 * M; ++j)
        if (j >= j0) dot = cadd(dot, cmul(a_at<M, REORDER>(a.A, i, j), gin[j]));
    const double2 t = cadd(cmul(g, a.B[int64_t(src_index<M, REORDER>(i)) * a.nb + col]), dot);
    return cdiv_by(t, a.sq[k], a.rsq[k]);
}

template <int M, bool REORDER>
__global__ __launch_bounds__(256) void diag_weight_kernel(DiagArgs a, int w, int64_t count) {
    const int64_t t = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t nb = a.nb;
    if (t >= count * nb) return;
    int64_t r = t / nb;
    const int64_t col = t - r * nb;

    // unrank: params of weight w in np.ndindex order
    const int W = a.S + 1;
    int p[M];
    int s = w;
#pragma unroll
    for (int j = 0; j < M - 1; ++j) {
        const int top = s < a.cut[j] - 1 ? s : a.cut[j] - 1;
        int v = 0;
        for (; v < top; ++v) {
            const int64_t n = a.T[int64_t(j + 1) * W + (s - v)];
            if (r < n) break;
            r -= n;
        }
        p[j] = v;
        s -= v;
    }
    p[M - 1] = s;
    int64_t f = 0;
#pragma unroll
    for (int j = 0; j < M; ++j) f += p[j] * a.stride[j];
    const int64_t P = a.P;

    double2 g0;
    if (w == 0) {
        g0 = a.c[a.c_batched ? col : 0];
        a.arr0[col] = g0;
    } else {
        g0 = a.arr0[f * nb + col];
    }

    // diagonal pivot (diagonal_amps.py:108-150)
    double2 ev[M];  // arr1[2d, p], read back by the off-diagonal pivots
#pragma unroll
    for (int m = 0; m < M; ++m) ev[m] = make_double2(0.0, 0.0);
    if (a.cut[0] == 1 || p[0] < a.cut[0] - 1) {
        double2 gin[2 * M];
#pragma unroll
        for (int i = 0; i < 2 * M; ++i) {
            const int m = i >> 1;
            gin[i] = make_double2(0.0, 0.0);
            if (p[m] > 0) {
                const double2 v = a.arr1[((i ^ 1) * P + f - a.stride[m]) * nb + col];
                gin[i] = cscale_promoted(v, a.sq[p[m]]);
            }
        }
#pragma unroll
        for (int i = 0; i < 2 * M; ++i) {
            const int m = i >> 1;
            if (p[m] + 1 < a.cut[m] && (i != 1 || p[0] + 2 < a.cut[0])) {
                const double2 v = pivot_out<M, REORDER>(a, i, 0, g0, gin, col, p[m] + 1);
                a.arr1[(i * P + f) * nb + col] = v;
                if (!(i & 1)) ev[m] = v;
            }
        }
    }

    // off-diagonal pivots for every d with p[:d] == 0 (diagonal_amps.py:19-105)
#pragma unroll
    for (int d = 0; d < M; ++d) {
        if (d > 0 && p[d - 1] != 0) break;
        if (p[d] >= a.cut[d] - 1) continue;
        const double2 g = ev[d];
        double2 gin[2 * M];
#pragma unroll
        for (int j = 0; j < 2 * M; ++j) gin[j] = make_double2(0.0, 0.0);
        gin[2 * d] = cscale_promoted(g0, a.sq[p[d] + 1]);
        if (p[d] > 0) gin[2 * d + 1] = cscale_promoted(a.arr2[(a.off2[d] + f - a.stride[d]) * nb + col], a.sq[p[d]]);
#pragma unroll
        for (int i = d + 1; i < M; ++i) {
            if (p[i] > 0) {
                const int64_t e = (a.off11[d] + (i - d - 1) * a.len[d] + f - a.stride[i]) * nb + col;
                gin[2 * i] = cscale_promoted(a.a1001[e], a.sq[p[i]]);
                gin[2 * i + 1] = cscale_promoted(a.a1010[e], a.sq[p[i]]);
            }
        }
        a.arr0[(f + a.stride[d]) * nb + col] = pivot_out<M, REORDER>(a, 2 * d + 1, 2 * d, g, gin, col, p[d] + 1);
        if (p[d] + 2 < a.cut[d])
            a.arr2[(a.off2[d] + f) * nb + col] = pivot_out<M, REORDER>(a, 2 * d, 2 * d, g, gin, col, p[d] + 2);
#pragma unroll
        for (int i = d + 1; i < M; ++i) {
            if (p[i] + 1 < a.cut[i]) {
                const int64_t e = (a.off11[d] + (i - d - 1) * a.len[d] + f) * nb + col;
                a.a1010[e] = pivot_out<M, REORDER>(a, 2 * i, 2 * d, g, gin, col, p[i] + 1);
                a.a1001[e] = pivot_out<M, REORDER>(a, 2 * i + 1, 2 * d, g, gin, col, p[i] + 1);
            }
        }
    }
}

}  // namespace mm

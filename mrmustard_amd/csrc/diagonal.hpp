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
// launch per weight runs all params of that weight in parallel.  Within a
// launch a params is a group of 2M lanes: lane i holds the i-th input of a
// pivot, the group exchanges them through LDS, and lane i forms output i (the
// dot of A's row i with the inputs): every output of both pivot kinds is row
// i of A for lane i.
//
// Params are numbered by their position in a list sorted by weight (np.ndindex
// order within a weight, helperFunctions.py:31-46), so the params of a launch
// are contiguous.  The producer of a value writes it straight into its one
// consumer's inbox, at the consumer's position: lane i of a pivot at p sends
// its output to slot i ^ 1 of p + e_m (m = i / 2), for the diagonal pivot's
// arr1 values (the consumer's diagonal inbox) and for the d = 0 off-diagonal
// pivot's arr0 / arr2 / arr1010 / arr1001 values (the consumer's off-diagonal
// inbox; slot 0 holds arr0[p], read as the consumer's G[p ⊗ p]).  A consumer
// then reads all its inputs at its own position: one coalesced round trip.
// Values of the off-diagonal pivots d >= 1, met only by params with p[0] == 0,
// stay in the reference's sub-lattices (indexed by flat p, of length
// prod(cutoffs[d:]) since p[:d] == 0) and are gathered at p - e_m.
//
// Positions come from count tables: T[j][s] = #{(p_j..p_{M-1}) : p_k <
// cutoff_k, sum = s} and their row prefix sums Q give the position of p, or of
// p + e_m, in O(M) (list_position).
#pragma once

#include "common.hpp"

namespace mm {

constexpr int DG_MAXM = 10;
constexpr int DG_THREADS = 256;

struct DiagArgs {
    const double2 *A, *B, *c;  // A: 2M x 2M; B: 2M x nb; c: 1 or nb values
    int c_batched;
    double2 *arr0;              // output, by flat p
    double2 *inD, *inO;         // 2M x P inboxes by position: diagonal and d = 0 off-diagonal inputs
    double2 *arr2, *a1010, *a1001;  // d >= 1 sub-lattices, by flat p
    int64_t *T, *Q;     // (M + 1) x (S + 1) counts and their row prefix sums (global copy, large S only)
    int32_t *list;      // P < 2^31 flat indices, sorted by weight, ndindex order within a weight
    int32_t *fwd;       // M x P: list position of p + e_m for the params at each position (-1: outside)
    const double *sq, *rsq;
    int64_t P, nb;
    int S;
    int cut[DG_MAXM];
    int64_t stride[DG_MAXM];
    int64_t len[DG_MAXM];   // prod(cut[d:])
    int64_t off2[DG_MAXM];  // start of arr2's slice d >= 1 (entries)
    int64_t off11[DG_MAXM]; // start of arr1010/arr1001's block d >= 1 (M - d - 1 slices of len[d])
};

__device__ __forceinline__ int64_t shfl_up_i64(int64_t v, int delta) {
    int lo = int(uint64_t(v) & 0xffffffffu), hi = int(uint64_t(v) >> 32);
    lo = __shfl_up(lo, delta, MM_WAVE);
    hi = __shfl_up(hi, delta, MM_WAVE);
    return int64_t((uint64_t(uint32_t(hi)) << 32) | uint32_t(lo));
}

// T[M][s] = (s == 0); Q[j] = prefix(T[j]); T[j-1][s] = Q[j][s] - Q[j][s - cut_{j-1}],
// down to Q[0] (the start of each weight in the list).  Run by the 64 lanes of
// one wave (a chunk of each row per lane, a shuffle scan of the chunk sums), so
// no block barrier is needed; T and Q may point to LDS or to global memory.
__device__ void diag_build_tables(const DiagArgs &a, int M, int64_t *T, int64_t *Q) {
    const int W = a.S + 1, lane = threadIdx.x & (MM_WAVE - 1);
    const int chunk = (W + MM_WAVE - 1) / MM_WAVE;
    const int lo = min(W, lane * chunk), hi = min(W, lo + chunk);
    for (int s = lane; s < W; s += MM_WAVE) T[int64_t(M) * W + s] = (s == 0);
    wave_sync();
    for (int j = M; j >= 0; --j) {
        const int64_t *row = T + int64_t(j) * W;
        int64_t *out = Q + int64_t(j) * W;
        int64_t own = 0;
        for (int s = lo; s < hi; ++s) own += row[s];
        int64_t incl = own;
#pragma unroll
        for (int d = 1; d < MM_WAVE; d <<= 1) {
            const int64_t v = shfl_up_i64(incl, d);
            if (lane >= d) incl += v;
        }
        int64_t acc = incl - own;
        for (int s = lo; s < hi; ++s) {
            acc += row[s];
            out[s] = acc;
        }
        wave_sync();
        if (j == 0) break;
        const int cj = a.cut[j - 1];
        for (int s = lane; s < W; s += MM_WAVE) T[int64_t(j - 1) * W + s] = out[s] - (s >= cj ? out[s - cj] : 0);
        wave_sync();
    }
}

// Global tables, for count tables too large for LDS (one wave).
__global__ __launch_bounds__(64) void diag_counts_kernel(DiagArgs a, int M) { diag_build_tables(a, M, a.T, a.Q); }

// List position of p (m < 0) or of p + e_m from the prefix tables:
//   start(w') + sum_{j < M-1} C_j(x_j, t_j),  C_j(x, t) = Q[j+1][t] - Q[j+1][t - x],
// with w' = |p| or |p| + 1 and (x_j, t_j) the digits and suffix sums of that
// params: for p + e_m, t_j = s_j + 1 for j <= m and x_m = p_m + 1.
template <int M>
__device__ __forceinline__ int64_t list_position(const int64_t *Q, int W, const int *p, int w, int m) {
    const int wp = m < 0 ? w : w + 1;
    int64_t g = wp > 0 ? Q[wp - 1] : 0;
    int s = w;
#pragma unroll
    for (int j = 0; j < M - 1; ++j) {
        const int64_t *Qr = Q + int64_t(j + 1) * W;
        const int hi = j <= m ? s + 1 : s;
        const int lo = j < m ? s + 1 - p[j] : s - p[j];
        g += Qr[hi] - Qr[lo];
        s -= p[j];
    }
    return g;
}

// For every p of the box at list position g: list[g] = flat(p) and
// fwd[m][g] = list position of p + e_m (-1 where p_m + 1 == cutoff_m).  With lds_tables
// each block first builds T and Q in its dynamic LDS; otherwise it reads the
// global Q that diag_counts_kernel built.
template <int M>
__global__ __launch_bounds__(256) void diag_rank_kernel(DiagArgs a, int lds_tables) {
    extern __shared__ int64_t tab[];
    const int W = a.S + 1;
    const int64_t *Q = a.Q;
    if (lds_tables) {
        if (threadIdx.x < MM_WAVE) diag_build_tables(a, M, tab, tab + int64_t(M + 1) * W);
        __syncthreads();
        Q = tab + int64_t(M + 1) * W;
    }
    const uint32_t P = uint32_t(a.P);
    const uint32_t inc = gridDim.x * blockDim.x;
    for (uint32_t f = blockIdx.x * blockDim.x + threadIdx.x; f < P; f += inc) {
        int p[M];
        uint32_t r = f;
        int w = 0;
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const uint32_t sj = uint32_t(a.stride[j]);
            p[j] = int(r / sj);
            r -= uint32_t(p[j]) * sj;
            w += p[j];
        }
        const int64_t g = list_position<M>(Q, W, p, w, -1);
        a.list[g] = int32_t(f);
#pragma unroll
        for (int m = 0; m < M; ++m) a.fwd[m * a.P + g] = p[m] + 1 < a.cut[m] ? int32_t(list_position<M>(Q, W, p, w, m)) : -1;
    }
}

// Compact index k -> index into the caller's A/b: identity, or the reference's
// reorder_AB_bargmann (backend_numpy.py:367-375) from [modes..., modes...].
template <int M, bool REORDER>
__device__ __forceinline__ int src_index(int k) {
    return REORDER ? (k & 1) * M + (k >> 1) : k;
}

// One launch per weight w: PB params per block, lane (slot, i) = (t % PB, t / PB)
// so that a wave's loads of one inbox slot cover consecutive positions.  All
// inputs of all pivots are loaded at once and parked in LDS, one slot per
// pivot: 0 for the diagonal pivot, 1 + d for the off-diagonal pivot d.  A sits
// in LDS too (lane i reads row i).  Two barriers: after the inputs land, and
// after the diagonal pivot's own outputs arr1[2d, p] (inputs of the
// off-diagonal pivots) are in LDS.
template <int M, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void diag_group_kernel(DiagArgs a, int w, int64_t count, int64_t start) {
    constexpr int G = 2 * M;            // lanes per params
    constexpr int PB = DG_THREADS / G;  // params per block
    __shared__ double2 sgin[M + 1][G][PB + 1];
    __shared__ double2 sev[M][PB];
    __shared__ double2 sA[G][G];  // A in the compact order
    const int t = threadIdx.x;
    for (int e = t; e < G * G; e += DG_THREADS)
        sA[e / G][e % G] = a.A[src_index<M, REORDER>(e / G) * G + src_index<M, REORDER>(e % G)];
    const int slot = t % PB, i = t / PB, m = i >> 1;
    const bool lane_ok = i < G;
    const int64_t nb = a.nb, P = a.P;
    const int64_t grp = int64_t(blockIdx.x) * PB + slot;
    const bool act = lane_ok && grp < count * nb;

    int64_t f = 0, col = 0, gs = 0, fw = -1;
    double2 g0 = make_double2(0.0, 0.0), xD = make_double2(0.0, 0.0), xO = make_double2(0.0, 0.0);
    int p[M];
#pragma unroll
    for (int j = 0; j < M; ++j) p[j] = 0;
    if (act) {
        const int64_t q = nb == 1 ? grp : grp / nb;
        col = grp - q * nb;
        gs = start + q;
        f = a.list[gs];
        fw = a.fwd[m * P + gs];  // position of p + e_m, where this lane's outputs go
        if (w == 0) {
            g0 = a.c[a.c_batched ? col : 0];
            if (i == 0) a.arr0[col] = g0;
        } else {
            g0 = a.inO[gs * nb + col];  // slot 0: arr0[p]
            xD = a.inD[(i * P + gs) * nb + col];
            xO = a.inO[(i * P + gs) * nb + col];
        }
        uint32_t r = uint32_t(f);
#pragma unroll
        for (int j = 0; j < M; ++j) {
            const uint32_t sj = uint32_t(a.stride[j]);
            p[j] = int(r / sj);
            r -= uint32_t(p[j]) * sj;
        }
    }
    // this lane's mode: p[m], stride[m], cut[m] without dynamic register indexing
    int pm = 0, cm = 1;
    int64_t sm = 0;
#pragma unroll
    for (int j = 0; j < M; ++j)
        if (j == m) {
            pm = p[j];
            cm = a.cut[j];
            sm = a.stride[j];
        }
    // off-diagonal pivots this params runs: d with p[:d] == 0 and p[d] < cut[d] - 1
    int runmask = 0;
    if (act) {
        bool zp = true;
#pragma unroll
        for (int d = 0; d < M; ++d) {
            if (zp && p[d] < a.cut[d] - 1) runmask |= 1 << d;
            zp = zp && p[d] == 0;
        }
    }

    // ---- inputs, scaled by sqrt of the pivot's entries (K_l).  Everything a
    // lane may need is loaded before any of it is used, in one round trip:
    // the inputs of the off-diagonal pivots d >= 1 (from a read-only dummy
    // address where the lane has none) and the three sqrt entries of its mode
    // (every sqrt a lane uses is of p_m, p_m + 1 or p_m + 2: the lanes 2d and
    // 2d + 1 of pivot d have m = d).
    double2 xo[M];
#pragma unroll
    for (int d = 1; d < M; ++d) {
        const double2 *src = a.A;
        int64_t e = 0;
        if (act && (runmask >> d & 1) && i > 2 * d && pm > 0) {
            if (i == 2 * d + 1) {
                src = a.arr2;
                e = (a.off2[d] + f - sm) * nb + col;
            } else {
                src = (i & 1) ? a.a1010 : a.a1001;
                e = (a.off11[d] + (m - d - 1) * a.len[d] + f - sm) * nb + col;
            }
        }
        xo[d] = src[e];
    }
    const double s0 = a.sq[pm], s1 = a.sq[pm + 1], r1 = a.rsq[pm + 1], s2 = a.sq[pm + 2], r2 = a.rsq[pm + 2];
    double2 bi = make_double2(0.0, 0.0);
    if (act) bi = a.B[int64_t(src_index<M, REORDER>(i)) * nb + col];
    const bool diag = act && (a.cut[0] == 1 || p[0] < a.cut[0] - 1);
    if (lane_ok) sgin[0][i][slot] = diag && pm > 0 ? cscale_promoted(xD, s0) : make_double2(0.0, 0.0);
#pragma unroll
    for (int d = 0; d < M; ++d) {
        if (!lane_ok || i < 2 * d) continue;
        double2 v = make_double2(0.0, 0.0);
        if (runmask >> d & 1) {
            if (i == 2 * d)
                v = cscale_promoted(g0, s1);
            else if (pm > 0)
                v = cscale_promoted(d == 0 ? xO : xo[d], s0);  // d == 0: the inbox
        }
        sgin[1 + d][i][slot] = v;
    }
    __syncthreads();

    // ---- diagonal pivot (diagonal_amps.py:108-150): arr1[i, p] -> slot i ^ 1 of p + e_m
    if (diag && pm + 1 < cm && (i != 1 || p[0] + 2 < a.cut[0])) {
        double2 dot = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 0; j < G; ++j) dot = cadd(dot, cmul(sA[i][j], sgin[0][j][slot]));
        const double2 out = cdiv_by(cadd(cmul(g0, bi), dot), s1, r1);
        a.inD[((i ^ 1) * P + fw) * nb + col] = out;
        if (!(i & 1)) sev[m][slot] = out;
    }
    __syncthreads();

    // ---- off-diagonal pivots (diagonal_amps.py:19-105)
#pragma unroll
    for (int d = 0; d < M; ++d) {
        if (!(runmask >> d & 1) || i < 2 * d) continue;
        double2 dot = make_double2(0.0, 0.0);
#pragma unroll
        for (int j = 2 * d; j < G; ++j) dot = cadd(dot, cmul(sA[i][j], sgin[1 + d][j][slot]));
        const double2 tt = cadd(cmul(sev[d][slot], bi), dot);
        if (i == 2 * d + 1) {
            // arr0[p + e_d]: the output, and slot 0 of that params' inbox (p_d = p_m)
            const double2 o = cdiv_by(tt, s1, r1);
            a.arr0[(f + a.stride[d]) * nb + col] = o;
            a.inO[fw * nb + col] = o;
        } else if (i == 2 * d) {
            if (pm + 2 < cm) {
                const double2 o = cdiv_by(tt, s2, r2);
                if (d == 0)
                    a.inO[(P + fw) * nb + col] = o;  // slot 1 of p + e_0
                else
                    a.arr2[(a.off2[d] + f) * nb + col] = o;
            }
        } else if (pm + 1 < cm) {
            const double2 o = cdiv_by(tt, s1, r1);
            if (d == 0) {
                a.inO[((i ^ 1) * P + fw) * nb + col] = o;
            } else {
                const int64_t e = (a.off11[d] + (m - d - 1) * a.len[d] + f) * nb + col;
                if (i & 1)
                    a.a1001[e] = o;
                else
                    a.a1010[e] = o;
            }
        }
    }
}

}  // namespace mm

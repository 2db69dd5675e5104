// Vector-Jacobian product of the diagonal strategy: (dA, dB, dc) = sum_p
// cot[p] d arr0[p] / d(A, B, c), holomorphic (no conjugation), A's entries
// independent, in the compact order, as fock_representation_diagonal_grad
// (compactFock/diagonal_grad.py:19-38, :41-258) differentiates the recursion
// and hermite_renormalized_diagonal_jax_bwd (jax_vjps/hermite.py:292-330)
// contracts it with the cotangent.
//
// Reverse sweep over the weights, highest first, after the forward
// (diag_group_kernel) has filled the inboxes.  Every pivot output is
// o_i = (g B_i + sum_l A_il x_l) / K_i with x_l = K_l[l] raw_l, so with
// w_i = o-bar_i / K_i:
//   dA_il += w_i x_l,  dB_i += w_i g,  g-bar += sum_i w_i B_i,
//   raw-bar_l = K_l[l] sum_i w_i A_il.
// A value's one consumer writes the adjoint of each input it read at the
// location it read it from (its own inbox slot, or the d >= 1 sub-lattice
// entry), and the producer reads the adjoints of its outputs back at the
// locations it wrote them to: the adjoint arrays mirror the forward's.  Lane
// i owns row i of dA and dB_i, kept in its own LDS column over a grid-stride
// loop; each block adds them to its own partial row (the same block slot in
// every launch), and the rows are summed in a fixed order at the end:
// deterministic.
#pragma once

#include "diagonal.hpp"

namespace mm {

struct DiagVjpArgs {
    DiagArgs a;                        // forward geometry and stored values (nb == 1)
    const double2 *cot;                // cotangent of arr0, by flat p
    double2 *adD, *adO;                // adjoints of the inboxes
    double2 *ad2, *ad1010, *ad1001;    // adjoints of the d >= 1 sub-lattices
    double2 *part;                     // nblk x 2M x (2M + 1): dA rows, then dB
    double2 *dc;                       // adjoint of c
};

template <int M, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void diag_vjp_kernel(DiagVjpArgs v, int w, int64_t count, int64_t start, int iters) {
    constexpr int G = 2 * M;
    constexpr int PB = DG_THREADS / G;
    __shared__ double2 sA[G][G];
    __shared__ double2 sB[G];
    __shared__ double2 sx[G][PB + 1];
    __shared__ double2 sw[G][PB + 1];
    __shared__ double2 sg[PB];
    __shared__ double2 sacc[G + 1][DG_THREADS];  // lane t's dA row (l < G) and dB (l == G), column t
    __shared__ int need;                          // off-diagonal pivots some params of the block runs
    const DiagArgs &a = v.a;
    const int t = threadIdx.x;
    for (int e = t; e < G * G; e += DG_THREADS) sA[e / G][e % G] = a.A[src_index<M, REORDER>(e / G) * G + src_index<M, REORDER>(e % G)];
    if (t < G) sB[t] = a.B[src_index<M, REORDER>(t)];
    const int slot = t % PB, i = t / PB, m = i >> 1;
    const bool lane_ok = i < G;
    const int64_t P = a.P;
#pragma unroll
    for (int l = 0; l <= G; ++l) sacc[l][t] = make_double2(0.0, 0.0);
    __syncthreads();

    for (int it = 0; it < iters; ++it) {
        const int64_t grp = (int64_t(it) * gridDim.x + blockIdx.x) * PB + slot;
        const bool act = lane_ok && grp < count;
        int64_t f = 0, gs = 0, fw = -1;
        int p[M];
#pragma unroll
        for (int j = 0; j < M; ++j) p[j] = 0;
        if (act) {
            gs = start + grp;
            f = a.list[gs];
            fw = a.fwd[m * P + gs];
            uint32_t r = uint32_t(f);
#pragma unroll
            for (int j = 0; j < M; ++j) {
                const uint32_t sj = uint32_t(a.stride[j]);
                p[j] = int(r / sj);
                r -= uint32_t(p[j]) * sj;
            }
        }
        int pm = 0, cm = 1;
        int64_t sm = 0;
#pragma unroll
        for (int j = 0; j < M; ++j)
            if (j == m) {
                pm = p[j];
                cm = a.cut[j];
                sm = a.stride[j];
            }
        int runmask = 0;
        if (act) {
            bool zp = true;
#pragma unroll
            for (int d = 0; d < M; ++d) {
                if (zp && p[d] < a.cut[d] - 1) runmask |= 1 << d;
                zp = zp && p[d] == 0;
            }
        }
        const double2 g0 = act ? (w == 0 ? a.c[0] : a.inO[gs]) : make_double2(0.0, 0.0);
        double2 gbar0 = make_double2(0.0, 0.0);  // this lane's share of arr0[p]'s adjoint
        double2 gext = make_double2(0.0, 0.0);   // adjoint of arr1[i, p] from the off-diagonal pivot i / 2

        if (t == 0) need = 0;
        __syncthreads();
        if (runmask && i == 0) atomicOr(&need, runmask);
        // every load of every pivot depends on p only: the addresses first,
        // then all the loads at once (from a read-only dummy address where the
        // lane needs none), then the arithmetic.  Every sqrt entry a lane uses
        // is of p_m, p_m + 1 or p_m + 2 (the lanes 2d and 2d + 1 of pivot d
        // have m = d, so p_d = p_m there).
        const double s0 = a.sq[pm], s1 = a.sq[pm + 1], r1 = a.rsq[pm + 1], s2 = a.sq[pm + 2], r2 = a.rsq[pm + 2];
        double2 xr[M], obr[M];
#pragma unroll
        for (int d = 0; d < M; ++d) {
            const bool run = (runmask >> d & 1) && i >= 2 * d;
            const double2 *xp = a.A, *op = a.A;
            int64_t xe = 0, oe = 0;
            if (run) {
                if (i == 2 * d) {
                    if (pm + 2 < cm) {
                        op = d == 0 ? v.adO : v.ad2;
                        oe = d == 0 ? P + fw : a.off2[d] + f;
                    }
                } else if (i == 2 * d + 1) {
                    if (pm > 0) {
                        xp = d == 0 ? a.inO : a.arr2;
                        xe = d == 0 ? P + gs : a.off2[d] + f - sm;
                    }
                    op = v.adO;  // arr0[p + e_d]: slot 0 of its consumer
                    oe = fw;
                } else {
                    const int64_t sl = a.off11[d] + (m - d - 1) * a.len[d];
                    if (pm > 0) {
                        xp = d == 0 ? a.inO : ((i & 1) ? a.a1010 : a.a1001);
                        xe = d == 0 ? i * P + gs : sl + f - sm;
                    }
                    if (pm + 1 < cm) {
                        op = d == 0 ? v.adO : ((i & 1) ? v.ad1001 : v.ad1010);
                        oe = d == 0 ? (i ^ 1) * P + fw : sl + f;
                    }
                }
            }
            xr[d] = xp[xe];
            obr[d] = op[oe];
        }
        // arr1[2d, p] for lane 2d (d = m), sent to p + e_d's slot 2d + 1
        const bool gon = !(i & 1) && (runmask >> m & 1) && i < G;
        const double2 gd = (gon ? a.inD : a.A)[gon ? (i + 1) * P + fw : 0];
        // the diagonal pivot's own inputs and output adjoint (without the off-diagonal pivots' share)
        const bool diag = act && (a.cut[0] == 1 || p[0] < a.cut[0] - 1);
        const bool dwrite = diag && pm + 1 < cm && (i != 1 || p[0] + 2 < a.cut[0]);
        const int c0p = p[0] + (m == 0 ? 1 : 0);  // the consumer p + e_m read it iff its own diagonal pivot ran
        const bool xon = diag && pm > 0, oon = dwrite && (a.cut[0] == 1 || c0p < a.cut[0] - 1);
        const double2 xDr = (xon ? a.inD : a.A)[xon ? i * P + gs : 0];
        const double2 obDr = (oon ? v.adD : a.A)[oon ? (i ^ 1) * P + fw : 0];
        double2 xs[M], wvs[M];
#pragma unroll
        for (int d = 0; d < M; ++d) {
            xs[d] = make_double2(0.0, 0.0);
            wvs[d] = make_double2(0.0, 0.0);
            if (!((runmask >> d & 1) && i >= 2 * d)) continue;
            if (i == 2 * d) {
                xs[d] = cscale_promoted(g0, s1);
                if (pm + 2 < cm) wvs[d] = cdiv_by(obr[d], s2, r2);
            } else if (i == 2 * d + 1) {
                if (pm > 0) xs[d] = cscale_promoted(xr[d], s0);
                wvs[d] = cdiv_by(obr[d], s1, r1);
            } else {
                if (pm > 0) xs[d] = cscale_promoted(xr[d], s0);
                if (pm + 1 < cm) wvs[d] = cdiv_by(obr[d], s1, r1);
            }
        }
        const double2 xD = xon ? cscale_promoted(xDr, s0) : make_double2(0.0, 0.0);
        const double2 obD = oon ? obDr : make_double2(0.0, 0.0);

        __syncthreads();
        const int need_all = need;

        // ---- off-diagonal pivots (diagonal_amps.py:19-105 in reverse)
#pragma unroll
        for (int d = 0; d < M; ++d) {
            if (!(need_all >> d & 1)) continue;  // block-uniform: no params of the block runs pivot d
            const bool run = (runmask >> d & 1) && i >= 2 * d;
            const double2 x = xs[d], wv = wvs[d];
            if (run && i == 2 * d) sg[slot] = gd;
            if (lane_ok) {
                sx[i][slot] = x;
                sw[i][slot] = wv;
            }
            __syncthreads();
            if (run) {
#pragma unroll
                for (int l = 2 * d; l < G; ++l) sacc[l][t] = cfma(wv, sx[l][slot], sacc[l][t]);
                sacc[G][t] = cfma(wv, sg[slot], sacc[G][t]);
                // lane i as column l = i: sum_i' w_i' A_i'l, times K_l[l]
                double2 s = make_double2(0.0, 0.0);
#pragma unroll
                for (int r = 2 * d; r < G; ++r) s = cfma(sw[r][slot], sA[r][i], s);
                if (i == 2 * d) {
                    gbar0 = caxpy(s1, s, gbar0);
                    double2 gb = make_double2(0.0, 0.0);  // adjoint of arr1[2d, p]
#pragma unroll
                    for (int r = 2 * d; r < G; ++r) gb = cfma(sw[r][slot], sB[r], gb);
                    gext = gb;
                } else if (i == 2 * d + 1) {
                    if (pm > 0) {
                        const double2 raw = make_double2(s.x * s0, s.y * s0);
                        if (d == 0)
                            v.adO[P + gs] = raw;
                        else
                            v.ad2[a.off2[d] + f - sm] = raw;
                    }
                } else if (pm > 0) {
                    const double2 raw = make_double2(s.x * s0, s.y * s0);
                    if (d == 0) {
                        v.adO[i * P + gs] = raw;
                    } else {
                        const int64_t en = a.off11[d] + (m - d - 1) * a.len[d] + f - sm;
                        if (i & 1)
                            v.ad1010[en] = raw;
                        else
                            v.ad1001[en] = raw;
                    }
                }
            }
            __syncthreads();
        }

        // ---- diagonal pivot (diagonal_amps.py:108-150 in reverse)
        double2 x = xD, wv = make_double2(0.0, 0.0);
        if (dwrite) {
            double2 ob = obD;
            if (!(i & 1) && (runmask >> m & 1)) ob = cadd(ob, gext);
            wv = cdiv_by(ob, s1, r1);
        }
        if (lane_ok) {
            sx[i][slot] = x;
            sw[i][slot] = wv;
        }
        __syncthreads();
        double2 gdiag = make_double2(0.0, 0.0);
        if (diag) {
#pragma unroll
            for (int l = 0; l < G; ++l) sacc[l][t] = cfma(wv, sx[l][slot], sacc[l][t]);
            sacc[G][t] = cfma(wv, g0, sacc[G][t]);
            double2 s = make_double2(0.0, 0.0);
#pragma unroll
            for (int r = 0; r < G; ++r) s = cfma(sw[r][slot], sA[r][i], s);
            if (pm > 0) v.adD[i * P + gs] = make_double2(s.x * s0, s.y * s0);
            if (i == 0) {
#pragma unroll
                for (int r = 0; r < G; ++r) gdiag = cfma(sw[r][slot], sB[r], gdiag);
            }
        }
        // arr0[p]'s adjoint: cotangent + diagonal pivot's g term + the off-diagonal pivots' x_2d terms
        if (lane_ok) sx[i][slot] = gbar0;
        __syncthreads();
        if (act && i == 0) {
            double2 gb = cadd(v.cot[f], gdiag);
#pragma unroll
            for (int d = 0; d < M; ++d) gb = cadd(gb, sx[2 * d][slot]);
            if (w == 0)
                *v.dc = gb;
            else
                v.adO[gs] = gb;
        }
        __syncthreads();
    }

    // per-block sums over the PB params slots of each row, added to this block's partial row
    double2 *prow = v.part + int64_t(blockIdx.x) * G * (G + 1);
    for (int e = t; e < G * (G + 1); e += DG_THREADS) {
        const int r = e / (G + 1), l = e % (G + 1);
        double2 s = prow[e];
        for (int k = 0; k < PB; ++k) s = cadd(s, sacc[l][r * PB + k]);
        prow[e] = s;
    }
}

// dA (caller's order), dB: one block per entry; thread k sums partial rows
// k, k + 256, ... in order, then a fixed tree over the 256 threads.
template <int M, bool REORDER>
__global__ __launch_bounds__(256) void diag_vjp_finalize(const double2 *part, int nblk, double2 *dA, double2 *dB) {
    constexpr int G = 2 * M;
    __shared__ double2 red[256];
    const int e = blockIdx.x, k = threadIdx.x;
    double2 s = make_double2(0.0, 0.0);
    for (int b = k; b < nblk; b += 256) s = cadd(s, part[int64_t(b) * G * (G + 1) + e]);
    red[k] = s;
    __syncthreads();
    for (int h = 128; h > 0; h >>= 1) {
        if (k < h) red[k] = cadd(red[k], red[k + h]);
        __syncthreads();
    }
    if (k == 0) {
        const int r = e / (G + 1), l = e % (G + 1);
        if (l < G)
            dA[src_index<M, REORDER>(r) * G + src_index<M, REORDER>(l)] = red[0];
        else
            dB[src_index<M, REORDER>(r)] = red[0];
    }
}

__global__ void zero_c128_span_kernel(double2 *p, int64_t n) {
    for (int64_t k = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; k < n; k += int64_t(gridDim.x) * blockDim.x)
        p[k] = make_double2(0.0, 0.0);
}

}  // namespace mm

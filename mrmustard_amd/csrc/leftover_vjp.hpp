// Vector-Jacobian product of the one-leftover-mode strategy: (dA, dB, dc) =
// sum cot * d out / d(A, B, c), holomorphic, A's entries independent, in the
// compact order, as fock_representation_1leftoverMode_grad
// (compactFock/singleLeftoverMode_grad.py:19-46, :50-225, :254-556, :560-692)
// differentiates the recursion and hermite_renormalized_1leftoverMode_jax_bwd
// (jax_vjps/hermite.py:397-440) contracts it with the cotangent.
//
// Reverse sweep over the weights, highest first, after the forward
// (lo_init / lo_diag / lo_off kernels) has filled the inboxes, two launches
// per weight in the reverse of the forward's order: off-diagonal pivots, then
// diagonal pivots.  Every pivot output at block entry (m, n) is
//   o_r(m, n) = (piv(m, n) B_r + A_r0 sqrt(m) piv(m-1, n) + A_r1 sqrt(n) piv(m, n-1)
//                + sum_l A_rl x_l(m, n)) / K_r,
// so with w_r = o-bar_r / K_r:
//   dA_rl += w_r x_l,  dB_r += w_r piv,  dA_r0 += w_r sqrt(m) piv(m-1, n),
//   dA_r1 += w_r sqrt(n) piv(m, n-1),  x-bar_l = sum_r w_r A_rl,
//   piv-bar(m, n) = sum_r (w_r(m, n) B_r + w_r(m+1, n) A_r0 sqrt(m+1)
//                          + w_r(m, n+1) A_r1 sqrt(n+1)).
// The pivot-block adjoint is a gather: each lane reads the output adjoints of
// its own block entry and of the two entries above it, which the launches of
// the weight above have completed.  Adjoint arrays mirror the forward's (a
// consumer writes the adjoint of an input where it read it, a producer reads
// the adjoint of an output where it wrote it) and are cleared first, so an
// output nobody consumed reads zero.  The zero-detection block is reversed
// last, column by column, by one block.  dA / dB rows go through per-block
// partial rows summed in a fixed order: deterministic.
#pragma once

#include "leftover.hpp"

namespace mm {

struct LeftVjpArgs {
    LeftArgs a;                       // forward geometry and stored values
    const double2 *cot;               // cotangent of out: entry (pos, flat p) at pos * P + f
    double2 *adD, *adO;               // adjoints of the inboxes
    double2 *ad2, *ad1010, *ad1001;   // adjoints of the d >= 1 sub-lattices
    double2 *adP;                     // adjoints of the arr1[2d, p] blocks from the off-diagonal pivots
    double2 *part;                    // (nblk + 1) x 2L x (2L + 1): dA rows, then dB
    double2 *dc;                      // adjoint of c
};

// Output adjoints of the lane's block entry and of the entries (m+1, n),
// (m, n+1) of the array at src (block entry innermost), divided by K.
__device__ __forceinline__ void lo_out_adjoints(const double2 *src, const double2 *ext, int pos, int c0, bool up_m, bool up_n,
                                                double K, double rK, double2 &wo, double2 &wm, double2 &wn) {
    auto ob = [&](int e) {
        double2 o = src[e];
        if (ext) o = cadd(o, ext[e]);
        return cdiv_by(o, K, rK);
    };
    wo = ob(pos);
    if (up_m) wm = ob(pos + c0);
    if (up_n) wn = ob(pos + 1);
}

// Lane (i, slot) of output row r = i + 2: its term of the pivot-block adjoint.
template <int L>
__device__ __forceinline__ double2 lo_piv_term(const double2 (*sA)[2 * L], const double2 *sB, int r, double2 wo, double2 wm,
                                               double2 wn, bool up_m, bool up_n, double sqm1, double sqn1) {
    double2 t = cmul(wo, sB[r]);
    if (up_m) t = cfma(cscale_promoted(wm, sqm1), sA[r][0], t);
    if (up_n) t = cfma(cscale_promoted(wn, sqn1), sA[r][1], t);
    return t;
}

// Per-block sums over the PB params slots of each pivot row (rows 2 ..), added
// to this block's partial row.
template <int MT>
__device__ __forceinline__ void lo_vjp_flush(double2 *part, const double2 (*sacc)[DG_THREADS]) {
    constexpr int G = 2 * MT, GL = G + 2;
    constexpr int PB = DG_THREADS / G;
    double2 *prow = part + int64_t(blockIdx.x) * GL * (GL + 1);
    for (int e = threadIdx.x; e < G * (GL + 1); e += DG_THREADS) {
        const int i = e / (GL + 1), l = e % (GL + 1);
        const int64_t pe = int64_t(i + 2) * (GL + 1) + l;
        double2 s = prow[pe];
        for (int k = 0; k < PB; ++k) s = cadd(s, sacc[l][i * PB + k]);
        prow[pe] = s;
    }
}

// Off-diagonal pivots of one weight in reverse (use_offDiag_pivot,
// singleLeftoverMode_amps.py:99-221; use_offDiag_pivot_grad, singleLeftoverMode_grad.py:254-461).
template <int MT, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void lo_off_vjp_kernel(LeftVjpArgs v, int64_t count, int64_t start, int iters) {
    constexpr int G = 2 * MT, L = MT + 1, GL = 2 * L;
    constexpr int PB = DG_THREADS / G;
    __shared__ double2 sA[GL][GL];
    __shared__ double2 sB[GL];
    __shared__ double2 sx[G][PB + 1];
    __shared__ double2 sw[G][PB + 1];
    __shared__ double2 st[G][PB + 1];
    __shared__ double2 sacc[GL + 1][DG_THREADS];  // lane t's dA row (l < GL) and dB (l == GL)
    const LeftArgs &a = v.a;
    const int t = threadIdx.x;
    for (int e = t; e < GL * GL; e += DG_THREADS)
        sA[e / GL][e % GL] = a.A[src_index<L, REORDER>(e / GL) * GL + src_index<L, REORDER>(e % GL)];
    if (t < GL) sB[t] = a.B[src_index<L, REORDER>(t)];
#pragma unroll
    for (int l = 0; l <= GL; ++l) sacc[l][t] = make_double2(0.0, 0.0);
    __syncthreads();
    const int64_t P = a.t.P, C = a.C;
    const int c0 = a.c0;
    const double2 z = make_double2(0.0, 0.0);

    for (int it = 0; it < iters; ++it) {
        const LoLane<MT> l = lo_lane<MT, PB>(a, count, start, int64_t(it) * gridDim.x + blockIdx.x);
        const int runmask = lo_runmask<MT>(a, l);
        const bool up_m = l.bm + 1 < c0, up_n = l.bn + 1 < c0;
        const double2 g0 = l.act ? a.t.inO[l.gs * C + l.pos] : z;
        double2 gbar0 = z;  // this lane's share of arr0[p]'s adjoint (lanes 2d)
#pragma unroll
        for (int d = 0; d < MT; ++d) {
            const bool run = (runmask >> d & 1) && l.i >= 2 * d;
            const int pd = l.p[d];
            double2 x = z, wo = z, wm = z, wn = z;
            if (run) {
                // the input this lane's column carries (as lo_off_kernel reads it)
                if (l.i == 2 * d) {
                    x = cscale_promoted(g0, a.t.sq[pd + 1]);
                } else if (l.pm > 0) {
                    double2 xr;
                    if (d == 0) {
                        xr = a.t.inO[(l.i * P + l.gs) * C + l.pos];
                    } else if (l.i == 2 * d + 1) {
                        xr = a.t.arr2[(a.t.off2[d] + l.f - l.sm) * C + l.pos];
                    } else {
                        const int64_t en = (a.t.off11[d] + (l.m - d - 1) * a.t.len[d] + l.f - l.sm) * C + l.pos;
                        xr = (l.i & 1) ? a.t.a1010[en] : a.t.a1001[en];
                    }
                    x = cscale_promoted(xr, a.t.sq[l.pm]);
                }
                // the adjoint of this lane's output, where lo_off_kernel wrote it
                const double2 *src = nullptr;
                int kk = 0;
                if (l.i == 2 * d + 1) {
                    src = v.adO + l.fw * C;  // arr0[p + e_d]
                    kk = pd + 1;
                } else if (l.i == 2 * d) {
                    if (pd + 2 < a.t.cut[d]) {
                        src = d == 0 ? v.adO + (P + l.fw) * C : v.ad2 + (a.t.off2[d] + l.f) * C;
                        kk = pd + 2;
                    }
                } else if (l.pm + 1 < l.cm) {
                    if (d == 0) {
                        src = v.adO + ((l.i ^ 1) * P + l.fw) * C;
                    } else {
                        const int64_t e = (a.t.off11[d] + (l.m - d - 1) * a.t.len[d] + l.f) * C;
                        src = (l.i & 1) ? v.ad1001 + e : v.ad1010 + e;
                    }
                    kk = l.pm + 1;
                }
                if (src) lo_out_adjoints(src, nullptr, l.pos, c0, up_m, up_n, a.t.sq[kk], a.t.rsq[kk], wo, wm, wn);
            }
            double2 term = z;
            if (run) term = lo_piv_term<L>(sA, sB, l.i + 2, wo, wm, wn, up_m, up_n, a.t.sq[l.bm + 1], a.t.sq[l.bn + 1]);
            if (l.lane_ok) {
                sx[l.i][l.slot] = x;
                sw[l.i][l.slot] = wo;
                st[l.i][l.slot] = term;
            }
            __syncthreads();
            if (run) {
                const int64_t pb = (d * P + l.gs) * C + l.pos;  // arr1[2d, p] block: the pivot
                sacc[GL][t] = cfma(wo, a.piv[pb], sacc[GL][t]);
                if (l.bm > 0) sacc[0][t] = cfma(wo, cscale_promoted(a.piv[pb - c0], a.t.sq[l.bm]), sacc[0][t]);
                if (l.bn > 0) sacc[1][t] = cfma(wo, cscale_promoted(a.piv[pb - 1], a.t.sq[l.bn]), sacc[1][t]);
#pragma unroll
                for (int j = 2 * d; j < G; ++j) sacc[j + 2][t] = cfma(wo, sx[j][l.slot], sacc[j + 2][t]);
                // lane i as column j = i: sum_r w_r A_r,j+2
                double2 s = z;
#pragma unroll
                for (int r = 2 * d; r < G; ++r) s = cfma(sw[r][l.slot], sA[r + 2][l.i + 2], s);
                if (l.i == 2 * d) {
                    gbar0 = caxpy(a.t.sq[pd + 1], s, gbar0);
                    double2 gb = z;
#pragma unroll
                    for (int r = 2 * d; r < G; ++r) gb = cadd(gb, st[r][l.slot]);
                    v.adP[pb] = gb;
                } else if (l.pm > 0) {
                    const double2 raw = cscale_promoted(s, a.t.sq[l.pm]);
                    if (d == 0) {
                        v.adO[(l.i * P + l.gs) * C + l.pos] = raw;
                    } else if (l.i == 2 * d + 1) {
                        v.ad2[(a.t.off2[d] + l.f - l.sm) * C + l.pos] = raw;
                    } else {
                        const int64_t en = (a.t.off11[d] + (l.m - d - 1) * a.t.len[d] + l.f - l.sm) * C + l.pos;
                        if (l.i & 1)
                            v.ad1010[en] = raw;
                        else
                            v.ad1001[en] = raw;
                    }
                }
            }
            __syncthreads();
        }
        // the off-diagonal pivots' share of arr0[p]'s adjoint, completed by lo_diag_vjp_kernel
        if (l.lane_ok) sx[l.i][l.slot] = gbar0;
        __syncthreads();
        if (l.act && l.i == 0) {
            double2 gb = z;
#pragma unroll
            for (int d = 0; d < MT; ++d) gb = cadd(gb, sx[2 * d][l.slot]);
            v.adO[l.gs * C + l.pos] = gb;
        }
        __syncthreads();
    }
    lo_vjp_flush<MT>(v.part, sacc);
}

// Diagonal pivots of one weight in reverse (use_diag_pivot,
// singleLeftoverMode_amps.py:224-285; use_diag_pivot_grad, singleLeftoverMode_grad.py:465-556), then arr0[p]'s full adjoint.
template <int MT, bool REORDER>
__global__ __launch_bounds__(DG_THREADS) void lo_diag_vjp_kernel(LeftVjpArgs v, int64_t count, int64_t start, int iters) {
    constexpr int G = 2 * MT, L = MT + 1, GL = 2 * L;
    constexpr int PB = DG_THREADS / G;
    __shared__ double2 sA[GL][GL];
    __shared__ double2 sB[GL];
    __shared__ double2 sx[G][PB + 1];
    __shared__ double2 sw[G][PB + 1];
    __shared__ double2 st[G][PB + 1];
    __shared__ double2 sacc[GL + 1][DG_THREADS];
    const LeftArgs &a = v.a;
    const int t = threadIdx.x;
    for (int e = t; e < GL * GL; e += DG_THREADS)
        sA[e / GL][e % GL] = a.A[src_index<L, REORDER>(e / GL) * GL + src_index<L, REORDER>(e % GL)];
    if (t < GL) sB[t] = a.B[src_index<L, REORDER>(t)];
#pragma unroll
    for (int l = 0; l <= GL; ++l) sacc[l][t] = make_double2(0.0, 0.0);
    __syncthreads();
    const int64_t P = a.t.P, C = a.C;
    const int c0 = a.c0;
    const double2 z = make_double2(0.0, 0.0);

    for (int it = 0; it < iters; ++it) {
        const LoLane<MT> l = lo_lane<MT, PB>(a, count, start, int64_t(it) * gridDim.x + blockIdx.x);
        const int runmask = lo_runmask<MT>(a, l);
        const bool up_m = l.bm + 1 < c0, up_n = l.bn + 1 < c0;
        const bool diag = l.act && (a.t.cut[0] == 1 || l.p[0] < a.t.cut[0] - 1);
        const bool dwrite = diag && l.pm + 1 < l.cm && (l.i != 1 || l.p[0] + 2 < a.t.cut[0]);
        const int64_t base = l.gs * C + l.pos;
        double2 g0 = z, gm = z, gn = z, x = z, wo = z, wm = z, wn = z;
        if (l.act) {
            g0 = a.t.inO[base];
            if (l.bm > 0) gm = a.t.inO[base - c0];
            if (l.bn > 0) gn = a.t.inO[base - 1];
        }
        if (diag && l.pm > 0) x = cscale_promoted(a.t.inD[(l.i * P + l.gs) * C + l.pos], a.t.sq[l.pm]);
        if (dwrite) {
            // consumer p + e_m's diagonal pivot, plus the off-diagonal pivot m of p for arr1[2m, p]
            const double2 *ext = (!(l.i & 1) && (runmask >> l.m & 1)) ? v.adP + (l.m * P + l.gs) * C : nullptr;
            lo_out_adjoints(v.adD + ((l.i ^ 1) * P + l.fw) * C, ext, l.pos, c0, up_m, up_n, a.t.sq[l.pm + 1], a.t.rsq[l.pm + 1], wo,
                            wm, wn);
        }
        double2 term = z;
        if (dwrite) term = lo_piv_term<L>(sA, sB, l.i + 2, wo, wm, wn, up_m, up_n, a.t.sq[l.bm + 1], a.t.sq[l.bn + 1]);
        if (l.lane_ok) {
            sx[l.i][l.slot] = x;
            sw[l.i][l.slot] = wo;
            st[l.i][l.slot] = term;
        }
        __syncthreads();
        if (diag) {
            sacc[GL][t] = cfma(wo, g0, sacc[GL][t]);
            if (l.bm > 0) sacc[0][t] = cfma(wo, cscale_promoted(gm, a.t.sq[l.bm]), sacc[0][t]);
            if (l.bn > 0) sacc[1][t] = cfma(wo, cscale_promoted(gn, a.t.sq[l.bn]), sacc[1][t]);
#pragma unroll
            for (int j = 0; j < G; ++j) sacc[j + 2][t] = cfma(wo, sx[j][l.slot], sacc[j + 2][t]);
            double2 s = z;
#pragma unroll
            for (int r = 0; r < G; ++r) s = cfma(sw[r][l.slot], sA[r + 2][l.i + 2], s);
            if (l.pm > 0) v.adD[(l.i * P + l.gs) * C + l.pos] = cscale_promoted(s, a.t.sq[l.pm]);
        }
        if (l.act && l.i == 0) {
            // cotangent + off-diagonal share + the diagonal pivot's block terms
            double2 gb = cadd(v.cot[int64_t(l.pos) * P + l.f], v.adO[base]);
            if (diag) {
#pragma unroll
                for (int r = 0; r < G; ++r) gb = cadd(gb, st[r][l.slot]);
            }
            v.adO[base] = gb;
        }
        __syncthreads();
    }
    lo_vjp_flush<MT>(v.part, sacc);
}

// The zero-detection block in reverse (singleLeftoverMode_amps.py:330-341):
// columns c0-1 .. 1 one step each, thread m owning row m of the adjoint
// (inbox slot 0 of position 0, entry (m, n) at m * c0 + n), then column 0 by
// thread 0.  Its dA rows 0-1 / dB go to partial row nblk.  blockDim.x is a
// power of two >= max(64, c0), at most 1024.
template <int MT, bool REORDER>
__global__ __launch_bounds__(1024) void lo_init_vjp_kernel(LeftVjpArgs v, int nblk) {
    constexpr int L = MT + 1, GL = 2 * L;
    __shared__ double2 sw[1025];
    __shared__ double2 red[1024];
    const LeftArgs &a = v.a;
    const int c0 = a.c0, m = threadIdx.x;
    const double2 A00 = a.A[src_index<L, REORDER>(0) * GL + src_index<L, REORDER>(0)];
    const double2 A10 = a.A[src_index<L, REORDER>(1) * GL + src_index<L, REORDER>(0)];
    const double2 A11 = a.A[src_index<L, REORDER>(1) * GL + src_index<L, REORDER>(1)];
    const double2 B0 = a.B[src_index<L, REORDER>(0)], B1 = a.B[src_index<L, REORDER>(1)];
    const double2 *val = a.t.inO;
    double2 *adj = v.adO;
    const double2 z = make_double2(0.0, 0.0);
    double2 dB1 = z, dA10 = z, dA11 = z;
    sw[m] = z;
    if (m == 0) sw[blockDim.x] = z;
    __syncthreads();
    for (int k = c0 - 1; k >= 1; --k) {
        // column k = n + 1 from columns n and n - 1
        const int n = k - 1;
        double2 wm = z;
        if (m < c0) wm = cdiv_by(adj[m * c0 + k], a.t.sq[k], a.t.rsq[k]);
        sw[m] = wm;
        __syncthreads();
        if (m < c0) {
            dB1 = cfma(wm, val[m * c0 + n], dB1);
            if (m > 0) dA10 = cfma(wm, cscale_promoted(val[(m - 1) * c0 + n], a.t.sq[m]), dA10);
            if (n > 0) dA11 = cfma(wm, cscale_promoted(val[m * c0 + n - 1], a.t.sq[n]), dA11);
            double2 an = cmul(wm, B1);
            if (m + 1 < c0) an = cfma(cscale_promoted(sw[m + 1], a.t.sq[m + 1]), A10, an);
            adj[m * c0 + n] = cadd(adj[m * c0 + n], an);
            if (n > 0) adj[m * c0 + n - 1] = cfma(cscale_promoted(wm, a.t.sq[n]), A11, adj[m * c0 + n - 1]);
        }
        __syncthreads();
    }
    if (m < c0) sw[m] = adj[m * c0];
    __syncthreads();
    double2 dB0 = z, dA00 = z;
    if (m == 0) {
        // v(k, 0) = (v(k-1, 0) B0 + sqrt(k-1) A00 v(k-2, 0)) / sqrt(k)
        for (int k = c0 - 1; k >= 1; --k) {
            const double2 w = cdiv_by(sw[k], a.t.sq[k], a.t.rsq[k]);
            dB0 = cfma(w, val[(k - 1) * c0], dB0);
            sw[k - 1] = cfma(w, B0, sw[k - 1]);
            if (k >= 2) {
                dA00 = cfma(w, cscale_promoted(val[(k - 2) * c0], a.t.sq[k - 1]), dA00);
                sw[k - 2] = cfma(cscale_promoted(w, a.t.sq[k - 1]), A00, sw[k - 2]);
            }
        }
        *v.dc = sw[0];
    }
    // block sums of the column terms, in a fixed tree
    double2 *prow = v.part + int64_t(nblk) * GL * (GL + 1);
    double2 acc[3] = {dB1, dA10, dA11};
    const int64_t dst[3] = {1 * (GL + 1) + GL, 1 * (GL + 1) + 0, 1 * (GL + 1) + 1};
    for (int q = 0; q < 3; ++q) {
        red[m] = acc[q];
        __syncthreads();
        for (int h = int(blockDim.x) >> 1; h > 0; h >>= 1) {
            if (m < h) red[m] = cadd(red[m], red[m + h]);
            __syncthreads();
        }
        if (m == 0) prow[dst[q]] = red[0];
        __syncthreads();
    }
    if (m == 0) {
        prow[0 * (GL + 1) + GL] = dB0;
        prow[0] = dA00;
    }
}

}  // namespace mm

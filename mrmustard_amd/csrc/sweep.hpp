// LDS-resident slab sweep of the vanilla recurrence (one team per triple).
//
// Recurrence (core.py:97-104, first-pivot rule of core.py:88-94):
//   G_k = ( b_i G_{k-1_i} + A_ii sqrt(k_i-1) G_{k-2_i}
//           + sum_{j>i} A_ij sqrt(k_j) G_{k-1_i-1_j} ) / sqrt(k_i),
//   i = first axis with k_i > 0.
// Every point with k_0 = m >= 1 has pivot axis 0, so the axis-0 slab m (a
// contiguous block of stride[0] points in C order) depends only on slabs m-1
// and m-2: all its points are computed in parallel.  The k_0 = 0 face is the
// same problem one axis down, and so on: "level" d sweeps the slabs k_d = m of
// the sub-lattice {k_0..k_{d-1} = 0}.  The number of dependent steps is
// sum_d (shape_d - 1), the same as an index-sum wavefront sweep, but every
// step is a contiguous, coalesced slab.
//
// LDS holds a ring of NR slabs of the top level (NR = 2 forward, 3 fused) plus
// the per-triple coefficient table (A_ij * sqrt(k), formed exactly as the
// reference's complex*float product), the packed digits of every slab point,
// and sqrt / 1/sqrt tables.  Levels below `top` live inside ring slot 0.
#pragma once

#include "common.hpp"
#include "exact.hpp"
#include "jobs.hpp"

// Phase markers of the sweep (set up, face levels, face VJP, level 0,
// reduction).  Empty in the library; a timing build may define them.
#ifndef MM_PHASE
#define MM_PHASE(i) ((void)0)
#endif

namespace mm {

struct Bases {
    const double2 *A, *b, *c, *g;
    double2 *G, *dA, *db, *dc, *partial;
    int *nf;  // forward-only sweeps that are part of a longer path: raise this flag on a non-finite value (else the team recomputes)
};

// Note a non-finite amplitude in the team's job header (LDS).
__device__ __forceinline__ void note_value(const Job &J, double2 v) {
    if (__builtin_expect(!finite2(v), 0)) const_cast<Job &>(J).nf = 1;
}

struct Tables {
    const double *sq, *rsq;  // sqrt(n), RN(1/sqrt(n)), n < 100000
};

template <bool WAVE>
struct Team {
    int rank, size, nwaves, wave;
    __device__ Team() {
        if (WAVE) {
            rank = threadIdx.x & (MM_WAVE - 1);
            size = MM_WAVE;
            nwaves = 1;
            wave = 0;
        } else {
            rank = threadIdx.x;
            size = blockDim.x;
            nwaves = blockDim.x / MM_WAVE;
            wave = threadIdx.x / MM_WAVE;
        }
    }
    __device__ void sync() const {
        if (WAVE)
            wave_sync();
        else
            __syncthreads();
    }
};

// VJP accumulators for one thread: upper triangle of dA (row-major i <= j),
// db, dc.  Indices are compile-time after unrolling.
template <int D>
struct Acc {
    double2 a[D * (D + 1) / 2];
    double2 b[D];
    double2 c;
    __device__ void zero() {
#pragma unroll
        for (int e = 0; e < D * (D + 1) / 2; ++e) a[e] = make_double2(0.0, 0.0);
#pragma unroll
        for (int e = 0; e < D; ++e) b[e] = make_double2(0.0, 0.0);
        c = make_double2(0.0, 0.0);
    }
    static __device__ constexpr int tri(int i, int j) { return i * D - i * (i - 1) / 2 + (j - i); }
};

__device__ __forceinline__ int digit(uint32_t kp, const Job &J, int j) { return int((kp >> J.kbit[j]) & uint32_t(J.kmask[j])); }

// acc += w * Gv * gk for real w: (w Gv) then a complex multiply-accumulate
// in four fused multiply-adds.
__device__ __forceinline__ double2 wacc(double2 acc, double w, double2 Gv, double2 gk) {
    const double ux = w * Gv.x, uy = w * Gv.y;
    acc.x = fma(ux, gk.x, acc.x);
    acc.x = fma(-uy, gk.y, acc.x);
    acc.y = fma(ux, gk.y, acc.y);
    acc.y = fma(uy, gk.x, acc.y);
    return acc;
}

// VJP terms of one point whose axes < `lo` are zero and whose digits for
// axes > lo come from `kp`; axis `lo` has value `klo`.  `cur` addresses the
// slab holding the point (offset q), `prev1`/`prev2` the slabs one and two
// steps back along axis `lo` (only read when klo >= 1 / >= 2).
// Reads mirror gradients.py:66-75: db_i uses G_{k-1_i}, dA_ii uses
// G_{k-2_i} with weight sqrt(k_i) sqrt(k_i-1)/2, dA_ij uses G_{k-1_i-1_j}.
template <int D>
__device__ __forceinline__ void vjp_point(Acc<D> &acc, const Job &J, const double *sq, uint32_t kp, int lo, int klo,
                                          const double2 *cur, const double2 *prev1, const double2 *prev2, int64_t q,
                                          double2 gk) {
    double2 Gv = cur[q];
    acc.c = cfma(Gv, gk, acc.c);
    // pivot axis lo
    if (klo > 0) {
        double sl = sq[klo];
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (i != lo) continue;
            acc.b[i] = wacc(acc.b[i], sl, prev1[q], gk);
            if (klo > 1) acc.a[Acc<D>::tri(i, i)] = wacc(acc.a[Acc<D>::tri(i, i)], 0.5 * sl * sq[klo - 1], prev2[q], gk);
#pragma unroll
            for (int j = i + 1; j < D; ++j) {
                int kj = digit(kp, J, j);
                if (kj > 0) acc.a[Acc<D>::tri(i, j)] = wacc(acc.a[Acc<D>::tri(i, j)], sl * sq[kj], prev1[q - J.stride[j]], gk);
            }
        }
    }
    // pivot axes i > lo (points inside the same slab)
#pragma unroll
    for (int i = 1; i < D; ++i) {
        if (i <= lo) continue;
        int ki = digit(kp, J, i);
        if (ki == 0) continue;
        double si = sq[ki];
        int64_t pi = q - J.stride[i];
        acc.b[i] = wacc(acc.b[i], si, cur[pi], gk);
        if (ki > 1) acc.a[Acc<D>::tri(i, i)] = wacc(acc.a[Acc<D>::tri(i, i)], 0.5 * si * sq[ki - 1], cur[pi - J.stride[i]], gk);
#pragma unroll
        for (int j = i + 1; j < D; ++j) {
            int kj = digit(kp, J, j);
            if (kj > 0) acc.a[Acc<D>::tri(i, j)] = wacc(acc.a[Acc<D>::tri(i, j)], si * sq[kj], cur[pi - J.stride[j]], gk);
        }
    }
}

// Sum the accumulators over the team and write either the final gradients
// (gradients.py:80-82: symmetrised dA, db, dc = sum(G*g)/c) or the raw sums.
// Returns the team's sum of G * dL/dG (before the division by c) to every
// thread of the team.
template <int D, bool WAVE>
__device__ double2 reduce_and_store(Acc<D> &acc, const Job &J, const Bases &P, const Team<WAVE> &tm, double2 *red, double2 c) {
    constexpr int NA = D * (D + 1) / 2 + D + 1;
    double2 v[NA];
#pragma unroll
    for (int e = 0; e < D * (D + 1) / 2; ++e) v[e] = acc.a[e];
#pragma unroll
    for (int e = 0; e < D; ++e) v[D * (D + 1) / 2 + e] = acc.b[e];
    v[NA - 1] = acc.c;
    wave_sum_dpp<NA>(v);
    const int lane = threadIdx.x & (MM_WAVE - 1);
    if (!WAVE) {
        // lane 63 of every wave publishes; thread e then sums accumulator e
        // over the waves in wave order, and thread 0 gathers the totals
        if (lane == MM_WAVE - 1) {
#pragma unroll
            for (int e = 0; e < NA; ++e) red[tm.wave * NA + e] = v[e];
        }
        __syncthreads();
        if (threadIdx.x < NA) {
            double2 t = red[threadIdx.x];
            for (int w = 1; w < tm.nwaves; ++w) t = cadd(t, red[w * NA + threadIdx.x]);
            red[threadIdx.x] = t;
        }
        __syncthreads();
        const double2 total = red[NA - 1];
        if (threadIdx.x != 0) return total;
#pragma unroll
        for (int e = 0; e < NA; ++e) v[e] = red[e];
    } else {
        const double2 total = make_double2(__shfl(v[NA - 1].x, MM_WAVE - 1, MM_WAVE), __shfl(v[NA - 1].y, MM_WAVE - 1, MM_WAVE));
        if (lane != MM_WAVE - 1) return total;
    }
    if (J.flags & JOB_WRITE_PARTIAL) {
        double2 *out = P.partial + J.p_off;
#pragma unroll
        for (int e = 0; e < NA; ++e) out[e] = v[e];
        return v[NA - 1];
    }
    double2 *dA = P.dA + J.a_off;
#pragma unroll
    for (int i = 0; i < D; ++i) {
#pragma unroll
        for (int j = 0; j < D; ++j) {
            double2 m = (j >= i) ? v[Acc<D>::tri(i, j)] : v[Acc<D>::tri(j, i)];
            // (M + M^T)/2 with the strictly lower part of M zero: diagonal M_ii, off-diagonal M_ij/2
            dA[i * D + j] = (i == j) ? m : make_double2(m.x * 0.5, m.y * 0.5);
        }
    }
    double2 *db = P.db + J.b_off;
#pragma unroll
    for (int i = 0; i < D; ++i) db[i] = v[D * (D + 1) / 2 + i];
    P.dc[J.c_off] = cdiv_py(v[NA - 1], c);
    return v[NA - 1];
}

// Level 0 of a triple swept whole in LDS, with every slab point of the team
// held in registers (RM points per thread): a point's own values of slabs
// m-1 and m-2, its coefficients A_0j sqrt(k_j) and digits stay in registers
// across steps, so LDS is read only for the neighbours k - 1_j.  Fused: the
// cotangent of slab m+1 is loaded during step m, and the slab-m neighbours
// read by the VJP of step m are kept for the forward of step m+1.  Two LDS
// slab buffers suffice.
//
// The loop body is branch-free so that all loads of a step can be issued
// together: a term whose digit is zero reads the point itself and is
// multiplied by a zero coefficient, which is what the reference does too
// (it adds A_ij * SQRT[0] * G[...], core.py:101-103); points past the slab
// end compute on a clamped index, store nothing and carry a zero cotangent.
template <int D, bool FUSED, bool WAVE, int RM>
__device__ __forceinline__ void level0_reg(const Job &J, const Team<WAVE> &tm, double2 *ring, const uint32_t *kv,
                                           const double *sq, const double *rsq, const double2 *bv, const double2 *A, double2 *G,
                                           const double2 *g, const double2 *coef, Acc<D> &acc) {
    constexpr bool CJ_REG = FUSED || (D - 1) * RM <= 8;
    const int smax = J.smax;
    const int S = J.S;
    const int s0 = J.shape[0];
    int st[D];
#pragma unroll
    for (int j = 0; j < D; ++j) st[j] = int(J.stride[j]);
    // rounds of `tm.size` points actually needed (uniform across the team)
    const int nr = (S + tm.size - 1) / tm.size;
    int q[RM];
    bool act[RM];
    int kd[RM][D];
    double sqk[RM][D];
    double2 cj[RM][D];
    double2 nb[RM][D];
    double2 p1v[RM], p2v[RM], gcur[RM], gnext[RM];
    const double2 b0 = bv[0];
    const double2 a00 = A[0];
    const double2 zero = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < RM; ++r) {
        const int qq = tm.rank + r * tm.size;
        act[r] = qq < S;
        q[r] = act[r] ? qq : S - 1;
        const uint32_t kp = kv[q[r]];
#pragma unroll
        for (int j = 1; j < D; ++j) {
            kd[r][j] = digit(kp, J, j);
            sqk[r][j] = sq[kd[r][j]];
            cj[r][j] = CJ_REG ? cscale_promoted(A[j], sqk[r][j]) : zero;
            nb[r][j] = zero;
        }
        p1v[r] = ring[q[r]];
        p2v[r] = zero;
        gcur[r] = zero;
        if (FUSED && s0 > 1) gcur[r] = g[int64_t(S) + q[r]];  // masked where used
        gnext[r] = zero;
    }
    for (int m = 1; m < s0; ++m) {
        double2 *cur = ring + (m & 1) * S;
        const double2 *p1 = ring + ((m - 1) & 1) * S;
        const double sm = sq[m], rm = rsq[m];
        const double sm1 = sq[m - 1];
        const double2 c00 = cscale_promoted(a00, sm1);
        const int64_t gbase = int64_t(m) * S;
        if (FUSED && m + 1 < s0) {
#pragma unroll
            for (int r = 0; r < RM; ++r) {
                if (r >= nr) break;
                gnext[r] = g[gbase + S + q[r]];  // masked where used, so nothing waits for it here
            }
        }
        double2 vnew[RM];
#pragma unroll
        for (int r = 0; r < RM; ++r) {
            vnew[r] = zero;
            if (r >= nr) break;
            double2 v = cmul(b0, p1v[r]);
            v = cadd(v, cmul(c00, p2v[r]));
#pragma unroll
            for (int j = 1; j < D; ++j) {
                if (!FUSED || m == 1) nb[r][j] = p1[q[r] - (kd[r][j] > 0 ? st[j] : 0)];
                const double2 cf = CJ_REG ? cj[r][j] : coef[coef_row(D, 0, 0, j) * smax + kd[r][j]];
                v = cadd(v, cmul(cf, nb[r][j]));
            }
            vnew[r] = cdiv_by(v, sm, rm);
        }
#pragma unroll
        for (int r = 0; r < RM; ++r) {
            if (r >= nr) break;
            if (act[r]) {
                cur[q[r]] = vnew[r];
                G[gbase + q[r]] = vnew[r];
                if (!FUSED) note_value(J, vnew[r]);
            }
        }
        tm.sync();
        if (FUSED) {
            const double w00 = 0.5 * sm * sm1;
#pragma unroll
            for (int r = 0; r < RM; ++r) {
                if (r >= nr) break;
                const double2 gk = csel(act[r], gcur[r], zero);
                acc.c = cfma(vnew[r], gk, acc.c);
                acc.b[0] = wacc(acc.b[0], sm, p1v[r], gk);
                acc.a[0] = wacc(acc.a[0], w00, p2v[r], gk);
#pragma unroll
                for (int j = 1; j < D; ++j) acc.a[Acc<D>::tri(0, j)] = wacc(acc.a[Acc<D>::tri(0, j)], sm * sqk[r][j], nb[r][j], gk);
#pragma unroll
                for (int i = 1; i < D; ++i) {
                    const int ki = kd[r][i];
                    const double si = sqk[r][i];
                    const int pi = q[r] - (ki > 0 ? st[i] : 0);
                    const double2 t = cur[pi];
                    acc.b[i] = wacc(acc.b[i], si, t, gk);
                    nb[r][i] = t;  // slab-m neighbour: the next step's forward needs it
                    const double2 t2 = cur[pi - (ki > 1 ? st[i] : 0)];
                    acc.a[Acc<D>::tri(i, i)] = wacc(acc.a[Acc<D>::tri(i, i)], 0.5 * si * sq[ki > 0 ? ki - 1 : 0], t2, gk);
#pragma unroll
                    for (int j = i + 1; j < D; ++j) {
                        const bool ok = ki > 0 && kd[r][j] > 0;
                        const double2 t3 = cur[ok ? pi - st[j] : q[r]];
                        acc.a[Acc<D>::tri(i, j)] = wacc(acc.a[Acc<D>::tri(i, j)], si * sqk[r][j], t3, gk);
                    }
                }
                gcur[r] = gnext[r];
            }
        }
#pragma unroll
        for (int r = 0; r < RM; ++r) {
            p2v[r] = p1v[r];
            p1v[r] = vnew[r];
        }
    }
}

// Straight-line sweep of one level with NR points per participating thread
// (NR compile-time, so the loads and arithmetic of all points of a step form
// one basic block).  Used for triples with D <= 3 swept whole in LDS.
//   L = 0: the level-0 slabs in the two-slot ring, with the fused VJP (as
//          level0_reg);
//   L > 0: a face level of the sub-lattice {k_0..k_{L-1} = 0}, inside ring
//          slot 0 (slab m at offset m * stride[L]).
// `rank`/`tsize` number the participating threads; smode is the ordering
// between steps: 0 = a single lane (no ordering needed), 1 = one wave,
// 2 = the team barrier.
template <int D, bool FUSED, bool WAVE, int L, int NR>
__device__ __forceinline__ void level_straight(const Job &J, const Team<WAVE> &tm, int rank, int tsize, int smode, double2 *ring,
                                               const uint32_t *kv, const double *sq, const double *rsq, const double2 *bv,
                                               const double2 *A, double2 *G, const double2 *g, Acc<D> &acc) {
    constexpr bool TOP = (L == 0);
    constexpr bool VJP = TOP && FUSED;
    const int Sd = int(J.stride[L]);
    const int sd = J.shape[L];
    const int S = J.S;
    int st[D];
#pragma unroll
    for (int j = 0; j < D; ++j) st[j] = int(J.stride[j]);
    int q[NR];
    bool act[NR];
    int kd[NR][D];
    double sqk[NR][D];
    double2 cj[NR][D];
    double2 nb[NR][D];
    double2 p1v[NR], p2v[NR], gcur[NR], gnext[NR];
    const double2 bL = bv[L];
    const double2 aLL = A[L * D + L];
    const double2 zero = make_double2(0.0, 0.0);
#pragma unroll
    for (int r = 0; r < NR; ++r) {
        const int qq = rank + r * tsize;
        act[r] = qq < Sd;
        q[r] = act[r] ? qq : Sd - 1;
        const uint32_t kp = kv[q[r]];
#pragma unroll
        for (int j = L + 1; j < D; ++j) {
            kd[r][j] = digit(kp, J, j);
            sqk[r][j] = sq[kd[r][j]];
            cj[r][j] = cscale_promoted(A[L * D + j], sqk[r][j]);
            nb[r][j] = zero;
        }
        p1v[r] = ring[q[r]];
        p2v[r] = zero;
        gcur[r] = zero;
        gnext[r] = zero;
        if (VJP && sd > 1) gcur[r] = g[int64_t(S) + q[r]];  // masked where used
    }
    for (int m = 1; m < sd; ++m) {
        double2 *cur = TOP ? ring + (m & 1) * S : ring + m * Sd;
        const double2 *p1 = TOP ? ring + ((m - 1) & 1) * S : ring + (m - 1) * Sd;
        const double sm = sq[m], rm = rsq[m];
        const double sm1 = sq[m - 1];
        const double2 cLL = cscale_promoted(aLL, sm1);
        const int64_t gbase = int64_t(m) * Sd;
        if (VJP && m + 1 < sd) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                gnext[r] = g[gbase + Sd + q[r]];  // masked where used, so nothing waits for it here
            }
        }
        double2 v[NR];
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            v[r] = cmul(bL, p1v[r]);
            v[r] = cadd(v[r], cmul(cLL, p2v[r]));
#pragma unroll
            for (int j = L + 1; j < D; ++j) {
                if (smode == 1) {
                    // one wave holds the whole slab: the neighbour is another lane's register
                    const int src = q[r] - (kd[r][j] > 0 ? st[j] : 0);
                    nb[r][j] = make_double2(__shfl(p1v[r].x, src, MM_WAVE), __shfl(p1v[r].y, src, MM_WAVE));
                } else if (!VJP || m == 1) {
                    nb[r][j] = p1[q[r] - (kd[r][j] > 0 ? st[j] : 0)];
                }
                v[r] = cadd(v[r], cmul(cj[r][j], nb[r][j]));
            }
        }
        cdiv_many<NR>(v, sm, rm);
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            if (act[r]) {
                cur[q[r]] = v[r];
                G[gbase + q[r]] = v[r];
                if (!VJP) note_value(J, v[r]);
            }
        }
        // VJP terms that need no other lane's slab-m value go before the
        // barrier: the cotangent is pre-scaled by sqrt(k_i) per pivot axis i.
        double2 u0[NR];
        if (VJP) {
            const double h00 = 0.5 * sm1;
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double2 gk = csel(act[r], gcur[r], zero);
                u0[r] = make_double2(sm * gk.x, sm * gk.y);
                acc.c = cfma(v[r], gk, acc.c);
                acc.b[0] = cfma(p1v[r], u0[r], acc.b[0]);
                acc.a[0] = cfma(make_double2(h00 * p2v[r].x, h00 * p2v[r].y), u0[r], acc.a[0]);
#pragma unroll
                for (int j = 1; j < D; ++j) {
                    const double w = sqk[r][j];
                    acc.a[Acc<D>::tri(0, j)] = cfma(make_double2(w * nb[r][j].x, w * nb[r][j].y), u0[r], acc.a[Acc<D>::tri(0, j)]);
                }
            }
        }
        if (smode == 2)
            tm.sync();
        if (VJP) {
#pragma unroll
            for (int r = 0; r < NR; ++r) {
                const double2 gk = csel(act[r], gcur[r], zero);
#pragma unroll
                for (int i = 1; i < D; ++i) {
                    const int ki = kd[r][i];
                    const double si = sqk[r][i];
                    const double2 ui = make_double2(si * gk.x, si * gk.y);
                    const int pi = q[r] - (ki > 0 ? st[i] : 0);
                    const double2 t = cur[pi];
                    acc.b[i] = cfma(t, ui, acc.b[i]);
                    nb[r][i] = t;  // slab-m neighbour: the next step's forward needs it
                    const double2 t2 = cur[pi - (ki > 1 ? st[i] : 0)];
                    const double hi = 0.5 * sq[ki > 0 ? ki - 1 : 0];
                    acc.a[Acc<D>::tri(i, i)] = cfma(make_double2(hi * t2.x, hi * t2.y), ui, acc.a[Acc<D>::tri(i, i)]);
#pragma unroll
                    for (int j = i + 1; j < D; ++j) {
                        const bool ok = ki > 0 && kd[r][j] > 0;
                        const double2 t3 = cur[ok ? pi - st[j] : q[r]];
                        const double w = sqk[r][j];
                        acc.a[Acc<D>::tri(i, j)] = cfma(make_double2(w * t3.x, w * t3.y), ui, acc.a[Acc<D>::tri(i, j)]);
                    }
                }
                gcur[r] = gnext[r];
            }
        }
#pragma unroll
        for (int r = 0; r < NR; ++r) {
            p2v[r] = p1v[r];
            p1v[r] = v[r];
        }
    }
}

// Pick the compile-time point count for level L and run it.
template <int D, bool FUSED, bool WAVE, int L, int RM>
__device__ __forceinline__ void level_dispatch(int nr, const Job &J, const Team<WAVE> &tm, int rank, int tsize, int smode,
                                               double2 *ring, const uint32_t *kv, const double *sq, const double *rsq,
                                               const double2 *bv, const double2 *A, double2 *G, const double2 *g, Acc<D> &acc) {
    switch (nr) {
        case 1:
            level_straight<D, FUSED, WAVE, L, 1>(J, tm, rank, tsize, smode, ring, kv, sq, rsq, bv, A, G, g, acc);
            break;
        case 2:
            if constexpr (RM >= 2) level_straight<D, FUSED, WAVE, L, 2>(J, tm, rank, tsize, smode, ring, kv, sq, rsq, bv, A, G, g, acc);
            break;
        case 3:
            if constexpr (RM >= 3) level_straight<D, FUSED, WAVE, L, 3>(J, tm, rank, tsize, smode, ring, kv, sq, rsq, bv, A, G, g, acc);
            break;
        case 4:
            if constexpr (RM >= 4) level_straight<D, FUSED, WAVE, L, 4>(J, tm, rank, tsize, smode, ring, kv, sq, rsq, bv, A, G, g, acc);
            break;
        case 5:
            if constexpr (RM >= 5) level_straight<D, FUSED, WAVE, L, 5>(J, tm, rank, tsize, smode, ring, kv, sq, rsq, bv, A, G, g, acc);
            break;
        default:
            break;
    }
}

// Face level L (> 0) of a triple with D <= 3 swept whole in LDS: a single
// lane when the slab is one point, wave 0 alone when it fits one wave,
// otherwise the whole team.  Ends with every thread of the team past a
// barrier.
template <int D, bool FUSED, bool WAVE, int L, int RM>
__device__ __forceinline__ void face_level(const Job &J, const Team<WAVE> &tm, double2 *ring, const uint32_t *kv, const double *sq,
                                           const double *rsq, const double2 *bv, const double2 *A, double2 *G, Acc<D> &acc) {
    const int Sd = int(J.stride[L]);
    if (J.shape[L] <= 1) return;
    if (Sd == 1) {
        if (tm.rank == 0) level_dispatch<D, false, WAVE, L, RM>(1, J, tm, 0, 1, 0, ring, kv, sq, rsq, bv, A, G, nullptr, acc);
        tm.sync();
    } else if (WAVE || Sd > MM_WAVE) {
        const int nr = (Sd + tm.size - 1) / tm.size;
        level_dispatch<D, false, WAVE, L, RM>(nr, J, tm, tm.rank, tm.size, 2, ring, kv, sq, rsq, bv, A, G, nullptr, acc);
    } else {
        if (tm.wave == 0) level_dispatch<D, false, WAVE, L, RM>(1, J, tm, tm.rank, MM_WAVE, 1, ring, kv, sq, rsq, bv, A, G, nullptr, acc);
        tm.sync();
    }
}

// Per-triple shift applied to a template job (uniform batches): offsets of
// triple k are the template's plus k times these strides.
struct JobShift {
    int64_t k, ua, ub, uc, ug;
};

// Copy a job descriptor (from the kernel arguments or the queue table) into
// the head of the team's LDS region and apply a uniform-batch shift.  Ends
// with the team's barrier, so the sweep can read it.
template <bool WAVE>
__device__ __forceinline__ void load_job(char *lds, const Job *src, const JobShift sh) {
    const Team<WAVE> tm;
    Job *dst = reinterpret_cast<Job *>(lds);
    const uint32_t *s32 = reinterpret_cast<const uint32_t *>(src);
    uint32_t *d32 = reinterpret_cast<uint32_t *>(dst);
    for (int t = tm.rank; t < int(sizeof(Job) / 4); t += tm.size) d32[t] = s32[t];
    tm.sync();
    if (tm.rank == 0 && sh.k > 0) {
        dst->a_off += sh.k * sh.ua;
        dst->b_off += sh.k * sh.ub;
        dst->c_off += sh.k * sh.uc;
        dst->g_off += sh.k * sh.ug;
    }
    tm.sync();
}

// The sweep.  D is the number of indices; the team is a workgroup (WAVE =
// false) or a single wave (WAVE = true) that owns `lds` exclusively.  The job
// descriptor sits in the head of the team's LDS region (load_job), so that
// fields indexed by a run-time level are LDS reads rather than private memory.
template <int D, bool FUSED, bool WAVE, int RM>
__device__ __forceinline__ void sweep_lds(const Bases P, const Tables T, char *lds) {
    const Team<WAVE> tm;
    const Job &J = *reinterpret_cast<const Job *>(lds);
    MM_PHASE(0);
    const int top = J.top;
    const int S = J.S;
    const int smax = J.smax;
    const int nw = WAVE ? 1 : tm.nwaves;
    const bool regp = top == 0 && reg_path(S, nw, RM);
    const LdsLayout L = lds_layout(D, top, S, smax, FUSED, nw, RM);
    double2 *ring = reinterpret_cast<double2 *>(lds + L.ring);
    uint32_t *kv = reinterpret_cast<uint32_t *>(lds + L.kv);
    double2 *coef = reinterpret_cast<double2 *>(lds + L.coef);
    double *sq = reinterpret_cast<double *>(lds + L.sq);
    double *rsq = reinterpret_cast<double *>(lds + L.rsq);
    double2 *bv = reinterpret_cast<double2 *>(lds + L.bvec);
    double2 *am = reinterpret_cast<double2 *>(lds + L.amat);
    double2 *red = reinterpret_cast<double2 *>(lds + L.red);

    const double2 *Ag = P.A + J.a_off;
    const double2 *A = am;  // the sweep reads A from LDS
    double2 *G = P.G + J.g_off;
    const double2 *g = FUSED ? P.g + J.g_off : nullptr;

    // ---- per-triple tables
    for (int t = tm.rank; t < smax; t += tm.size) {
        sq[t] = T.sq[t];
        rsq[t] = T.rsq[t];
    }
    if (tm.rank < D) bv[tm.rank] = P.b[J.b_off + tm.rank];
    if (tm.rank == 0) bv[D] = P.c[J.c_off];
    const int lda = J.lda > 0 ? J.lda : D;
    for (int t = tm.rank; t < D * D; t += tm.size) am[t] = Ag[(t / D) * lda + t % D];
    const int nrows = coef_rows(D, top);
    for (int t = tm.rank; t < nrows * smax; t += tm.size) {
        int r = t / smax, k = t - r * smax;
        int i = top, rr = r;
        while (rr >= D - i) {
            rr -= D - i;
            ++i;
        }
        int j = i + rr;
        double2 a = Ag[i * lda + j];
        coef[t] = (k < J.shape[j]) ? cscale_promoted(a, T.sq[k]) : make_double2(0.0, 0.0);
    }
    for (int f = tm.rank; f < S; f += tm.size) {
        uint32_t kp = 0;
        int rem = f;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (j <= top) continue;
            const int sj = int(J.stride[j]);
            const int dj = rem / sj;
            rem -= dj * sj;
            kp |= uint32_t(dj) << J.kbit[j];
        }
        kv[f] = kp;
    }
    if (tm.rank == 0) {
        ring[0] = P.c[J.c_off];
        G[0] = P.c[J.c_off];
        if (!FUSED) note_value(J, P.c[J.c_off]);
    }
    tm.sync();

    const double2 c = bv[D];
    Acc<D> acc;
    acc.zero();
    MM_PHASE(1);

    // straight-line path: D <= 3, whole triple in LDS, level 0 in registers
    constexpr bool STRAIGHT_OK = D <= 3;
    const bool straight = STRAIGHT_OK && regp;
    if (straight) {
        if constexpr (STRAIGHT_OK) {
            if constexpr (D >= 3) face_level<D, FUSED, WAVE, 2, RM>(J, tm, ring, kv, sq, rsq, bv, A, G, acc);
            if constexpr (D >= 2) face_level<D, FUSED, WAVE, 1, RM>(J, tm, ring, kv, sq, rsq, bv, A, G, acc);
        }
    } else {
        // ---- levels d = D-1 .. top+1 inside ring slot 0
        for (int d = D - 1; d > top; --d) {
            const int Sd = int(J.stride[d]);
            const int sd = J.shape[d];
            if (sd <= 1) continue;
            const double2 bd = bv[d];
            const double2 *cdd = coef + coef_row(D, top, d, d) * smax;
            if (Sd == 1) {
                // every later axis has size 1: a one-index chain, kept in one lane's registers
                if (tm.rank == 0) {
                    double2 m1 = ring[0], m2 = make_double2(0.0, 0.0);
                    for (int m = 1; m < sd; ++m) {
                        double2 v = cmul(bd, m1);
                        if (m >= 2) v = cadd(v, cmul(cdd[m - 1], m2));
                        v = cdiv_by(v, sq[m], rsq[m]);
                        ring[m] = v;
                        G[m] = v;
                        if (!FUSED) note_value(J, v);
                        m2 = m1;
                        m1 = v;
                    }
                }
                tm.sync();
                continue;
            }
            // slabs of at most one wave's width are swept by wave 0 alone, with
            // wave-level ordering between steps; the other waves wait at the end
            const bool one_wave = !WAVE && Sd <= MM_WAVE;
            if (one_wave && tm.wave != 0) {
                tm.sync();
                continue;
            }
            const int stride_q = one_wave ? MM_WAVE : tm.size;
            for (int m = 1; m < sd; ++m) {
                const double sm = sq[m], rm = rsq[m];
                for (int q = tm.rank; q < Sd; q += stride_q) {
                    const int f = m * Sd + q;
                    const uint32_t kp = kv[f];
                    double2 v = cmul(bd, ring[f - Sd]);
                    if (m >= 2) v = cadd(v, cmul(cdd[m - 1], ring[f - 2 * Sd]));
    #pragma unroll
                    for (int j = 1; j < D; ++j) {
                        if (j <= d) continue;
                        const int kj = digit(kp, J, j);
                        if (kj > 0) v = cadd(v, cmul(coef[coef_row(D, top, d, j) * smax + kj], ring[f - Sd - int(J.stride[j])]));
                    }
                    v = cdiv_by(v, sm, rm);
                    ring[f] = v;
                    G[f] = v;
                    if (!FUSED) note_value(J, v);
                }
                if (one_wave)
                    wave_sync();
                else
                    tm.sync();
            }
            if (one_wave) __syncthreads();
        }
    }
    MM_PHASE(2);
    // ---- VJP of the sub-lattice below `top` (all of it sits in ring slot 0)
    if (FUSED) {
        for (int f = tm.rank; f < S; f += tm.size) vjp_point<D>(acc, J, sq, kv[f], top, 0, ring, ring, ring, f, g[f]);
    }

    MM_PHASE(3);
    // ---- level `top`
    if (straight) {
        if constexpr (STRAIGHT_OK) {
            const int nr = (S + tm.size - 1) / tm.size;
            level_dispatch<D, FUSED, WAVE, 0, RM>(nr, J, tm, tm.rank, tm.size, 2, ring, kv, sq, rsq, bv, A, G, g, acc);
        }
    } else if (regp) {
        level0_reg<D, FUSED, WAVE, RM>(J, tm, ring, kv, sq, rsq, bv, A, G, g, coef, acc);
    } else {
        const int NR = FUSED ? 3 : 2;
        const int s0 = J.shape[top];
        const double2 b0 = bv[top];
        const double2 *c00 = coef + coef_row(D, top, top, top) * smax;
        for (int m = 1; m < s0; ++m) {
            double2 *cur = ring + (m % NR) * S;
            const double2 *p1 = ring + ((m - 1) % NR) * S;
            const double2 *p2 = ring + ((m + NR - 2) % NR) * S;
            const double sm = sq[m], rm = rsq[m];
            const int64_t gbase = int64_t(m) * S;
            for (int q = tm.rank; q < S; q += tm.size) {
                const uint32_t kp = kv[q];
                double2 v = cmul(b0, p1[q]);
                if (m >= 2) v = cadd(v, cmul(c00[m - 1], p2[q]));
#pragma unroll
                for (int j = 1; j < D; ++j) {
                    if (j <= top) continue;
                    const int kj = digit(kp, J, j);
                    if (kj > 0) v = cadd(v, cmul(coef[coef_row(D, top, top, j) * smax + kj], p1[q - int(J.stride[j])]));
                }
                v = cdiv_by(v, sm, rm);
                cur[q] = v;
                G[gbase + q] = v;
                if (!FUSED) note_value(J, v);
            }
            tm.sync();
            if (FUSED) {
                for (int q = tm.rank; q < S; q += tm.size) vjp_point<D>(acc, J, sq, kv[q], top, m, cur, p1, p2, q, g[gbase + q]);
            }
        }
    }

    MM_PHASE(4);
    double2 total = make_double2(0.0, 0.0);
    if (FUSED) total = reduce_and_store<D, WAVE>(acc, J, P, tm, red, c);
    tm.sync();
    MM_PHASE(5);

    // ---- non-finite values: rerun the triple in the reference's arithmetic
    // (the fused sum of G * dL/dG is non-finite whenever G or dL/dG holds a
    // non-finite value)
    const bool bad = FUSED ? (!finite2(total) && top == 0 && !(J.flags & JOB_WRITE_PARTIAL)) : (J.nf != 0);
    if (__builtin_expect(bad, 0)) {
        if (!FUSED && P.nf) {
            if (tm.rank == 0) atomicOr(P.nf, 1);
        } else {
            Lat lat;
            lat.D = D;
            for (int j = 0; j < 8; ++j) {
                lat.shape[j] = J.shape[j];
                lat.stride[j] = J.stride[j];
            }
            lat.N = int64_t(J.shape[0]) * J.stride[0];
            const double2 *bg = P.b + J.b_off;
            if (WAVE) {
                const WaveOps ops;
                exact_forward(ops, lat, top, Ag, bg, c, G, T.sq, lda);
                if (FUSED) exact_vjp(ops, lat, G, g, c, P.dA + J.a_off, P.db + J.b_off, P.dc + J.c_off, T.sq);
            } else {
                const BlockOps ops(red);
                exact_forward(ops, lat, top, Ag, bg, c, G, T.sq, lda);
                if (FUSED) exact_vjp(ops, lat, G, g, c, P.dA + J.a_off, P.db + J.b_off, P.dc + J.c_off, T.sq);
            }
        }
        tm.sync();
    }
}

// One-index chain in a single lane (D = 1: core.py:108-122 with D = 1).
// Fields are passed as scalars so that nothing is copied to private memory.
template <bool FUSED>
__device__ __forceinline__ void chain_lane(int s0, int64_t a_off, int64_t b_off, int64_t c_off, int64_t g_off, const Bases P, const Tables T) {
    const double2 a = P.A[a_off];
    const double2 b = P.b[b_off];
    const double2 c = P.c[c_off];
    double2 *G = P.G + g_off;
    const double2 *g = FUSED ? P.g + g_off : nullptr;
    double2 m1 = c, m2 = make_double2(0.0, 0.0);
    G[0] = c;
    double2 ab = make_double2(0.0, 0.0), bb = make_double2(0.0, 0.0), cc = make_double2(0.0, 0.0);
    if (FUSED) cc = cfma(c, g[0], cc);
    for (int n = 1; n < s0; ++n) {
        const double sn = T.sq[n];
        double2 v = cmul(b, m1);
        if (n >= 2) v = cadd(v, cmul(cscale_promoted(a, T.sq[n - 1]), m2));
        v = cdiv_by(v, sn, T.rsq[n]);
        G[n] = v;
        if (FUSED) {
            const double2 gk = g[n];
            bb = wacc(bb, sn, m1, gk);
            if (n >= 2) ab = wacc(ab, 0.5 * sn * T.sq[n - 1], m2, gk);
            cc = cfma(v, gk, cc);
        }
        m2 = m1;
        m1 = v;
    }
    if (FUSED) {
        P.dA[a_off] = ab;
        P.db[b_off] = bb;
        P.dc[c_off] = cdiv_py(cc, c);
        // the forward of one index has no zero-weight terms; the VJP has
        if (__builtin_expect(!finite2(cc), 0)) {
            Lat lat;
            lat.D = 1;
            lat.shape[0] = s0;
            lat.stride[0] = 1;
            lat.N = s0;
            exact_vjp(LaneOps(), lat, G, g, c, P.dA + a_off, P.db + b_off, P.dc + c_off, T.sq);
        }
    }
}

}  // namespace mm

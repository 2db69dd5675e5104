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
#include "jobs.hpp"

namespace mm {

struct Bases {
    const double2 *A, *b, *c, *g;
    double2 *G, *dA, *db, *dc, *partial;
};

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

// G*gk products accumulated with weight w:  acc += w * (Gv * gk)
__device__ __forceinline__ double2 wacc(double2 acc, double w, double2 Gv, double2 gk) {
    double2 p = make_double2(Gv.x * gk.x - Gv.y * gk.y, Gv.x * gk.y + Gv.y * gk.x);
    return caxpy(w, p, acc);
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
template <int D, bool WAVE>
__device__ void reduce_and_store(Acc<D> &acc, const Job &J, const Bases &P, const Team<WAVE> &tm, double2 *red, double2 c) {
    constexpr int NA = D * (D + 1) / 2 + D + 1;
    double2 v[NA];
#pragma unroll
    for (int e = 0; e < D * (D + 1) / 2; ++e) v[e] = acc.a[e];
#pragma unroll
    for (int e = 0; e < D; ++e) v[D * (D + 1) / 2 + e] = acc.b[e];
    v[NA - 1] = acc.c;
#pragma unroll
    for (int e = 0; e < NA; ++e) v[e] = wave_sum(v[e]);
    if (!WAVE) {
        int lane = threadIdx.x & (MM_WAVE - 1);
        if (lane == 0) {
#pragma unroll
            for (int e = 0; e < NA; ++e) red[tm.wave * NA + e] = v[e];
        }
        __syncthreads();
        if (threadIdx.x == 0) {
            for (int w = 1; w < tm.nwaves; ++w)
#pragma unroll
                for (int e = 0; e < NA; ++e) v[e] = cadd(v[e], red[w * NA + e]);
        }
    }
    if (tm.rank != 0) return;
    if (J.flags & JOB_WRITE_PARTIAL) {
        double2 *out = P.partial + J.p_off;
#pragma unroll
        for (int e = 0; e < NA; ++e) out[e] = v[e];
        return;
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
}

// Per-triple shift applied to a template job (uniform batches): offsets of
// triple k are the template's plus k times these strides.
struct JobShift {
    int64_t k, ua, ub, uc, ug;
};

// The sweep.  D is the number of indices; the team is a workgroup (WAVE =
// false) or a single wave (WAVE = true) that owns `lds` exclusively.  The job
// descriptor is first copied into the head of the team's LDS region, so that
// fields indexed by a run-time level are LDS reads rather than private memory.
template <int D, bool FUSED, bool WAVE>
__device__ void sweep_lds(const Job *src, const JobShift sh, const Bases &P, const Tables &T, char *lds) {
    const Team<WAVE> tm;
    {
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
    const Job &J = *reinterpret_cast<const Job *>(lds);
    const int top = J.top;
    const int S = J.S;
    const int smax = J.smax;
    const LdsLayout L = lds_layout(D, top, S, smax, FUSED, WAVE ? 1 : tm.nwaves);
    double2 *ring = reinterpret_cast<double2 *>(lds + L.ring);
    uint32_t *kv = reinterpret_cast<uint32_t *>(lds + L.kv);
    double2 *coef = reinterpret_cast<double2 *>(lds + L.coef);
    double *sq = reinterpret_cast<double *>(lds + L.sq);
    double *rsq = reinterpret_cast<double *>(lds + L.rsq);
    double2 *bv = reinterpret_cast<double2 *>(lds + L.bvec);
    double2 *red = reinterpret_cast<double2 *>(lds + L.red);

    const double2 *A = P.A + J.a_off;
    double2 *G = P.G + J.g_off;
    const double2 *g = FUSED ? P.g + J.g_off : nullptr;

    // ---- per-triple tables
    for (int t = tm.rank; t < smax; t += tm.size) {
        sq[t] = T.sq[t];
        rsq[t] = T.rsq[t];
    }
    if (tm.rank < D) bv[tm.rank] = P.b[J.b_off + tm.rank];
    if (tm.rank == 0) bv[D] = P.c[J.c_off];
    const int nrows = coef_rows(D, top);
    for (int t = tm.rank; t < nrows * smax; t += tm.size) {
        int r = t / smax, k = t - r * smax;
        // row r -> (i, j)
        int i = top, rr = r;
        while (rr >= D - i) {
            rr -= D - i;
            ++i;
        }
        int j = i + rr;
        double2 a = A[i * D + j];
        coef[t] = (k < J.shape[j]) ? cscale_promoted(a, T.sq[k]) : make_double2(0.0, 0.0);
    }
    for (int f = tm.rank; f < S; f += tm.size) {
        uint32_t kp = 0;
        int64_t rem = f;
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (j <= top) continue;
            int64_t dj = rem / J.stride[j];
            rem -= dj * J.stride[j];
            kp |= uint32_t(dj) << J.kbit[j];
        }
        kv[f] = kp;
    }
    tm.sync();

    const double2 c = bv[D];
    Acc<D> acc;
    if (FUSED) acc.zero();

    // ---- levels d = D-1 .. top+1 inside ring slot 0
    if (tm.rank == 0) {
        ring[0] = c;
        G[0] = c;
    }
    tm.sync();
    for (int d = D - 1; d > top; --d) {
        const int Sd = int(J.stride[d]);
        const int sd = J.shape[d];
        const double2 bd = bv[d];
        const double2 *cdd = coef + coef_row(D, top, d, d) * smax;
        for (int m = 1; m < sd; ++m) {
            const double sm = sq[m], rm = rsq[m];
            for (int q = tm.rank; q < Sd; q += tm.size) {
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
            }
            tm.sync();
        }
    }

    // ---- VJP of the sub-lattice below `top` (all of it sits in ring slot 0)
    if (FUSED) {
        for (int f = tm.rank; f < S; f += tm.size) {
            const uint32_t kp = kv[f];
            // lowest non-zero axis > top is the "pivot" role; vjp_point with lo = top, klo = 0
            vjp_point<D>(acc, J, sq, kp, top, 0, ring, ring, ring, f, g[f]);
        }
    }

    // ---- level `top`: slabs m = 1 .. shape[top]-1 in the ring
    {
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
            }
            tm.sync();
            if (FUSED) {
                for (int q = tm.rank; q < S; q += tm.size)
                    vjp_point<D>(acc, J, sq, kv[q], top, m, cur, p1, p2, q, g[gbase + q]);
            }
        }
    }

    if (FUSED) reduce_and_store<D, WAVE>(acc, J, P, tm, red, c);
    tm.sync();
}

// One-index chain in a single lane (D = 1: core.py:108-122 with D = 1).
template <bool FUSED>
__device__ void chain_lane(const Job &J, const Bases &P, const Tables &T) {
    const int s0 = J.shape[0];
    const double2 a = P.A[J.a_off];
    const double2 b = P.b[J.b_off];
    const double2 c = P.c[J.c_off];
    double2 *G = P.G + J.g_off;
    const double2 *g = FUSED ? P.g + J.g_off : nullptr;
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
        if (J.flags & JOB_WRITE_PARTIAL) {
            double2 *out = P.partial + J.p_off;
            out[0] = ab;
            out[1] = bb;
            out[2] = cc;
        } else {
            P.dA[J.a_off] = ab;
            P.db[J.b_off] = bb;
            P.dc[J.c_off] = cdiv_py(cc, c);
        }
    }
}

}  // namespace mm

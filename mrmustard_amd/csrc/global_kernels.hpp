// Kernels that work on G in global memory: the per-slab forward for lattices
// whose slabs do not fit in LDS, the closed-form VJP reduction, and the
// all-pivot ("stable") recurrence swept by index-sum wavefronts.
#pragma once

#include "common.hpp"
#include "exact.hpp"
#include "jobs.hpp"

namespace mm {

struct Geom {
    int32_t D;
    int32_t shape[8];
    int64_t stride[8];
    int64_t N;
};

// Mixed-radix digits of x over axes [lo, D).  Lattices below 2^31 points use
// 32-bit divisions (a 64-bit one is a long software sequence, and a thread of
// a slab launch with one point per thread would spend most of its time here).
template <int D>
__device__ __forceinline__ void digits_of(int64_t x, const Geom &g, int lo, int (&k)[D]) {
    if (g.N <= int64_t(0x7fffffff)) {
        uint32_t r = uint32_t(x);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (j < lo) {
                k[j] = 0;
                continue;
            }
            const uint32_t sj = uint32_t(g.stride[j]);
            const uint32_t dj = r / sj;
            r -= dj * sj;
            k[j] = int(dj);
        }
        return;
    }
#pragma unroll
    for (int j = 0; j < D; ++j) {
        if (j < lo) {
            k[j] = 0;
            continue;
        }
        int64_t dj = x / g.stride[j];
        x -= dj * g.stride[j];
        k[j] = int(dj);
    }
}

// k += inc (both as digits over axes [lo, D)); returns false on overflow past axis lo.
template <int D>
__device__ __forceinline__ bool digits_add(int (&k)[D], const int (&inc)[D], const Geom &g, int lo) {
    int carry = 0;
#pragma unroll
    for (int j = D - 1; j >= 0; --j) {
        if (j < lo) break;
        int v = k[j] + inc[j] + carry;
        carry = v >= g.shape[j];
        k[j] = carry ? v - g.shape[j] : v;
    }
    return carry == 0;
}

// ------------------------------------------------------------ per-slab forward
// Slab k_d = m of the sub-lattice {k_0..k_{d-1} = 0}: every point has pivot
// axis d (core.py:88-104 with i = d) and reads only slabs m-1 and m-2.
struct SlabArgs {
    Geom g;
    const double2 *A, *b;
    double2 *G;
    const double *sq, *rsq;
    int *nf;  // raised when a stored value is not finite
    int d, m;
};

template <int D>
__global__ __launch_bounds__(256) void fwd_slab_kernel(SlabArgs a) {
    const int d = a.d, m = a.m;
    const int64_t Sd = a.g.stride[d];
    const Span sp = xcd_span(xcd_block(blockIdx.x, gridDim.x), gridDim.x, Sd);
    const int64_t q0 = sp.q0, inc = sp.inc, hi = sp.hi;
    if (q0 >= hi) return;
    int k[D], ki[D];
    digits_of<D>(q0, a.g, d + 1, k);
    digits_of<D>(inc < Sd ? inc : 0, a.g, d + 1, ki);
    const double2 bd = a.b[d];
    const double2 cdd = cscale_promoted(a.A[d * D + d], a.sq[m > 0 ? m - 1 : 0]);
    const double sm = a.sq[m], rm = a.rsq[m];
    double2 Adj[D];
#pragma unroll
    for (int j = 0; j < D; ++j) Adj[j] = (j > d) ? a.A[d * D + j] : make_double2(0.0, 0.0);
    double2 *G = a.G;
    const int64_t base = int64_t(m) * Sd;
    bool bad = false;
    for (int64_t q = q0; q < hi; q += inc) {
        const int64_t f = base + q;
        // every load of the point first (one round trip); a term with k_j = 0
        // is the zero coefficient A_dj sqrt(0) times the point's own value,
        // as the reference adds it
        double2 nbv[D];
        double wv[D];
#pragma unroll
        for (int j = 1; j < D; ++j) {
            if (j <= d) continue;
            nbv[j] = G[f - Sd - (k[j] > 0 ? a.g.stride[j] : 0)];
            wv[j] = a.sq[k[j]];
        }
        double2 v = cmul(bd, G[f - Sd]);
        if (m >= 2) v = cadd(v, cmul(cdd, G[f - 2 * Sd]));
#pragma unroll
        for (int j = 1; j < D; ++j) {
            if (j <= d) continue;
            v = cadd(v, cmul(cscale_promoted(Adj[j], wv[j]), nbv[j]));
        }
        const double2 out = cdiv_by(v, sm, rm);
        G[f] = out;
        bad = bad || !finite2(out);
        if (q + inc < hi) digits_add<D>(k, ki, a.g, d + 1);
    }
    if (__builtin_expect(bad, 0) && a.nf) atomicOr(a.nf, 1);
}

// --------------------------------------------- several slabs per launch
// Slabs m0 .. m0 + T - 1 of level d in one launch (T = 1 or 2), for slabs
// small enough that a launch per slab costs more than the slab's work: one
// block per CU, 1024 threads, packed digits and an LDS coefficient table
// (32-bit index arithmetic; fwd_slab_kernel takes any slab size).  A point
// of slab m reads slab m - 1 at most h = stride[d+1] points back
// (k - 1_d - 1_j, j > d) and slab m - 2 at itself, so block b, owning the
// points [a_b, e_b) of every slab, computes slab m0 + s (phase s) on
// [a_b - (T-1-s) h, e_b): its own points and those its later phases read.
// A phase's values stay in LDS for the next phase (two buffers); a point's
// own values of the two slabs below
// stay in the registers of the thread that owns the point in every phase
// (points are numbered from e_b down, so a region that shrinks from the
// front keeps the numbering); phase 0 loads them and the point's neighbours
// in slab m0 - 1 in one round trip.  Only owned points are written to G,
// each by the expression of fwd_slab_kernel, so G is bit for bit the same;
// a term with k_j = 0 is added as the reference adds it, a zero coefficient
// A_dj sqrt(0) times a value read at the point itself (no branch per lane).
constexpr int MULTI_THREADS = 1024;
constexpr int MULTI_KMAX = 2;  // points per thread: regions of at most 2048 points

struct SlabMultiArgs {
    Geom g;
    const double2 *A, *b;
    double2 *G;
    const double *sq, *rsq;
    int *nf;  // raised when a stored value is not finite
    int d, m0, nphase;
    int tile, halo, cap;    // owned points per block, neighbour reach, points per LDS buffer
    int smax;               // row length of the coefficient table (largest shape of the axes > d)
    int kbit[8], kmask[8];  // packed digits of the axes > d
};

template <int D>
__global__ __launch_bounds__(MULTI_THREADS) void fwd_slab_multi_kernel(SlabMultiArgs a) {
    extern __shared__ __align__(16) char lds[];
    double2 *buf0 = reinterpret_cast<double2 *>(lds);
    double2 *buf1 = buf0 + a.cap;
    double2 *ctab = buf1 + a.cap;  // A_dj sqrt(k_j) for j > d, row j - d - 1
    const double2 zero = make_double2(0.0, 0.0);
    const int d = a.d, T = a.nphase, h = a.halo;
    const int Sd = int(a.g.stride[d]);
    for (int t = threadIdx.x; t < (D - 1 - d) * a.smax; t += blockDim.x) {
        const int j = d + 1 + t / a.smax, k = t % a.smax;
        ctab[t] = k < a.g.shape[j] ? cscale_promoted(a.A[d * D + j], a.sq[k]) : zero;
    }
    const int ab = int(xcd_block(blockIdx.x, gridDim.x)) * a.tile;
    const int eb = min(ab + a.tile, Sd);
    const int lo0 = max(0, ab - (T - 1) * h);
    int st[D];
#pragma unroll
    for (int j = 0; j < D; ++j) st[j] = int(a.g.stride[j]);
    uint32_t kp[MULTI_KMAX];
    double2 p1v[MULTI_KMAX], p2v[MULTI_KMAX], nb[MULTI_KMAX][D];
    const int64_t base0 = int64_t(a.m0) * Sd;
    const double2 *G0 = a.G + base0 - Sd;  // slab m0 - 1
#pragma unroll
    for (int r = 0; r < MULTI_KMAX; ++r) {
        const int q = eb - 1 - int(threadIdx.x) - r * MULTI_THREADS;
        kp[r] = 0;
        p1v[r] = zero;
        p2v[r] = zero;
#pragma unroll
        for (int j = 0; j < D; ++j) nb[r][j] = zero;
        if (q >= lo0) {
            int rem = q;
#pragma unroll
            for (int j = 0; j < D; ++j) {
                if (j <= d) continue;
                const int kj = rem / st[j];
                rem -= kj * st[j];
                kp[r] |= uint32_t(kj) << a.kbit[j];
                nb[r][j] = G0[kj > 0 ? q - st[j] : q];  // phase 0's neighbours, loaded with the point's own values
            }
            p1v[r] = G0[q];
            if (a.m0 >= 2) p2v[r] = G0[q - Sd];
        }
    }
    __syncthreads();  // the coefficient table
    const double2 bd = a.b[d];
    bool bad = false;
    for (int s = 0; s < T; ++s) {
        const int m = a.m0 + s;
        const int lo = max(0, ab - (T - 1 - s) * h);
        const int plo = max(0, ab - (T - s) * h);  // origin of the previous phase's buffer
        const double2 *prev = (s & 1) ? buf0 : buf1;
        double2 *cur = (s & 1) ? buf1 : buf0;
        const double2 cdd = cscale_promoted(a.A[d * D + d], a.sq[m > 0 ? m - 1 : 0]);
        const double sm = a.sq[m], rm = a.rsq[m];
#pragma unroll
        for (int r = 0; r < MULTI_KMAX; ++r) {
            const int q = eb - 1 - int(threadIdx.x) - r * MULTI_THREADS;
            if (q < lo) continue;
            double2 v = cmul(bd, p1v[r]);
            if (m >= 2) v = cadd(v, cmul(cdd, p2v[r]));
#pragma unroll
            for (int j = 1; j < D; ++j) {
                if (j <= d) continue;
                const int kj = int((kp[r] >> a.kbit[j]) & uint32_t(a.kmask[j]));
                // k_j = 0: the point itself times the zero coefficient A_dj sqrt(0)
                const double2 nbv = s == 0 ? nb[r][j] : prev[(kj > 0 ? q - st[j] : q) - plo];
                v = cadd(v, cmul(ctab[(j - d - 1) * a.smax + kj], nbv));
            }
            const double2 out = cdiv_by(v, sm, rm);
            if (s + 1 < T) cur[q - lo] = out;
            if (q >= ab) {
                a.G[int64_t(m) * Sd + q] = out;
                bad = bad || !finite2(out);
            }
            p2v[r] = p1v[r];
            p1v[r] = out;
        }
        __syncthreads();
    }
    if (__builtin_expect(bad, 0) && a.nf) atomicOr(a.nf, 1);
}

// Level-0 slab m with the three slabs in separate buffers (a rolling window
// that never holds the whole lattice): cur <- f(p1, p2).
struct Slab0Args {
    Geom g;
    const double2 *A, *b;
    const double2 *p1, *p2;  // slabs m-1 and m-2 (p2 unused for m = 1)
    double2 *cur;
    const double *sq, *rsq;
    int m;
};

// Block `blk` of `nblk` (256 threads each) over the slab's points.
template <int D>
__device__ __forceinline__ void fwd_slab0_body(const Slab0Args &a, int64_t blk, int64_t nblk) {
    const int m = a.m;
    const int64_t S = a.g.stride[0];
    const Span sp = xcd_span(blk, nblk, S);
    const int64_t q0 = sp.q0, inc = sp.inc, hi = sp.hi;
    if (q0 >= hi) return;
    int k[D], ki[D];
    digits_of<D>(q0, a.g, 1, k);
    digits_of<D>(inc < S ? inc : 0, a.g, 1, ki);
    const double2 b0 = a.b[0];
    const double2 c00 = cscale_promoted(a.A[0], a.sq[m - 1]);
    const double sm = a.sq[m], rm = a.rsq[m];
    double2 A0j[D];
#pragma unroll
    for (int j = 0; j < D; ++j) A0j[j] = (j > 0) ? a.A[j] : make_double2(0.0, 0.0);
    for (int64_t q = q0; q < hi; q += inc) {
        // every load of the point first, zero-digit terms as in fwd_slab_kernel
        double2 nbv[D];
        double wv[D];
#pragma unroll
        for (int j = 1; j < D; ++j) {
            nbv[j] = a.p1[q - (k[j] > 0 ? a.g.stride[j] : 0)];
            wv[j] = a.sq[k[j]];
        }
        double2 v = cmul(b0, a.p1[q]);
        if (m >= 2) v = cadd(v, cmul(c00, a.p2[q]));
#pragma unroll
        for (int j = 1; j < D; ++j) v = cadd(v, cmul(cscale_promoted(A0j[j], wv[j]), nbv[j]));
        a.cur[q] = cdiv_by(v, sm, rm);
        if (q + inc < hi) digits_add<D>(k, ki, a.g, 1);
    }
}

template <int D>
__global__ __launch_bounds__(256) void fwd_slab0_kernel(Slab0Args a) {
    fwd_slab0_body<D>(a, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

__global__ void zero_c128_kernel(double2 *p, int64_t n) {
    for (int64_t i = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; i < n; i += int64_t(gridDim.x) * blockDim.x)
        p[i] = make_double2(0.0, 0.0);
}

// ------------------------------------------------------------------- the VJP
// gradients.py:64-80 as a parallel reduction: each thread accumulates the
// terms of its points (fused multiply-adds), a block reduces through shuffles
// and LDS, and writes one partial per block; vjp_finalize sums the partials
// of a triple in block order (deterministic) and symmetrises.
struct VjpArgs {
    Geom g;
    const double2 *G, *gg;
    double2 *partial;
    const double *sq;
    int nchunks;
};

// Accumulators of one group of pivot rows [R0, R1): dA(i, j) for R0 <= i < R1
// and all j >= i, db_i for those rows, and dc with the first group.  Large D
// splits the rows over several launches so each thread holds few sums.
template <int D, int R0, int R1>
struct RowGroup {
    static constexpr int NT = D * (D + 1) / 2;
    static constexpr int NA = NT + D + 1;
    static constexpr int NROWA = (R1 - R0) * D - (R1 * (R1 - 1) - R0 * (R0 - 1)) / 2;  // sum_{i=R0}^{R1-1} (D - i)
    static constexpr int NL = NROWA + (R1 - R0) + (R0 == 0 ? 1 : 0);
    static __host__ __device__ constexpr int la(int i, int j) { return (i - R0) * D - (i * (i - 1) - R0 * (R0 - 1)) / 2 + (j - i); }
    static __host__ __device__ constexpr int lb(int i) { return NROWA + (i - R0); }
    static __host__ __device__ constexpr int lc() { return NROWA + (R1 - R0); }
    // column of local sum l in the full layout (upper triangle of dA, db, dc)
    static __device__ int column(int l) {
        if (R0 == 0 && l == lc()) return NA - 1;
        if (l >= NROWA) return NT + R0 + (l - NROWA);
        int i = R0, off = 0;
        while (l - off >= D - i) {
            off += D - i;
            ++i;
        }
        return i * D - i * (i - 1) / 2 + (l - off);
    }
};

// Reduce NL sums over the block (DPP within waves, waves in order) and
// either store or add them into the block's partial row, at the group's columns.
template <int D, int R0, int R1, bool ADD, int NL>
__device__ __forceinline__ void block_store_group(double2 (&acc)[NL], double2 *row) {
    using RG = RowGroup<D, R0, R1>;
    static_assert(NL == RG::NL, "accumulator count");
    __shared__ double2 red[(256 / MM_WAVE) * NL];
    wave_sum_dpp<NL>(acc);
    const int lane = threadIdx.x & (MM_WAVE - 1), w = threadIdx.x / MM_WAVE;
    if (lane == MM_WAVE - 1) {
#pragma unroll
        for (int e = 0; e < NL; ++e) red[w * NL + e] = acc[e];
    }
    __syncthreads();
    if (threadIdx.x < NL) {
        double2 t = red[threadIdx.x];
        for (int ww = 1; ww < int(blockDim.x / MM_WAVE); ++ww) t = cadd(t, red[ww * NL + threadIdx.x]);
        const int col = RG::column(threadIdx.x);
        row[col] = ADD ? cadd(row[col], t) : t;
    }
}

// gradients.py:64-80 as a parallel reduction over a whole lattice held in
// memory.  Branch-free: a term whose digit is zero reads the point itself
// with weight sqrt(0) = 0, so every gather of a point issues together.
template <int D, int R0, int R1>
__global__ __launch_bounds__(256) void vjp_kernel(VjpArgs a) {
    using RG = RowGroup<D, R0, R1>;
    constexpr int NL = RG::NL;
    const int64_t t = blockIdx.y;
    const double2 *G = a.G + t * a.g.N;
    const double2 *gg = a.gg + t * a.g.N;
    double2 acc[NL];
#pragma unroll
    for (int e = 0; e < NL; ++e) acc[e] = make_double2(0.0, 0.0);
    const int64_t f0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    const int64_t inc = int64_t(gridDim.x) * blockDim.x;
    if (f0 < a.g.N) {
        int64_t st[D];
#pragma unroll
        for (int i = 0; i < D; ++i) st[i] = a.g.stride[i];
        int k[D], ki[D];
        digits_of<D>(f0, a.g, 0, k);
        digits_of<D>(inc < a.g.N ? inc : 0, a.g, 0, ki);
        for (int64_t f = f0; f < a.g.N; f += inc) {
            const double2 gk = gg[f];
            if (R0 == 0) acc[RG::lc()] = cfma(G[f], gk, acc[RG::lc()]);
#pragma unroll
            for (int i = R0; i < R1; ++i) {
                const int kk = k[i];
                const double si = a.sq[kk];
                const double2 ui = make_double2(si * gk.x, si * gk.y);
                const int64_t p = f - (kk > 0 ? st[i] : 0);
                acc[RG::lb(i)] = cfma(G[p], ui, acc[RG::lb(i)]);
                const double h = 0.5 * a.sq[kk > 0 ? kk - 1 : 0];
                const double2 t2 = G[p - (kk > 1 ? st[i] : 0)];
                acc[RG::la(i, i)] = cfma(make_double2(h * t2.x, h * t2.y), ui, acc[RG::la(i, i)]);
#pragma unroll
                for (int j = i + 1; j < D; ++j) {
                    const bool ok = kk > 0 && k[j] > 0;
                    const double2 t3 = G[ok ? p - st[j] : f];
                    const double w = a.sq[k[j]];
                    acc[RG::la(i, j)] = cfma(make_double2(w * t3.x, w * t3.y), ui, acc[RG::la(i, j)]);
                }
            }
            if (f + inc < a.g.N) digits_add<D>(k, ki, a.g, 0);
        }
    }
    block_store_group<D, R0, R1, false>(acc, a.partial + (t * a.nchunks + blockIdx.x) * RG::NA);
}

// VJP terms of the points of level-0 slab m (k_0 = m), the three slabs given
// as separate buffers (the recompute window); each block adds its sums into
// its own partial row (launches are stream-ordered, so the order of the
// additions is fixed).  Row 0 reads slabs m-1 and m-2; with m < 1 / m < 2
// their weights are sqrt(0) = 0 and the pointers are only dereferenced at
// valid slab-0 addresses.
struct SlabVjpArgs {
    Geom g;
    const double2 *cur, *p1, *p2, *gg;
    double2 *partial;
    const double *sq;
    int m;
};

// Block `blk` of `nblk` (256 threads each) adds its sums into partial row blk.
template <int D, int R0, int R1>
__device__ __forceinline__ void vjp_slab_body(const SlabVjpArgs &a, int64_t blk, int64_t nblk) {
    using RG = RowGroup<D, R0, R1>;
    constexpr int NL = RG::NL;
    const int m = a.m;
    const int64_t S = a.g.stride[0];
    double2 acc[NL];
#pragma unroll
    for (int e = 0; e < NL; ++e) acc[e] = make_double2(0.0, 0.0);
    const Span sp = xcd_span(blk, nblk, S);
    const int64_t q0 = sp.q0, inc = sp.inc, hi = sp.hi;
    if (q0 < hi) {
        int64_t st[D];
#pragma unroll
        for (int i = 0; i < D; ++i) st[i] = a.g.stride[i];
        int k[D], ki[D];
        digits_of<D>(q0, a.g, 1, k);
        digits_of<D>(inc < S ? inc : 0, a.g, 1, ki);
        k[0] = m;
        ki[0] = 0;
        const double2 *p1 = m >= 1 ? a.p1 : a.cur;
        const double2 *p2 = m >= 2 ? a.p2 : a.cur;
        for (int64_t q = q0; q < hi; q += inc) {
            const double2 gk = a.gg[q];
            if (R0 == 0) {
                acc[RG::lc()] = cfma(a.cur[q], gk, acc[RG::lc()]);
                const double s0 = a.sq[m];
                const double2 u0 = make_double2(s0 * gk.x, s0 * gk.y);
                acc[RG::lb(0)] = cfma(p1[q], u0, acc[RG::lb(0)]);
                const double h = 0.5 * a.sq[m > 0 ? m - 1 : 0];
                const double2 t2 = p2[q];
                acc[RG::la(0, 0)] = cfma(make_double2(h * t2.x, h * t2.y), u0, acc[RG::la(0, 0)]);
#pragma unroll
                for (int j = 1; j < D; ++j) {
                    const double2 t3 = p1[k[j] > 0 ? q - st[j] : q];
                    const double w = a.sq[k[j]];
                    acc[RG::la(0, j)] = cfma(make_double2(w * t3.x, w * t3.y), u0, acc[RG::la(0, j)]);
                }
            }
#pragma unroll
            for (int i = (R0 > 1 ? R0 : 1); i < R1; ++i) {
                const int kk = k[i];
                const double si = a.sq[kk];
                const double2 ui = make_double2(si * gk.x, si * gk.y);
                const int64_t p = q - (kk > 0 ? st[i] : 0);
                acc[RG::lb(i)] = cfma(a.cur[p], ui, acc[RG::lb(i)]);
                const double h = 0.5 * a.sq[kk > 0 ? kk - 1 : 0];
                const double2 t2 = a.cur[p - (kk > 1 ? st[i] : 0)];
                acc[RG::la(i, i)] = cfma(make_double2(h * t2.x, h * t2.y), ui, acc[RG::la(i, i)]);
#pragma unroll
                for (int j = i + 1; j < D; ++j) {
                    const bool ok = kk > 0 && k[j] > 0;
                    const double2 t3 = a.cur[ok ? p - st[j] : q];
                    const double w = a.sq[k[j]];
                    acc[RG::la(i, j)] = cfma(make_double2(w * t3.x, w * t3.y), ui, acc[RG::la(i, j)]);
                }
            }
            if (q + inc < hi) digits_add<D>(k, ki, a.g, 1);
        }
    }
    block_store_group<D, R0, R1, true>(acc, a.partial + blk * RG::NA);
}

template <int D, int R0, int R1>
__global__ __launch_bounds__(256) void vjp_slab_kernel(SlabVjpArgs a) {
    vjp_slab_body<D, R0, R1>(a, xcd_block(blockIdx.x, gridDim.x), gridDim.x);
}

// Row groups per D: one for D <= 4, two for D = 5, 6, three for D = 7, 8.
template <int D>
struct RowSplit {
    static constexpr int NG = D <= 4 ? 1 : (D <= 6 ? 2 : 3);
    static constexpr int B1 = D <= 4 ? D : 2;
    static constexpr int B2 = D <= 6 ? D : 4;
};

// One level-0 slab per launch, all of its work in that launch: the VJP terms
// of slab m in every pivot-row group and (FWD) the forward of slab m + 1 from
// slabs m and m - 1.  Logical block l (xcd_block) has role l % NR (a row
// group, or the forward) and covers chunk l / NR of the slab, so the NR
// blocks of one chunk run at the same time on one XCD and read the same slab
// data: each slab is fetched about once per launch instead of once per row
// group.  The forward writes
// slab m + 1, which no VJP term of this launch reads.
template <int D, bool FWD>
__global__ __launch_bounds__(256) void slab_fused_kernel(SlabVjpArgs v, Slab0Args f) {
    using RS = RowSplit<D>;
    constexpr int NR = RS::NG + (FWD ? 1 : 0);
    const int64_t l = xcd_block(blockIdx.x, gridDim.x);  // the NR roles of a chunk on one XCD
    const int64_t role = l % NR, blk = l / NR, nblk = gridDim.x / NR;
    if (role == 0) {
        vjp_slab_body<D, 0, RS::B1>(v, blk, nblk);
    } else if (RS::NG >= 2 && role == 1) {
        if constexpr (RS::NG >= 2) vjp_slab_body<D, RS::B1, RS::B2>(v, blk, nblk);
    } else if (RS::NG >= 3 && role == 2) {
        if constexpr (RS::NG >= 3) vjp_slab_body<D, RS::B2, D>(v, blk, nblk);
    } else if (FWD) {
        fwd_slab0_body<D>(f, blk, nblk);
    }
}

// One block per triple.  npart partial rows of NA sums each (the VJP kernel's
// chunks, optionally followed by extra rows from other kernels).  Wave w sums
// columns w, w + nwaves, ...: lane l adds rows l, l + 64, ... in order, then
// the lanes combine in a fixed butterfly, so the result does not depend on
// timing.
template <int D>
__global__ __launch_bounds__(1024) void vjp_finalize(const double2 *partial, int npart, const double2 *c, double2 *dA, double2 *db,
                                                     double2 *dc) {
    constexpr int NA = D * (D + 1) / 2 + D + 1;
    constexpr int NT = D * (D + 1) / 2;
    __shared__ double2 s[NA];
    const int64_t t = blockIdx.x;
    const double2 *p = partial + t * npart * NA;
    const int lane = threadIdx.x & (MM_WAVE - 1);
    const int nw = blockDim.x / MM_WAVE;
    for (int e = threadIdx.x / MM_WAVE; e < NA; e += nw) {
        double2 acc = make_double2(0.0, 0.0);
#pragma unroll 8
        for (int r = lane; r < npart; r += MM_WAVE) acc = cadd(acc, p[int64_t(r) * NA + e]);
        acc = wave_sum(acc);
        if (lane == 0) s[e] = acc;
    }
    __syncthreads();
    if (threadIdx.x < D * D) {
        const int i = threadIdx.x / D, j = threadIdx.x % D;
        const int lo = i < j ? i : j, hi = i < j ? j : i;
        const double2 m = s[lo * D - lo * (lo - 1) / 2 + (hi - lo)];
        dA[t * D * D + threadIdx.x] = (i == j) ? m : make_double2(m.x * 0.5, m.y * 0.5);
    }
    if (threadIdx.x < D) db[t * D + threadIdx.x] = s[NT + threadIdx.x];
    if (threadIdx.x == 0) dc[t] = cdiv_py(s[NA - 1], c[t]);
}

// ------------------------------------------------------------- stable sweep
// core.py:183-211.  All pivots of a point lie on the previous index-sum
// wavefront, so the lattice is swept by |k| = n.  A "row" fixes k_0..k_{D-2};
// on wavefront n row r holds exactly one point, k_{D-1} = n - |row r|, when
// that is inside the box.  The row table stores each row's digits (uint16).
struct StableArgs {
    Geom g;
    double rstride[8];     // 1 / stride[j], for row digits of lattices below 2^31 points
    const double2 *A, *b, *c;
    double2 *G;
    const uint16_t *rows;  // nrows x 8 digits
    int64_t nrows;         // rows of the last axis: N / shape[D - 1]
    const double *sq, *rsq;
    int n;                 // wavefront (per-wavefront launch)
    int64_t tstride_a, tstride_b, tstride_g;  // batch strides (complex elements)
    int *nf;               // per-wavefront launches: raised when a stored value is not finite
};

// Digits of every row, read by the per-wavefront and block kernels (a load
// from L2 is quicker than the row's divisions on their critical path).
__global__ void row_table_kernel(Geom g, uint16_t *rows, int64_t nrows) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= nrows) return;
    const int D = g.D;
    const int64_t slast = g.shape[D - 1];
    int64_t x = r * slast;
    uint16_t *out = rows + r * 8;
    for (int j = 0; j < 8; ++j) {
        if (j < D - 1) {
            int64_t dj = x / g.stride[j];
            x -= dj * g.stride[j];
            out[j] = uint16_t(dj);
        } else {
            out[j] = 0;
        }
    }
}

// Digits k_0..k_{D-2} of row r (the points r * shape[D-1] + k_{D-1}) and their
// sum, computed (the corner sweeps, once per thread).  Below 2^31 points each quotient is x * (1 / stride) in double
// precision, off by at most one (the product's error is below 2^-21), then
// corrected from the remainder.
template <int D>
__device__ __forceinline__ int row_digits(const Geom &g, const double (&rinv)[8], int64_t r, int (&k)[D]) {
    int p = 0;
    const int64_t f = r * g.shape[D - 1];
    if (g.N <= int64_t(0x7fffffff)) {
        int x = int(f);
#pragma unroll
        for (int j = 0; j < D - 1; ++j) {
            const int sj = int(g.stride[j]);
            int dj = int(double(x) * rinv[j]);
            const int64_t rem = int64_t(x) - int64_t(dj) * sj;
            dj += rem >= sj ? 1 : (rem < 0 ? -1 : 0);
            x -= dj * sj;
            k[j] = dj;
            p += dj;
        }
    } else {
        int64_t x = f;
#pragma unroll
        for (int j = 0; j < D - 1; ++j) {
            const int64_t dj = x / g.stride[j];
            x -= dj * g.stride[j];
            k[j] = int(dj);
            p += k[j];
        }
    }
    k[D - 1] = 0;
    return p;
}

// Returns true when the stored value is not finite.  BF (branch-free): the
// compiler can issue the loads of every pivot together; a term of the
// reference whose weight is sqrt(0) (k_j = 0, or the A_ii term at k_i = 1)
// is added as the reference adds it, the zero coefficient times a value
// already computed (the pivot's own), and a pivot axis with k_i = 0 is
// evaluated on the point's own address and left out of the sum and of the
// count.  Without BF (fewer registers: the 1024-thread kernels at larger D)
// zero-weight terms and pivots are skipped, which gives the same values for
// finite G.
template <int D, bool BF>
__device__ __forceinline__ bool stable_point(const StableArgs &a, const double2 *A, const double2 *b, double2 *G, const int (&k)[D],
                                             int64_t f) {
    int num = 0;
    double2 vals = make_double2(0.0, 0.0);
    if constexpr (!BF) {
#pragma unroll
        for (int i = 0; i < D; ++i) {
            if (k[i] == 0) continue;
            ++num;
            const int64_t pivot = f - a.g.stride[i];
            double2 val = cmul(b[i], G[pivot]);
#pragma unroll
            for (int j = 0; j < D; ++j) {
                if (j >= i) break;
                if (k[j] > 0) val = cadd(val, cmul(cscale_promoted(A[i * D + j], a.sq[k[j]]), G[pivot - a.g.stride[j]]));
            }
            if (k[i] >= 2) val = cadd(val, cmul(cscale_promoted(A[i * D + i], a.sq[k[i] - 1]), G[pivot - a.g.stride[i]]));
#pragma unroll
            for (int j = i + 1; j < D; ++j)
                if (k[j] > 0) val = cadd(val, cmul(cscale_promoted(A[i * D + j], a.sq[k[j]]), G[pivot - a.g.stride[j]]));
            vals = cadd(vals, cdiv_by(val, a.sq[k[i]], a.rsq[k[i]]));
        }
        const double2 out = cdiv_py(vals, make_double2(double(num), 0.0));
        G[f] = out;
        return !finite2(out);
    }
#pragma unroll
    for (int i = 0; i < D; ++i) {
        const int64_t pivot = f - (k[i] > 0 ? a.g.stride[i] : 0);
        double2 val = cmul(b[i], G[pivot]);
#pragma unroll
        for (int j = 0; j < D; ++j) {
            if (j >= i) break;
            val = cadd(val, cmul(cscale_promoted(A[i * D + j], a.sq[k[j]]), G[pivot - (k[j] > 0 ? a.g.stride[j] : 0)]));
        }
        val = cadd(val, cmul(cscale_promoted(A[i * D + i], a.sq[k[i] > 0 ? k[i] - 1 : 0]), G[pivot - (k[i] >= 2 ? a.g.stride[i] : 0)]));
#pragma unroll
        for (int j = i + 1; j < D; ++j)
            val = cadd(val, cmul(cscale_promoted(A[i * D + j], a.sq[k[j]]), G[pivot - (k[j] > 0 ? a.g.stride[j] : 0)]));
        const double2 term = cdiv_by(val, a.sq[k[i]], a.rsq[k[i]]);
        const bool on = k[i] > 0;
        vals = make_double2(on ? vals.x + term.x : vals.x, on ? vals.y + term.y : vals.y);
        num += on ? 1 : 0;
    }
    const double2 out = cdiv_py(vals, make_double2(double(num), 0.0));
    G[f] = out;
    return !finite2(out);
}

template <int D, bool BF>
__device__ __forceinline__ bool stable_row(const StableArgs &a, const double2 *A, const double2 *b, double2 *G, int64_t r, int n) {
    const uint16_t *rd = a.rows + r * 8;
    int k[D];
    int p = 0;
#pragma unroll
    for (int j = 0; j < D - 1; ++j) {
        k[j] = rd[j];
        p += k[j];
    }
    const int kl = n - p;
    if (kl < 0 || kl >= a.g.shape[D - 1]) return false;
    k[D - 1] = kl;
    return stable_point<D, BF>(a, A, b, G, k, r * a.g.shape[D - 1] + kl);
}

// One wavefront of one triple per launch (large lattices).
template <int D>
__global__ __launch_bounds__(256) void stable_wavefront_kernel(StableArgs a) {
    const int64_t r0 = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    bool bad = false;
    for (int64_t r = r0; r < a.nrows; r += int64_t(gridDim.x) * blockDim.x) bad |= stable_row<D, true>(a, a.A, a.b, a.G, r, a.n);
    if (__builtin_expect(bad, 0) && a.nf) atomicOr(a.nf, 1);
}

// The rows of a per-wavefront sweep whose digit sum p = k_0 + ... + k_{D-2} is
// at most plo (list lo) or at least phi (list hi), appended in any order
// (counts[0..1] start at zero; the host knows both totals, at most cap each).
template <int D>
__global__ __launch_bounds__(256) void stable_corner_rows_kernel(StableArgs a, int plo, int phi, int32_t *lo, int32_t *hi, int *counts,
                                                                 int cap) {
    const int64_t r = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (r >= a.nrows) return;
    int k[D];
    const int p = row_digits<D>(a.g, a.rstride, r, k);
    if (p <= plo) {
        const int e = atomicAdd(counts, 1);
        if (e < cap) lo[e] = int32_t(r);
    }
    if (p >= phi) {
        const int e = atomicAdd(counts + 1, 1);
        if (e < cap) hi[e] = int32_t(r);
    }
}

// Wavefronts n0..n1 of one triple in one workgroup, one listed row per
// thread: the first and last wavefronts of a large lattice hold few points,
// and at D = 2 a workgroup step costs about a third of a launch (used for
// D = 2 only).  Every point of these wavefronts lies on a listed row.
template <int D>
__global__ __launch_bounds__(1024) void stable_corner_kernel(StableArgs a, const int32_t *rows, int count, int n0, int n1) {
    const int64_t r = int(threadIdx.x) < count ? rows[threadIdx.x] : -1;
    const bool has = r >= 0 && r < a.nrows;
    int k[D];
    const int p = row_digits<D>(a.g, a.rstride, has ? r : 0, k);
    const int slast = a.g.shape[D - 1];
    const int64_t f0 = (has ? r : 0) * slast;
    bool bad = false;
    for (int n = n0; n <= n1; ++n) {
        const int kl = n - p;
        if (has && kl >= 0 && kl < slast) {
            k[D - 1] = kl;
            bad |= stable_point<D, (D <= 4)>(a, a.A, a.b, a.G, k, f0 + kl);
        }
        __syncthreads();
    }
    if (__builtin_expect(bad, 0) && a.nf) atomicOr(a.nf, 1);
}

// Whole sweep of one triple per workgroup (small lattices, batches).
template <int D>
__global__ __launch_bounds__(1024) void stable_block_kernel(StableArgs a) {
    const int64_t t = blockIdx.x;
    const double2 *A = a.A + t * a.tstride_a;
    const double2 *b = a.b + t * a.tstride_b;
    double2 *G = a.G + t * a.tstride_g;
    if (threadIdx.x == 0) G[0] = a.c[t];
    __syncthreads();
    int nmax = 0;
    for (int j = 0; j < D; ++j) nmax += a.g.shape[j] - 1;
    bool bad = !finite2(a.c[t]);
    for (int n = 1; n <= nmax; ++n) {
        for (int64_t r = threadIdx.x; r < a.nrows; r += blockDim.x) bad |= stable_row<D, (D <= 6)>(a, A, b, G, r, n);
        __syncthreads();
    }
    // non-finite values: the reference's sequential sweep with every term
    if (__syncthreads_or(bad)) {
        __shared__ double2 red[16];
        Lat lat;
        lat.D = D;
        for (int j = 0; j < 8; ++j) {
            lat.shape[j] = a.g.shape[j];
            lat.stride[j] = a.g.stride[j];
        }
        lat.N = a.g.N;
        exact_stable_forward(BlockOps(red), lat, A, b, a.c[t], G, a.sq);
    }
}

// G[0] = c of a per-wavefront sweep; the first one of a call also clears the
// call's non-finite flag.
__global__ void set_first_kernel(double2 *G, const double2 *c, int *nf) {
    G[0] = c[0];
    if (nf) *nf = 0;
}

// ------------------------------------------------ non-finite values (exact.hpp)
// Last launch of a multi-launch vanilla path: one workgroup per triple
// reruns the triple in the reference's arithmetic when the path met a
// non-finite value (the flag of a forward, or a non-finite dL/dc of a VJP,
// whose sum of G * dL/dG every non-finite G or dL/dG value poisons).
struct FixArgs {
    Lat g;
    const double2 *A, *b, *c, *gg;
    double2 *G, *dA, *db, *dc;
    const double *sq;
    int64_t ua, ub, uc, ug;  // per-triple strides (complex elements)
    const int *nf;           // forward: the path's flag, or NULL
    int fwd, vjp, stable;
};

__global__ __launch_bounds__(1024) void exact_fixup_kernel(FixArgs a) {
    __shared__ double2 red[16];
    __shared__ int s_bad;
    const int64_t t = blockIdx.x;
    if (threadIdx.x == 0) {
        bool bad = a.nf && __hip_atomic_load(a.nf, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT) != 0;
        if (a.vjp) bad = bad || !finite2(a.dc[t * a.uc]);
        s_bad = bad;
    }
    __syncthreads();
    if (!s_bad) return;
    const BlockOps ops(red);
    const double2 c = a.c[t * a.uc];
    if (a.fwd && a.stable) exact_stable_forward(ops, a.g, a.A + t * a.ua, a.b + t * a.ub, c, a.G + t * a.ug, a.sq);
    else if (a.fwd) exact_forward(ops, a.g, 0, a.A + t * a.ua, a.b + t * a.ub, c, a.G + t * a.ug, a.sq);
    if (a.vjp) exact_vjp(ops, a.g, a.G + t * a.ug, a.gg + t * a.ug, c, a.dA + t * a.ua, a.db + t * a.ub, a.dc + t * a.uc, a.sq);
}

__global__ void zero_ints_kernel(int *p, int n) {
    if (int(threadIdx.x) < n) p[threadIdx.x] = 0;
}

}  // namespace mm

// Exact recomputation of a lattice and of its VJP for inputs or results
// that are not finite.
//
// The fast kernels skip the reference's zero-weight terms: where a digit
// k_j is zero the reference still adds A_ij * SQRT[0] * G[p - s_i - s_j]
// (core.py:101-103, 117-118), reading a position that numpy wraps around,
// and the VJP forms SQRT[0] * G[...] products and zero matrix entries times
// the cotangent (gradients.py:66-78).  With finite values these terms are
// zeros and change nothing; once G holds an inf or a NaN they turn into NaN
// (0 * inf), and so does the other part of a complex number with an infinite
// part after numba's promoted division.  The fast paths detect non-finite
// values (stored amplitudes, or the sum of G * dL/dG, which every bad value
// of G or dL/dG makes non-finite) and then rerun the triple here, term for
// term in the reference's arithmetic:
//   forward: the recurrence of core.py:85-122 with every term, reads of
//     negative flat indices giving zero (numpy wraps them to positions at or
//     past the point, which the reference's zero-initialised G has not
//     written yet); points are visited level by level, slab by slab, which
//     keeps every read behind the point (all reads lie below the start of
//     the point's slab);
//   stable forward: core.py:164-213 with every term, in flat order by one
//     thread (exact_stable_forward);
//   VJP: gradients.py:60-82 with every term, negative indices wrapped
//     around the whole lattice, sums over the team in any order (a sum of
//     values with an inf or NaN is the same in every order; finite parts
//     agree with the reference to rounding).
// Team operations: a lane alone, one wave, or a workgroup.
#pragma once

#include "common.hpp"

namespace mm {

struct Lat {
    int D;
    int32_t shape[8];
    int64_t stride[8];
    int64_t N;
};

__device__ __forceinline__ bool finite2(double2 v) { return __builtin_isfinite(v.x) && __builtin_isfinite(v.y); }

// One lane is the team (a one-index chain per lane).
struct LaneOps {
    int rank = 0, size = 1;
    __device__ void sync() const {}
    __device__ double2 sum(double2 v) const { return v; }
};

// One wave is the team.
struct WaveOps {
    int rank, size;
    __device__ WaveOps() : rank(int(threadIdx.x) & (MM_WAVE - 1)), size(MM_WAVE) {}
    __device__ void sync() const {
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
        __builtin_amdgcn_wave_barrier();
        __builtin_amdgcn_fence(__ATOMIC_SEQ_CST, "workgroup");
    }
    __device__ double2 sum(double2 v) const {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            v.x += shfl_xor_d(v.x, m);
            v.y += shfl_xor_d(v.y, m);
        }
        return v;
    }
};

// The workgroup is the team; `red` is LDS for one value per wave.
struct BlockOps {
    int rank, size;
    double2 *red;
    __device__ explicit BlockOps(double2 *r) : rank(int(threadIdx.x)), size(int(blockDim.x)), red(r) {}
    __device__ void sync() const { __syncthreads(); }
    // total in rank 0
    __device__ double2 sum(double2 v) const {
#pragma unroll
        for (int m = 32; m >= 1; m >>= 1) {
            v.x += shfl_xor_d(v.x, m);
            v.y += shfl_xor_d(v.y, m);
        }
        __syncthreads();
        if ((threadIdx.x & (MM_WAVE - 1)) == 0) red[threadIdx.x / MM_WAVE] = v;
        __syncthreads();
        if (threadIdx.x == 0)
            for (int w = 1; w < int(blockDim.x) / MM_WAVE; ++w) v = cadd(v, red[w]);
        return v;
    }
};

__device__ __forceinline__ int lat_digit(const Lat &g, int64_t p, int j) { return int((p / g.stride[j]) % g.shape[j]); }

// G of the sub-lattice {k_0 .. k_{top-1} = 0} (the whole lattice for top = 0);
// row i of A starts at A + i * lda (lda = 0: D).
template <class Ops>
__device__ void exact_forward(const Ops &t, const Lat &g, int top, const double2 *A, const double2 *b, double2 c, double2 *G,
                              const double *sq, int lda = 0) {
    const int D = g.D;
    const int ld = lda > 0 ? lda : D;
    const double2 zero = make_double2(0.0, 0.0);
    if (t.rank == 0) G[0] = c;
    t.sync();
    for (int L = D - 1; L >= top; --L) {
        const int64_t sL = g.stride[L];
        const double2 bL = b[L];
        for (int k = 1; k < g.shape[L]; ++k) {
            const double2 wLL = cscale_promoted(A[L * ld + L], sq[k - 1]);
            const double sk = sq[k];
            for (int64_t r = t.rank; r < sL; r += t.size) {
                const int64_t p = int64_t(k) * sL + r;
                double2 v = cmul(bL, G[p - sL]);
                const int64_t q = p - 2 * sL;
                const double2 gq = G[q >= 0 ? q : 0];
                v = cadd(v, cmul(wLL, q >= 0 ? gq : zero));
                int64_t rem = r;
                for (int j = L + 1; j < D; ++j) {
                    const int64_t kj = rem / g.stride[j];
                    rem -= kj * g.stride[j];
                    const int64_t qj = p - sL - g.stride[j];
                    const double2 gj = G[qj >= 0 ? qj : 0];
                    v = cadd(v, cmul(cscale_promoted(A[L * ld + j], sq[kj]), qj >= 0 ? gj : zero));
                }
                G[p] = cdiv_promoted(v, sk);
            }
            t.sync();
        }
    }
}

// G of the all-pivot average (core.py:164-213) with every term, in the
// reference's flat order by one thread (rank 0): a zero-weight read can
// reach any earlier flat position, so no two points are independent a priori.
// A negative index wraps around the lattice; it reads a value only when that
// position is already written (earlier in flat order), else zero, as the
// reference's zero-initialised G holds there.
template <class Ops>
__device__ void exact_stable_forward(const Ops &t, const Lat &g, const double2 *A, const double2 *b, double2 c, double2 *G,
                                     const double *sq) {
    if (t.rank == 0) {
        const int D = g.D;
        const int64_t N = g.N;
        const double2 zero = make_double2(0.0, 0.0);
        int k[8] = {0, 0, 0, 0, 0, 0, 0, 0};
        G[0] = c;
        for (int64_t p = 1; p < N; ++p) {
            for (int a = D - 1; a >= 0; --a) {  // np.ndindex order
                if (++k[a] < g.shape[a]) break;
                k[a] = 0;
            }
            auto rd = [&](int64_t q) {
                if (q < 0) q += N;
                return (q >= 0 && q < p) ? G[q] : zero;
            };
            int num = 0;
            double2 vals = zero;
            for (int i = 0; i < D; ++i) {
                if (k[i] == 0) continue;
                ++num;
                const int64_t pivot = p - g.stride[i];
                double2 val = cmul(b[i], G[pivot]);
                for (int j = 0; j < i; ++j) val = cadd(val, cmul(cscale_promoted(A[i * D + j], sq[k[j]]), rd(pivot - g.stride[j])));
                val = cadd(val, cmul(cscale_promoted(A[i * D + i], sq[k[i] - 1]), rd(pivot - g.stride[i])));
                for (int j = i + 1; j < D; ++j) val = cadd(val, cmul(cscale_promoted(A[i * D + j], sq[k[j]]), rd(pivot - g.stride[j])));
                vals = cadd(vals, cdiv_promoted(val, sq[k[i]]));
            }
            G[p] = cdiv_promoted(vals, double(num));
        }
    }
    t.sync();
}

// dL/dA, dL/db, dL/dc of gradients.py:41-82, one accumulator at a time.
template <class Ops>
__device__ void exact_vjp(const Ops &t, const Lat &g, const double2 *G, const double2 *gg, double2 c, double2 *dA, double2 *db,
                          double2 *dc, const double *sq) {
    const int D = g.D;
    const int64_t N = g.N;
    const double2 zero = make_double2(0.0, 0.0);
    auto wrap = [N](int64_t x) { return ((x % N) + N) % N; };
    // dLdA before the symmetrisation: entries j < i stay the zero matrix
    // entries times the cotangent
    for (int i = 0; i < D; ++i) {
        for (int j = 0; j < D; ++j) {
            double2 acc = zero;
            for (int64_t p = 1 + t.rank; p < N; p += t.size) {
                double2 u = zero;
                if (j == i) {
                    const int ki = lat_digit(g, p, i);
                    if (ki > 1) u = cmul(make_double2(0.5 * sq[ki] * sq[ki - 1], 0.0), G[p - 2 * g.stride[i]]);
                } else if (j > i) {
                    const double w = sq[lat_digit(g, p, i)] * sq[lat_digit(g, p, j)];
                    u = cmul(make_double2(w, 0.0), G[wrap(p - g.stride[i] - g.stride[j])]);
                }
                acc = cadd(acc, cmul(u, gg[p]));
            }
            acc = t.sum(acc);
            if (t.rank == 0) dA[i * D + j] = acc;
        }
    }
    for (int i = 0; i < D; ++i) {
        double2 acc = zero;
        for (int64_t p = 1 + t.rank; p < N; p += t.size) {
            const double2 u = cmul(make_double2(sq[lat_digit(g, p, i)], 0.0), G[wrap(p - g.stride[i])]);
            acc = cadd(acc, cmul(u, gg[p]));
        }
        acc = t.sum(acc);
        if (t.rank == 0) db[i] = acc;
    }
    double2 s = zero;
    for (int64_t p = t.rank; p < N; p += t.size) s = cadd(s, cmul(G[p], gg[p]));
    s = t.sum(s);
    if (t.rank == 0) {
        // (dLdA + dLdA^T) / 2 with numba's complex / int division
        for (int i = 0; i < D; ++i) {
            for (int j = i; j < D; ++j) {
                const double2 m = cdiv_promoted(cadd(dA[i * D + j], dA[j * D + i]), 2.0);
                dA[i * D + j] = m;
                dA[j * D + i] = m;
            }
        }
        dc[0] = cdiv_py(s, c);
    }
    t.sync();
}

}  // namespace mm

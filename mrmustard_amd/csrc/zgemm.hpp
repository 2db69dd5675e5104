// Complex128 matrix product C = A * B on the FP64 matrix cores (MFMA).
//
// The consumer of the amplitudes: a density matrix of n modes, reshaped to
// (bra multi-index) x (ket multi-index), contracted with an operator in the
// number basis (ArrayAnsatz.contract -> math.einsum, array_ansatz.py:159+,
// backend_manager.py:526-548).  Row-major A (M x K, lda), B (K x N, ldb),
// C (M x N, ldc), all interleaved (re, im) doubles.
//
// Tiling: a workgroup of 4 waves owns a 64 x 128 tile of C (8 waves, 128 x 128,
// for 64 < M <= 128); wave w owns the 32 x 64 sub-tile (w / 2, w % 2), i.e.
// 2 x 4 tiles of 16 x 16, each with a
// real and an imaginary accumulator (v_mfma_f64_16x16x4_f64, K = 4 per
// instruction).  A complex product is four real ones:
//   C_re += A_re B_re - A_im B_im,   C_im += A_re B_im + A_im B_re.
// K advances in chunks of 8: the A chunk (64 x 8) and the B chunk (8 x 128)
// are staged in LDS as separate real and imaginary planes, A transposed
// (k-major) so that an operand fragment is 16 consecutive doubles per k.
// Operand layout of the instruction (16 x 4 A, 4 x 16 B, 16 x 16 C):
//   lane l: A[l % 16][l / 16], B[l / 16][l % 16],
//           C[l / 16 + 4 j][l % 16] in accumulator element j
// (measured on gfx950 with a one-wave probe against a host product).
// Tiles past the matrix edge are loaded as zeros and not stored.
#pragma once

#include "common.hpp"

namespace mm {

typedef double dbl4 __attribute__((ext_vector_type(4)));

struct ZgemmArgs {
    const double2 *A, *B;
    double2 *C;
    int64_t M, N, K, lda, ldb, ldc;
    int64_t sA, sB, sC;  // element strides between the products of a batch (grid z; 0 broadcasts an operand)
};

constexpr int ZG_BN = 128, ZG_BK = 8;

// ZG_BM = 64 (4 waves, two workgroups per CU) in general; ZG_BM = 128 (8 waves,
// one workgroup per CU, the same two waves per SIMD) when one tile row covers
// M (64 < M <= 128), so that B is read once instead of twice.
template <int ZG_BM>
__global__ __launch_bounds__(ZG_BM * 4, ZG_BM == 64 ? 2 : 1) void zgemm_mfma_kernel(ZgemmArgs a) {
    constexpr int ZG_THREADS = ZG_BM * 4;
    a.A += int64_t(blockIdx.z) * a.sA;
    a.B += int64_t(blockIdx.z) * a.sB;
    a.C += int64_t(blockIdx.z) * a.sC;
    // LDS planes, two buffers: A transposed [k][m] (rows padded by one double
    // so that the transposing stores of consecutive lanes fall in different
    // banks), B [k][n]
    __shared__ double sAr[2][ZG_BK][ZG_BM + 1], sAi[2][ZG_BK][ZG_BM + 1];
    __shared__ double sBr[2][ZG_BK][ZG_BN], sBi[2][ZG_BK][ZG_BN];
    const int t = threadIdx.x;
    const int lane = t & (MM_WAVE - 1);
    const int wave = t / MM_WAVE;
    const int64_t m0 = int64_t(blockIdx.y) * ZG_BM;
    const int64_t n0 = int64_t(blockIdx.x) * ZG_BN;
    const int wm = (wave >> 1) * 32;  // wave's row offset in the tile
    const int wn = (wave & 1) * 64;   // wave's column offset in the tile
    const int lr = lane & 15, lk = lane >> 4;

    dbl4 cre[2][4], cim[2][4];
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) {
            cre[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
            cim[i][j] = dbl4{0.0, 0.0, 0.0, 0.0};
        }

    // the next chunk is loaded into registers while the current one is used
    constexpr int NA = (ZG_BM * ZG_BK) / ZG_THREADS, NB = (ZG_BK * ZG_BN) / ZG_THREADS;
    double2 pa[NA], pb[NB];
    auto load_chunk = [&](int64_t k0) {
        // A chunk: 64 x 8 complex, rows of 8 contiguous; B chunk: 8 x 128
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = t + i * ZG_THREADS;
            const int r = e / ZG_BK, c = e % ZG_BK;
            const int64_t gm = m0 + r, gk = k0 + c;
            pa[i] = (gm < a.M && gk < a.K) ? a.A[gm * a.lda + gk] : make_double2(0.0, 0.0);
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = t + i * ZG_THREADS;
            const int r = e / ZG_BN, c = e % ZG_BN;
            const int64_t gk = k0 + r, gn = n0 + c;
            pb[i] = (gk < a.K && gn < a.N) ? a.B[gk * a.ldb + gn] : make_double2(0.0, 0.0);
        }
    };
    auto store_chunk = [&](int buf) {
#pragma unroll
        for (int i = 0; i < NA; ++i) {
            const int e = t + i * ZG_THREADS;
            const int r = e / ZG_BK, c = e % ZG_BK;
            sAr[buf][c][r] = pa[i].x;
            sAi[buf][c][r] = pa[i].y;
        }
#pragma unroll
        for (int i = 0; i < NB; ++i) {
            const int e = t + i * ZG_THREADS;
            const int r = e / ZG_BN, c = e % ZG_BN;
            sBr[buf][r][c] = pb[i].x;
            sBi[buf][r][c] = pb[i].y;
        }
    };
    // chunk k is read from buffer k & 1 while chunk k+1 is loaded and then
    // stored into the other buffer, whose last readers (chunk k-1) passed the
    // previous barrier: one barrier per chunk
    load_chunk(0);
    store_chunk(0);
    __syncthreads();
    int buf = 0;
    for (int64_t k0 = 0; k0 < a.K; k0 += ZG_BK) {
        const bool more = k0 + ZG_BK < a.K;
        if (more) load_chunk(k0 + ZG_BK);
#pragma unroll
        for (int ks = 0; ks < ZG_BK; ks += 4) {
            const int kk = ks + lk;
            double ar[2], ai[2], nai[2], br[4], bi[4];
#pragma unroll
            for (int i = 0; i < 2; ++i) {
                ar[i] = sAr[buf][kk][wm + i * 16 + lr];
                ai[i] = sAi[buf][kk][wm + i * 16 + lr];
                nai[i] = -ai[i];
            }
#pragma unroll
            for (int j = 0; j < 4; ++j) {
                br[j] = sBr[buf][kk][wn + j * 16 + lr];
                bi[j] = sBi[buf][kk][wn + j * 16 + lr];
            }
            // all 16 accumulators once, then again: an instruction never
            // waits for the previous one's result
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], br[j], cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ar[i], bi[j], cim[i][j], 0, 0, 0);
                }
#pragma unroll
            for (int i = 0; i < 2; ++i)
#pragma unroll
                for (int j = 0; j < 4; ++j) {
                    cre[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(nai[i], bi[j], cre[i][j], 0, 0, 0);
                    cim[i][j] = __builtin_amdgcn_mfma_f64_16x16x4f64(ai[i], br[j], cim[i][j], 0, 0, 0);
                }
        }
        if (more) store_chunk(buf ^ 1);
        __syncthreads();
        buf ^= 1;
    }
    // store: accumulator element e of lane l is row l / 16 + 4 e, column l % 16
#pragma unroll
    for (int i = 0; i < 2; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int64_t gm = m0 + wm + i * 16 + lk + 4 * e;
                const int64_t gn = n0 + wn + j * 16 + lr;
                if (gm < a.M && gn < a.N) a.C[gm * a.ldc + gn] = make_double2(cre[i][j][e], cim[i][j][e]);
            }
}

}  // namespace mm

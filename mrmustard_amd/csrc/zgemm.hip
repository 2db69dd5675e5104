// libmmhip: C ABI of the FP64 matrix-core complex product (zgemm_mfma_kernel).

#include "host.hpp"
#include "zgemm.hpp"

using namespace mm;

extern "C" {

int mm_zgemm_batched(const double *A, int64_t lda, int64_t sA, const double *B, int64_t ldb, int64_t sB, double *C, int64_t ldc,
                     int64_t sC, int64_t M, int64_t N, int64_t K, int64_t batch, void *stream) {
    if (M < 0 || N < 0 || K < 0 || batch < 0) return fail(MM_ERR_INVALID, "negative matrix size or batch");
    if (M == 0 || N == 0 || batch == 0) return MM_OK;
    if (!A || !B || !C) return fail(MM_ERR_INVALID, "NULL matrix pointer");
    if (lda < K || ldb < N || ldc < N) return fail(MM_ERR_INVALID, "leading dimension smaller than the row length");
    if (sA < 0 || sB < 0 || sC < 0) return fail(MM_ERR_INVALID, "negative batch stride");
    if (batch > 1 && sC < (M - 1) * ldc + N) return fail(MM_ERR_INVALID, "the products of a batch overlap in C");
    const bool tall = M > 64 && M <= 128;  // one 128-row tile covers M: B is read once
    const int bm = tall ? 128 : 64;
    const int64_t gx = (N + ZG_BN - 1) / ZG_BN, gy = (M + bm - 1) / bm;
    if (gx > 0x7fffffff || gy > 65535) return fail(MM_ERR_UNSUPPORTED, "matrix too large for one launch");
    ZgemmArgs a;
    a.M = M;
    a.N = N;
    a.K = K;
    a.lda = lda;
    a.ldb = ldb;
    a.ldc = ldc;
    a.sA = sA;
    a.sB = sB;
    a.sC = sC;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    for (int64_t b0 = 0; b0 < batch; b0 += 65535) {  // grid z holds at most 65535 products
        const int64_t nz = std::min<int64_t>(65535, batch - b0);
        a.A = reinterpret_cast<const double2 *>(A) + b0 * sA;
        a.B = reinterpret_cast<const double2 *>(B) + b0 * sB;
        a.C = reinterpret_cast<double2 *>(C) + b0 * sC;
        if (tall)
            hipLaunchKernelGGL(zgemm_mfma_kernel<128>, dim3(unsigned(gx), unsigned(gy), unsigned(nz)), dim3(512), 0, s, a);
        else
            hipLaunchKernelGGL(zgemm_mfma_kernel<64>, dim3(unsigned(gx), unsigned(gy), unsigned(nz)), dim3(256), 0, s, a);
    }
    LAUNCH_CHECK("zgemm_mfma_kernel");
    return MM_OK;
}

int mm_zgemm(const double *A, int64_t lda, const double *B, int64_t ldb, double *C, int64_t ldc, int64_t M, int64_t N, int64_t K,
             void *stream) {
    return mm_zgemm_batched(A, lda, 0, B, ldb, 0, C, ldc, 0, M, N, K, 1, stream);
}

}  // extern "C"

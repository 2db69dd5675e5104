// libmmhip: C ABI of the Fock cutoffs of Gaussian states (autoshape kernels).

#include "host.hpp"
#include "autoshape.hpp"

using namespace mm;

extern "C" {

size_t mm_autoshape_workspace(int64_t B, int M) { return B > 0 && M > 0 ? size_t(B) * size_t(M) * 8 * 16 : 0; }

int mm_autoshape(const double *A, const double *b, const double *c, int64_t B, int M, double max_prob, int max_shape,
                 int min_shape, int64_t *shape, void *workspace, size_t ws_bytes, void *stream) {
    if (M < 1 || M > AS_MAXM) return fail(MM_ERR_INVALID, "autoshape needs 1 <= M <= %d modes, got %d", AS_MAXM, M);
    if (B < 0) return fail(MM_ERR_INVALID, "negative batch size");
    if (max_shape < 1 || int64_t(max_shape) + 2 >= MM_MAX_CUTOFF)
        return fail(MM_ERR_INVALID, "max_shape must be in [1, %d)", MM_MAX_CUTOFF - 2);
    if (B == 0) return MM_OK;
    if (!A || !b || !c || !shape) return fail(MM_ERR_INVALID, "NULL array pointer");
    if (B * M > 0x7fffffff) return fail(MM_ERR_UNSUPPORTED, "too many states for one launch");
    if (!workspace || ws_bytes < mm_autoshape_workspace(B, M))
        return fail(MM_ERR_WORKSPACE, "autoshape workspace of %zu bytes needed, %zu given", mm_autoshape_workspace(B, M), ws_bytes);
    DevState *ds = nullptr;
    int rc = device_state(&ds);
    if (rc) return rc;
    AutoshapeArgs a;
    a.marg = reinterpret_cast<double2 *>(workspace);
    a.A = reinterpret_cast<const double2 *>(A);
    a.b = reinterpret_cast<const double2 *>(b);
    a.c = reinterpret_cast<const double2 *>(c);
    a.B = B;
    a.M = M;
    a.max_prob = max_prob;
    a.max_shape = max_shape;
    a.min_shape = min_shape;
    a.sq = ds->sq;
    a.shape = shape;
    hipStream_t s = reinterpret_cast<hipStream_t>(stream);
    hipLaunchKernelGGL(autoshape_schur_kernel, dim3(unsigned(B * M)), dim3(64), 0, s, a);
    hipLaunchKernelGGL(autoshape_recursion_kernel, dim3(unsigned((B * M + 255) / 256)), dim3(256), 0, s, a);
    LAUNCH_CHECK("autoshape kernels");
    return MM_OK;
}

}  // extern "C"

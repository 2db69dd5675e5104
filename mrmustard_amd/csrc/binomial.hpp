// The binomial strategy (mrmustard/math/lattice/strategies/binomial.py:30-74):
// the lattice filled by weight |k| = sum_j k_j, stopping after the first
// weight at which the accumulated squared norm exceeds max_l2, or at weight
// global_cutoff - 1.  Each value is the first-pivot step of the vanilla
// recurrence (steps.py:37-67), so the amplitudes are vanilla's; this build
// computes the whole box with the vanilla sweep and then
//   weight_norm_kernel:  per-block sums of |G_k|^2 by weight (LDS histogram),
//   binomial_stop_kernel: the partials summed in block order, the running norm
//                         |c|^2 + sum_w n_w as the reference accumulates it,
//                         and the stopping weight,
//   weight_mask_kernel:  zeros beyond the stopping weight (the reference's
//                        tensor starts as zeros and is never written there).
#pragma once

#include "global_kernels.hpp"

namespace mm {

constexpr int BINOM_MAX_W = 8192;  // weights 0 .. W-1 held in one LDS histogram

template <int D>
__device__ __forceinline__ int weight_of(int64_t q, const Geom &g) {
    int w = 0;
#pragma unroll
    for (int j = 0; j < D; ++j) {
        const int64_t dj = q / g.stride[j];
        q -= dj * g.stride[j];
        w += int(dj);
    }
    return w;
}

template <int D>
__global__ __launch_bounds__(256) void weight_norm_kernel(const double2 *G, Geom g, int W, int64_t chunk, double *part) {
    extern __shared__ double h[];
    for (int w = threadIdx.x; w < W; w += blockDim.x) h[w] = 0.0;
    __syncthreads();
    const int64_t lo = int64_t(blockIdx.x) * chunk;
    const int64_t hi = lo + chunk < g.N ? lo + chunk : g.N;
    for (int64_t q = lo + threadIdx.x; q < hi; q += blockDim.x) {
        const double2 v = G[q];
        const double a = hypot(v.x, v.y);
        atomicAdd(&h[weight_of<D>(q, g)], a * a);
    }
    __syncthreads();
    for (int w = threadIdx.x; w < W; w += blockDim.x) part[int64_t(blockIdx.x) * W + w] = h[w];
}

// One block: weight sums over the blocks (in block order), then the
// reference's loop over weights 1 .. global_cutoff-1 (binomial.py:56-72).
__global__ __launch_bounds__(256) void binomial_stop_kernel(const double *part, int nblk, int W, const double2 *c, double max_l2,
                                                            int64_t global_cutoff, int *wstop, double *norm_out) {
    extern __shared__ double tot[];
    for (int w = threadIdx.x; w < W; w += blockDim.x) {
        double s = 0.0;
        for (int bk = 0; bk < nblk; ++bk) s += part[int64_t(bk) * W + w];
        tot[w] = s;
    }
    __syncthreads();
    if (threadIdx.x != 0) return;
    const double ac = hypot(c->x, c->y);
    double norm = ac * ac;
    int64_t last = 0;
    for (int64_t w = 1; w < global_cutoff; ++w) {
        if (w < W) norm += tot[w];
        last = w;
        if (norm > max_l2) break;
    }
    *wstop = int(last < W ? last : W - 1);
    if (norm_out) *norm_out = norm;
}

template <int D>
__global__ __launch_bounds__(256) void weight_mask_kernel(double2 *G, Geom g, const int *wstop) {
    const int ws = *wstop;
    const int64_t inc = int64_t(gridDim.x) * blockDim.x;
    for (int64_t q = int64_t(blockIdx.x) * blockDim.x + threadIdx.x; q < g.N; q += inc)
        if (weight_of<D>(q, g) > ws) G[q] = make_double2(0.0, 0.0);
}

}  // namespace mm

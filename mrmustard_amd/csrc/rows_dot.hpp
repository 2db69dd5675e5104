// out[r] = sum_k G[r * K + k] w[k]: the contraction of a block of lattice
// rows with a vector over their trailing (flattened) indices, run on each
// slab as soon as it is computed so that the lattice is never stored whole
// (CircuitComponent.fock_array's derived-variables branch,
// mrmustard/lab/circuit_components.py:516-528, contracts the lattice on
// shape + shape_derived_vars with c over the derived indices).
//
// HBM-bound: 16 K bytes read per row (the slab, usually still in L2 / MALL
// after the kernel that wrote it) plus 16 bytes written.  Rows of K >= 64
// entries take one wave each (lanes strided over k, then a fixed butterfly
// over the 64 lanes); shorter rows take one lane each, summed in index order.
#pragma once

#include "common.hpp"

namespace mm {

__global__ __launch_bounds__(256) void rows_dot_wave_kernel(const double2 *G, const double2 *w, double2 *out, int64_t rows,
                                                            int64_t K) {
    const int lane = threadIdx.x & 63;
    const int64_t row = (int64_t(blockIdx.x) * blockDim.x + threadIdx.x) >> 6;
    if (row >= rows) return;  // uniform per wave
    const double2 *g = G + row * K;
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t k = lane; k < K; k += 64) acc = cfma(g[k], w[k], acc);
#pragma unroll
    for (int o = 32; o > 0; o >>= 1) {
        acc.x += __shfl_xor(acc.x, o);
        acc.y += __shfl_xor(acc.y, o);
    }
    if (lane == 0) out[row] = acc;
}

__global__ __launch_bounds__(256) void rows_dot_lane_kernel(const double2 *G, const double2 *w, double2 *out, int64_t rows,
                                                            int64_t K) {
    const int64_t row = int64_t(blockIdx.x) * blockDim.x + threadIdx.x;
    if (row >= rows) return;
    const double2 *g = G + row * K;
    double2 acc = make_double2(0.0, 0.0);
    for (int64_t k = 0; k < K; ++k) acc = cfma(g[k], w[k], acc);
    out[row] = acc;
}

}  // namespace mm

// Launchers of the LDS sweep kernels, one translation unit per kernel
// instantiation (lds_*.hip) so that they compile in parallel.
#pragma once

#include "lds_kernel.hpp"

namespace mm {

// Raise a kernel's dynamic LDS limit to `bytes` (needed above 64 KiB).
template <typename K>
void raise_lds_limit(K kernel, int bytes) {
    (void)hipFuncSetAttribute(reinterpret_cast<const void *>(kernel), hipFuncAttributeMaxDynamicSharedMemorySize, bytes);
}

// lds_kernel<DMAX, FUSED, MAXT>: DMAX = 3 or 8, FUSED = forward + VJP.
void launch_lds_3f(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);  // <3, true, 512>
void launch_lds_3(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);   // <3, false, 768>
void launch_lds_8f(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);  // <8, true, 256>
void launch_lds_8(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);   // <8, false, 512>
// lds_block_kernel<D, false, 1024, 1>, D = 2 or 3 (forward only).
void launch_lds_block(int D, const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);
// lds_block_kernel<3, false, 1024, 2>: a single D = 3 sub-lattice with 1024-2048-point slabs.
void launch_lds_block2(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s);

}  // namespace mm

// libmmhip: the LDS sweep kernel lds_kernel<3, false, 768> in a translation unit of its own.
#include "lds_launch.hpp"

namespace mm {

void launch_lds_3(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s) {
    raise_lds_limit(lds_kernel<3, false, 768>, int(shmem));
    hipLaunchKernelGGL((lds_kernel<3, false, 768>), grid, block, shmem, s, a);
}

void launch_lds_block2(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s) {
    raise_lds_limit(lds_block_kernel<3, false, 1024, 2>, int(shmem));
    hipLaunchKernelGGL((lds_block_kernel<3, false, 1024, 2>), grid, block, shmem, s, a);
}

void launch_lds_block(int D, const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s) {
    if (D == 2) {
        raise_lds_limit(lds_block_kernel<2, false, 1024, 1>, int(shmem));
        hipLaunchKernelGGL((lds_block_kernel<2, false, 1024, 1>), grid, block, shmem, s, a);
    } else {
        raise_lds_limit(lds_block_kernel<3, false, 1024, 1>, int(shmem));
        hipLaunchKernelGGL((lds_block_kernel<3, false, 1024, 1>), grid, block, shmem, s, a);
    }
}

}  // namespace mm

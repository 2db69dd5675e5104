// libmmhip: the LDS sweep kernel lds_kernel<8, false, 512> in a translation unit of its own.
#include "lds_launch.hpp"

namespace mm {

void launch_lds_8(const LdsArgs &a, dim3 grid, dim3 block, size_t shmem, hipStream_t s) {
    raise_lds_limit(lds_kernel<8, false, 512>, int(shmem));
    hipLaunchKernelGGL((lds_kernel<8, false, 512>), grid, block, shmem, s, a);
}

}  // namespace mm

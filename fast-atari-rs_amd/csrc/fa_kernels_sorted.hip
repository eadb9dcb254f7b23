// fa_kernels_sorted.hip -- the opcode-grouping instantiation of the frame
// kernel (set_sorted).  Lanes are served one distinct opcode at a time
// (64-bit ballot + readlane) while the groups are large; a translation unit
// of its own so that each kernel gets its own launch attributes.
#include "fa_frame.hpp"

namespace fa {

hipError_t prepare_sorted_kernel(size_t max_lds) {
    return hipFuncSetAttribute((const void *)fa_frame_kernel<true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                               (int)max_lds);
}

hipError_t launch_sorted(const StepArgs &g, uint32_t bs, uint32_t grid, size_t lds, hipStream_t s) {
    hipLaunchKernelGGL(fa_frame_kernel<true>, dim3(grid), dim3(bs), lds, s, g);
    return hipGetLastError();
}

}  // namespace fa

// Batch step kernel, 16 contact slots: the worlds of several models (kinds of different dof
// counts, at most 48 padded dofs) in one launch, each world on one wave (see sapx_wave_impl.h).
#include "sapx_wave_impl.h"

namespace sapx {
namespace wv {

__global__ __launch_bounds__(64, 1) void sapx_batch_kernel_16(BatchMap B) {
    const int b = (int)blockIdx.x;
    int g = 0;
#pragma unroll
    for (int k = 1; k < kMaxBatchGroups; ++k)
        if (k < B.groups && b >= B.block_start[k]) g = k;
    const StepParams& P = B.params[g];
    const int local = b - B.block_start[g];
    switch (B.npad[g]) {
        case 16: wave_body<16, 16>(P, local); break;
        case 32: wave_body<32, 16>(P, local); break;
        default: wave_body<48, 16>(P, local); break;
    }
}

}  // namespace wv

hipError_t launch_batch_16(const BatchMap& B, int blocks, size_t smem, hipStream_t stream) {
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)wv::sapx_batch_kernel_16, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL(wv::sapx_batch_kernel_16, dim3(blocks), dim3(64), smem, stream, B);
    return hipGetLastError();
}

}  // namespace sapx

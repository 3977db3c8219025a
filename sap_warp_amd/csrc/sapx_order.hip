// Dispatch order of the worlds for the step kernels (gfx950).
//
// Worlds are independent and their step cost varies several-fold (Newton
// iterations, contact count).  Workgroups start in blockIdx order, so handing
// the step kernels a permutation with the most expensive worlds first shortens
// the tail of the launch (longest-processing-time-first list scheduling).  The
// estimate uses the previous step's Newton count of each world and this step's
// contact count; it only changes which wave slot a world runs in, never a
// world's arithmetic.
#include <hip/hip_runtime.h>

#include "sapx_params.h"

namespace sapx {

namespace {

constexpr int kOrderThreads = 1024;
constexpr int kOrderBins = 256;

// estimated step cost of one world in units of ~10k cycles: fixed part (kinematics,
// dynamics, free motion, contact setup, integration) plus per-Newton-iteration work
// that grows with the contact count; bin 0 = most expensive
__device__ __forceinline__ int order_bin(const int* cworld_start, const int* iters, int w, int C) {
    int nc = cworld_start ? cworld_start[w + 1] - cworld_start[w] : 0;
    nc = nc < 0 ? 0 : (nc > C ? C : nc);
    const int it = iters[w] < 0 ? 0 : (iters[w] > 255 ? 255 : iters[w]);
    int est = (100 + it * (30 + nc)) / 5;
    est = est > kOrderBins - 1 ? kOrderBins - 1 : est;
    return kOrderBins - 1 - est;
}

__global__ __launch_bounds__(kOrderThreads) void sapx_world_order_kernel(const int* cworld_start, const int* iters,
                                                                        int W, int C, int* order) {
    __shared__ int count[kOrderBins];
    __shared__ int base[kOrderBins];
    const int tid = threadIdx.x;
    for (int b = tid; b < kOrderBins; b += kOrderThreads) count[b] = 0;
    __syncthreads();
    for (int w = tid; w < W; w += kOrderThreads) atomicAdd(&count[order_bin(cworld_start, iters, w, C)], 1);
    __syncthreads();
    if (tid < 64) {
        // exclusive scan of the 256 bin counts: four bins per lane, then a wave scan
        int c[4], s = 0;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            c[k] = count[tid * 4 + k];
            s += c[k];
        }
        int incl = s;
#pragma unroll
        for (int d = 1; d < 64; d <<= 1) {
            const int t = __shfl_up(incl, d, 64);
            if (tid >= d) incl += t;
        }
        int run = incl - s;
#pragma unroll
        for (int k = 0; k < 4; ++k) {
            base[tid * 4 + k] = run;
            run += c[k];
        }
    }
    __syncthreads();
    for (int w = tid; w < W; w += kOrderThreads) {
        const int pos = atomicAdd(&base[order_bin(cworld_start, iters, w, C)], 1);
        order[pos] = w;
    }
}

}  // namespace

hipError_t launch_world_order(const int* cworld_start, const int* iters, int W, int C, int* order, hipStream_t stream) {
    hipLaunchKernelGGL(sapx_world_order_kernel, dim3(1), dim3(kOrderThreads), 0, stream, cworld_start, iters, W, C,
                       order);
    return hipGetLastError();
}

}  // namespace sapx

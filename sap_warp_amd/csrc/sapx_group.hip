// Grouping of an arbitrary flat rigid-contact list by world, for callers that
// hand SolverSAP a contact list without per-world offsets.  The reference
// assigns per-world slots with atomics (contact_jacobian.py:404-476,
// _contact_env_from_shapes sap_helpers.py:439-464); here a stable radix sort
// on the world id keeps flat order inside each world, so results are
// deterministic.  Contacts that span two worlds or none are dropped.
#include <hipcub/hipcub.hpp>

#include "sapx_params.h"

namespace sapx {

__global__ void sapx_group_keys_kernel(GroupParams G) {
    const int i = blockIdx.x * blockDim.x + threadIdx.x;
    if (i >= G.capacity) return;
    int key = G.W;
    if (i < G.count[0]) {
        const int s0 = G.in_shape0[i], s1 = G.in_shape1[i];
        const int b0 = s0 >= 0 ? G.shape_body[s0] : -1;
        const int b1 = s1 >= 0 ? G.shape_body[s1] : -1;
        int env = b0 >= 0 ? b0 / G.NB : (b1 >= 0 ? b1 / G.NB : -1);
        if (env >= 0 && b0 >= 0 && b0 / G.NB != env) env = -1;
        if (env >= 0 && b1 >= 0 && b1 / G.NB != env) env = -1;
        if (env >= 0 && env < G.W) key = env;
    }
    G.keys[i] = key;
    G.vals[i] = i;
}

__global__ void sapx_group_starts_kernel(GroupParams G) {
    const int w = blockIdx.x * blockDim.x + threadIdx.x;
    if (w > G.W) return;
    int lo = 0, hi = G.capacity;
    while (lo < hi) {
        const int mid = (lo + hi) >> 1;
        if (G.keys_out[mid] < w) lo = mid + 1;
        else hi = mid;
    }
    G.world_start[w] = lo;
}

__global__ void sapx_group_gather_kernel(GroupParams G) {
    const int k = blockIdx.x * blockDim.x + threadIdx.x;
    if (k >= G.capacity || G.keys_out[k] >= G.W) return;
    const int i = G.vals_out[k];
    G.out_shape0[k] = G.in_shape0[i];
    G.out_shape1[k] = G.in_shape1[i];
    for (int e = 0; e < 3; ++e) {
        G.out_p0[k * 3 + e] = G.in_p0[i * 3 + e];
        G.out_p1[k * 3 + e] = G.in_p1[i * 3 + e];
        G.out_n[k * 3 + e] = G.in_n[i * 3 + e];
    }
    G.out_m0[k] = G.in_m0[i];
    G.out_m1[k] = G.in_m1[i];
}

hipError_t launch_group_contacts(const GroupParams& G, hipStream_t stream) {
    const int blk = 256;
    const int grid = (G.capacity + blk - 1) / blk;
    hipLaunchKernelGGL(sapx_group_keys_kernel, dim3(grid), dim3(blk), 0, stream, G);
    size_t tmp = G.cub_tmp_bytes;
    hipError_t e = hipcub::DeviceRadixSort::SortPairs(G.cub_tmp, tmp, G.keys, G.keys_out, G.vals, G.vals_out,
                                                      G.capacity, 0, 32, stream);
    if (e != hipSuccess) return e;
    hipLaunchKernelGGL(sapx_group_starts_kernel, dim3((G.W + 1 + blk - 1) / blk), dim3(blk), 0, stream, G);
    hipLaunchKernelGGL(sapx_group_gather_kernel, dim3(grid), dim3(blk), 0, stream, G);
    return hipGetLastError();
}

}  // namespace sapx

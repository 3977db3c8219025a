// Blocked Cholesky factor and solve of large symmetric positive definite matrices (gfx950).
//
// Counterpart of BlockCholeskySolverBatched (blocked_cholesky.py:725-901), whose
// factorisation runs one workgroup per matrix (blocked_cholesky.py:803-809): fine for a
// 35-dof world, a serial bottleneck for a single world of thousands of dofs.  Here every
// stage spreads over the whole GPU:
//
//   right-looking, 64 x 64 blocks, float32, per block column k:
//     panel        (T - k workgroups)  each factors the diagonal block A_kk in LDS and
//                                      inverts L_kk (redundantly: no grid barrier);
//                                      workgroup 0 stores L_kk and L_kk^-1, the others the
//                                      panel blocks L_ik = A_ik L_kk^-T by MFMA
//     Schur update                     A_ij -= L_ik L_jk^T of every lower trailing block,
//                                      v_mfma_f32_16x16x4f32, split with look-ahead: the
//                                      tiles of column k + 1 first (chol_schur_next), then
//                                      the other tiles in the same launch as panel k + 1
//                                      (chol_panel_schur), which touches none of them
//   solves, one launch each: workgroup i computes block i of y (forward) or x (backward)
//   as soon as the blocks it depends on are published (per-block ready flags with
//   release / acquire ordering and a bounded wait), using the stored diagonal inverses, so
//   the dependent chain is one 64 x 64 mat-vec per block; then float64-residual refinement
//   sweeps r = b - A x, x += solve(r) against a kept copy of the input.
//   The factor is of D A D with D = diag(A)^-1/2 (the solves scale in and out), which
//   takes the spread of masses, inertias and contact stiffnesses out of the float32
//   condition number.
//
// Rows and columns from n to the padded order are identity, as in the reference's
// padded layout (blocked_cholesky.py:299-313).
#include <hip/hip_runtime.h>

#include <cfloat>
#include <cstdint>
#include <string>

#include "../../include/sapx.h"

namespace sapx {
void set_last_error(const char* msg);
}

namespace {

constexpr int NB = 64;       // block order
constexpr int LS = NB + 1;   // LDS row stride (floats): conflict-free column access
constexpr int NT = 256;      // threads per workgroup
constexpr int kSpinLimit = 1 << 24;  // polls before a wait gives up (about a second)
constexpr int kMaxResident = 512;    // workgroups of one solve launch (all resident at once)

typedef float f4 __attribute__((ext_vector_type(4)));

int32_t cfail(int32_t code, const std::string& msg) {
    sapx::set_last_error(msg.c_str());
    return code;
}

__device__ __forceinline__ int flag_load(const int* p) {
    return __hip_atomic_load(p, __ATOMIC_ACQUIRE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ void flag_store(int* p, int v) {
    __hip_atomic_store(p, v, __ATOMIC_RELEASE, __HIP_MEMORY_SCOPE_AGENT);
}
__device__ __forceinline__ float coherent_load(const float* p) {
    return __hip_atomic_load(p, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
}

__device__ __forceinline__ float bcast(float v, int lane) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(v), lane));
}

constexpr int SB = 16;       // sub-block order inside the 64 x 64 diagonal block
constexpr int YS = SB + 1;   // LDS row stride of the 16 x 16 work tiles

// acc += P Q^T (qt) or P Q over one 16 x 16 x 16 product on one wave (v_mfma_f32_16x16x4f32):
// P rows at P + i * ps, Q rows at Q + r * qs; acc[e] is C[4 (lane >> 4) + e][lane & 15]
__device__ __forceinline__ void mfma16(const float* P, int ps, const float* Q, int qs, bool qt, f4& acc) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int kk = 0; kk < SB; kk += 4) {
        const int k = kk + (lane >> 4), r = lane & 15;
        const float a = P[r * ps + k];
        const float b = qt ? Q[r * qs + k] : Q[k * qs + r];
        acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a, b, acc, 0, 0, 0);
    }
}
__device__ __forceinline__ void store16(float* C, int cs, const f4& acc, float scale) {
    const int lane = threadIdx.x & 63;
#pragma unroll
    for (int e = 0; e < 4; ++e) C[(4 * (lane >> 4) + e) * cs + (lane & 15)] = scale * acc[e];
}

// Factor the 64 x 64 diagonal block held in D (lower triangle) and invert the factor, on the
// workgroup's four waves, in 16 x 16 sub-blocks: per sub-block column, wave 0 factors the
// 16 x 16 diagonal block and inverts it in registers (lane r = row r, other rows by readlane),
// the waves form the sub-blocks below it (A Y^T, MFMA) and apply their Schur updates (MFMA);
// then X = L^-1 by block forward substitution, X_ij = -Y_i sum_{m=j..i-1} L_im X_mj, one
// column of blocks per wave.  L stays in D and X goes to Xs (upper triangles are zeroed when
// stored).  Y (4 x 16 x YS) and tmp (4 x 16 x YS) are LDS work tiles.  Returns the first column
// whose pivot was not positive (replaced by FLT_MIN), or -1 (valid on wave 0).
__device__ __forceinline__ int diag_factor_invert(float* D, float* Xs, float* Y, float* tmp) {
    const int lane = threadIdx.x & 63, wave = threadIdx.x >> 6;
    int bad = -1;
    for (int kb = 0; kb < NB / SB; ++kb) {
        const int k0 = kb * SB;
        if (wave == 0) {
            const int r = lane & 15;
            float a[SB];
#pragma unroll
            for (int c = 0; c < SB; ++c) a[c] = D[(k0 + r) * LS + k0 + c];
            float dinv = 1.0f;
#pragma unroll
            for (int j = 0; j < SB; ++j) {
                const float d = bcast(a[j], j);
                const bool ok = d > 0.0f && d <= FLT_MAX;
                if (!ok && bad < 0) bad = k0 + j;
                const float sq = sqrtf(ok ? d : FLT_MIN);
                const float inv = 1.0f / sq;
                dinv = r == j ? inv : dinv;
                a[j] = r == j ? sq : (r > j ? a[j] * inv : a[j]);
#pragma unroll
                for (int l = j + 1; l < SB; ++l) a[l] -= a[j] * bcast(a[j], l);
            }
            float x[SB];
#pragma unroll
            for (int c = 0; c < SB; ++c) x[c] = r == c ? 1.0f : 0.0f;
#pragma unroll
            for (int k = 0; k < SB; ++k) {
#pragma unroll
                for (int c = 0; c <= k; ++c) x[c] = r == k ? x[c] * dinv : x[c];
#pragma unroll
                for (int c = 0; c <= k; ++c) {
                    const float xk = bcast(x[c], k);
                    x[c] = r > k ? x[c] - a[k] * xk : x[c];
                }
            }
            if (lane < SB) {
#pragma unroll
                for (int c = 0; c < SB; ++c) {
                    D[(k0 + r) * LS + k0 + c] = c <= r ? a[c] : 0.0f;
                    Y[(kb * SB + r) * YS + c] = c <= r ? x[c] : 0.0f;
                }
            }
        }
        __syncthreads();
        // sub-blocks below: L_ib = A_ib Y^T
        const int below = NB / SB - 1 - kb;
        if (wave < below) {
            const int i0 = k0 + SB * (wave + 1);
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            mfma16(D + i0 * LS + k0, LS, Y + kb * SB * YS, YS, true, acc);
            store16(tmp + wave * SB * YS, YS, acc, 1.0f);
        }
        __syncthreads();
        if (wave < below) {
            const int i0 = k0 + SB * (wave + 1);
            for (int e = lane; e < SB * SB; e += 64) {
                const int rr = e >> 4, cc = e & 15;
                D[(i0 + rr) * LS + k0 + cc] = tmp[(wave * SB + rr) * YS + cc];
            }
        }
        __syncthreads();
        // Schur updates of the trailing sub-blocks (i, j), kb < j <= i
        const int ntile = below * (below + 1) / 2;
        for (int t = wave; t < ntile; t += 4) {
            int ib = 0, rem = t;
            while (rem > ib) {
                rem -= ib + 1;
                ++ib;
            }
            const int i0 = k0 + SB * (ib + 1), j0 = k0 + SB * (rem + 1);
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            mfma16(D + i0 * LS + k0, LS, D + j0 * LS + k0, LS, true, acc);
#pragma unroll
            for (int e = 0; e < 4; ++e) D[(i0 + 4 * (lane >> 4) + e) * LS + j0 + (lane & 15)] -= acc[e];
        }
        __syncthreads();
    }
    // X = L^-1: diagonal blocks Y_i, below them X_ij for i > j (wave j), upper blocks zero
    for (int e = threadIdx.x; e < NB * NB; e += NT) {
        const int r = e >> 6, c = e & 63, bi = r >> 4, bj = c >> 4;
        Xs[r * LS + c] = bi == bj ? Y[(bi * SB + (r & 15)) * YS + (c & 15)] : 0.0f;
    }
    __syncthreads();
    {
        const int j = wave, j0 = j * SB;
        float* tw = tmp + wave * SB * YS;
        for (int i = j + 1; i < NB / SB; ++i) {
            const int i0 = i * SB;
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
            for (int m = j; m < i; ++m) mfma16(D + i0 * LS + m * SB, LS, Xs + m * SB * LS + j0, LS, false, acc);
            store16(tw, YS, acc, 1.0f);
            __builtin_amdgcn_wave_barrier();
            f4 acc2 = {0.0f, 0.0f, 0.0f, 0.0f};
            mfma16(Y + i * SB * YS, YS, tw, YS, false, acc2);
            __builtin_amdgcn_wave_barrier();
            store16(Xs + i0 * LS + j0, LS, acc2, -1.0f);
            __builtin_amdgcn_wave_barrier();
        }
    }
    __syncthreads();
    return bad;
}

// publish: every thread's stores are visible device-wide before the flag
__device__ __forceinline__ void publish(int* flag) {
    __threadfence();
    __syncthreads();
    if (threadIdx.x == 0) flag_store(flag, 1);
}

// symmetric diagonal scaling d_i = 1 / sqrt(A_ii) (1 where A_ii is not positive, and on the
// padding): the factor works on D A D, whose unit diagonal removes the scale spread of masses,
// rotational inertias and stiff contact blocks from the float32 condition number
__global__ __launch_bounds__(NT) void chol_scale(const float* __restrict__ A, int lda, long long a_stride, int n,
                                                 float* __restrict__ dsc, int npad) {
    const int b = blockIdx.y;
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= npad) return;
    const float a = i < n ? A[b * a_stride + (long long)i * lda + i] : 1.0f;
    dsc[(size_t)b * npad + i] = (a > 0.0f && a <= FLT_MAX) ? 1.0f / sqrtf(a) : 1.0f;
}

// the scaled padded work matrix D A D and the kept original A (identity outside n x n)
__global__ __launch_bounds__(NT) void chol_load(const float* __restrict__ A, int lda, long long a_stride, int n,
                                                const float* __restrict__ dsc, float* __restrict__ W,
                                                float* __restrict__ A0, int npad, int* __restrict__ info) {
    const int b = blockIdx.y;
    const long long total = (long long)npad * npad;
    const float* Ab = A + b * a_stride;
    const float* db = dsc + (size_t)b * npad;
    for (long long e = (long long)blockIdx.x * NT + threadIdx.x; e < total; e += (long long)gridDim.x * NT) {
        const int i = (int)(e / npad), j = (int)(e - (long long)i * npad);
        const float v = (i < n && j < n) ? Ab[(long long)i * lda + j] : (i == j ? 1.0f : 0.0f);
        W[b * total + e] = (v * db[i]) * db[j];
        A0[b * total + e] = v;
    }
    if (blockIdx.x == 0 && threadIdx.x == 0) info[b] = 0;
}

// Block structure: bit (i, j) of the mask (row i: words of kMaskWords bits) says that block
// (i, j), j <= i, of the work matrix may be nonzero.  chol_mask marks the loaded blocks with a
// nonzero entry (one workgroup per lower block), chol_symbolic adds the fill of the block
// elimination (block (i, j) fills when blocks (i, k) and (j, k) are nonzero for some k < j).  A
// block outside the mask is exactly zero throughout the factorization, so the kernels skip its
// panel product and every Schur update that would subtract a product with it (each such update
// subtracts exact zeros); a dense matrix has every bit set and runs every tile.
constexpr int kMaskWords = (kMaxResident + 31) / 32;

__device__ __forceinline__ bool mask_bit(const unsigned* m, int i, int j) {
    return (m[i * kMaskWords + (j >> 5)] >> (j & 31)) & 1u;
}

// first block column >= from (and < limit) of mask row `row` with its bit set; limit if none
__device__ __forceinline__ int mask_next(const unsigned* row, int from, int limit) {
    for (int w = from >> 5; w * 32 < limit; ++w) {
        unsigned bits = row[w];
        if (w == (from >> 5)) bits &= ~0u << (from & 31);
        if (bits) {
            const int j = w * 32 + __builtin_ctz(bits);
            return j < limit ? j : limit;
        }
    }
    return limit;
}

__global__ __launch_bounds__(NT) void chol_mask(const float* __restrict__ W, int npad, int T,
                                                unsigned* __restrict__ mask) {
    const int b = blockIdx.y;
    int ib, jb;
    {
        const int t = (int)blockIdx.x;
        int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
        while ((i + 1) * (i + 2) / 2 <= t) ++i;
        while (i * (i + 1) / 2 > t) --i;
        ib = i;
        jb = t - i * (i + 1) / 2;
    }
    const float* Wb = W + (size_t)b * npad * npad;
    bool nz = false;
    for (int e = threadIdx.x; e < NB * NB; e += NT) {
        const int r = e >> 6, c = e & 63;
        nz = nz || Wb[(size_t)(ib * NB + r) * npad + jb * NB + c] != 0.0f;
    }
    nz = __syncthreads_or(nz);
    unsigned* m = mask + (size_t)b * T * kMaskWords;
    if (threadIdx.x == 0) {
        const unsigned bit = 1u << (jb & 31);
        if (nz || ib == jb) atomicOr(m + ib * kMaskWords + (jb >> 5), bit);
    }
}

__global__ __launch_bounds__(1024) void chol_symbolic(int T, unsigned* __restrict__ mask) {
    __shared__ unsigned sm[kMaxResident * kMaskWords];
    unsigned* m = mask + (size_t)blockIdx.x * T * kMaskWords;
    for (int e = threadIdx.x; e < T * kMaskWords; e += 1024) sm[e] = m[e];
    __syncthreads();
    for (int k = 0; k + 1 < T; ++k) {
        const int r = T - k - 1;  // rows and columns after k
        for (int t = threadIdx.x; t < r * (r + 1) / 2; t += 1024) {
            int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
            while ((i + 1) * (i + 2) / 2 <= t) ++i;
            while (i * (i + 1) / 2 > t) --i;
            const int j = t - i * (i + 1) / 2;
            const int ig = k + 1 + i, jg = k + 1 + j;
            if (((sm[ig * kMaskWords + (k >> 5)] >> (k & 31)) & 1u) && ((sm[jg * kMaskWords + (k >> 5)] >> (k & 31)) & 1u))
                atomicOr(&sm[ig * kMaskWords + (jg >> 5)], 1u << (jg & 31));
        }
        __syncthreads();
    }
    for (int e = threadIdx.x; e < T * kMaskWords; e += 1024) m[e] = sm[e];
}

// 64 x 64 x 64 product on four waves: out = P Q^T with P, Q in LDS (row stride LS); wave w
// owns the 32 x 32 quadrant (w >> 1, w & 1) as 2 x 2 MFMA tiles
__device__ __forceinline__ void mfma_pqt(const float* P, const float* Q, f4 (&acc)[2][2]) {
    const int lane = threadIdx.x & 63, w = threadIdx.x >> 6;
    const int qr = (w >> 1) * 32, qc = (w & 1) * 32;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c) acc[a][c] = f4{0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll 4
    for (int kk = 0; kk < NB; kk += 4) {
        const int col = kk + (lane >> 4);
        float pa[2], qb[2];
#pragma unroll
        for (int t = 0; t < 2; ++t) {
            pa[t] = P[(qr + t * 16 + (lane & 15)) * LS + col];
            qb[t] = Q[(qc + t * 16 + (lane & 15)) * LS + col];
        }
#pragma unroll
        for (int a = 0; a < 2; ++a)
#pragma unroll
            for (int c = 0; c < 2; ++c) acc[a][c] = __builtin_amdgcn_mfma_f32_16x16x4f32(pa[a], qb[c], acc[a][c], 0, 0, 0);
    }
}

// LDS of the panel role (D, X, Y, tmp) and of the Schur role (two 64 x 64 operands)
constexpr int kPanelFloats = 2 * NB * LS + NB * YS + 4 * SB * YS;
constexpr int kSmemFloats = kPanelFloats > 2 * NB * LS ? kPanelFloats : 2 * NB * LS;

// Block column k, block row k + blk: factor and invert the diagonal block (every panel
// workgroup, redundantly: no grid barrier); blk 0 stores L_kk and L_kk^-1, the others the
// panel block L_ik = A_ik L_kk^-T (MFMA).
__device__ __forceinline__ void panel_body(float* smem, int* first_bad, float* Wb, float* Xi, int b, int npad, int T, int k, int blk,
                           int* info) {
    float* D = smem;
    float* X = smem + NB * LS;
    float* Y = smem + 2 * NB * LS;
    float* tmp = Y + NB * YS;
    const int tid = threadIdx.x;
    const int k0 = k * NB;
    for (int e = tid; e < NB * NB; e += NT) {
        const int i = e >> 6, j = e & 63;
        D[i * LS + j] = Wb[(size_t)(k0 + i) * npad + k0 + j];
    }
    __syncthreads();
    {
        const int bad = diag_factor_invert(D, X, Y, tmp);
        if (tid == 0) *first_bad = bad;
    }
    __syncthreads();
    if (blk == 0) {
        float* Xo = Xi + ((size_t)b * T + k) * NB * NB;
        for (int e = tid; e < NB * NB; e += NT) {
            const int i = e >> 6, j = e & 63;
            Wb[(size_t)(k0 + i) * npad + k0 + j] = j <= i ? D[i * LS + j] : 0.0f;
            Xo[e] = j <= i ? X[i * LS + j] : 0.0f;
        }
        if (tid == 0 && *first_bad >= 0) atomicCAS(info + b, 0, k0 + *first_bad + 1);
        return;
    }
    // panel block: D <- A_ik, then L_ik = A_ik X^T
    const int i0 = (k + blk) * NB;
    for (int e = tid; e < NB * NB; e += NT) {
        const int r = e >> 6, m = e & 63;
        D[r * LS + m] = Wb[(size_t)(i0 + r) * npad + k0 + m];
    }
    __syncthreads();
    f4 acc[2][2];
    mfma_pqt(D, X, acc);
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = (w >> 1) * 32 + a * 16 + 4 * (lane >> 4) + e;
                const int cc = (w & 1) * 32 + c * 16 + (lane & 15);
                Wb[(size_t)(i0 + r) * npad + k0 + cc] = acc[a][c][e];
            }
}

// one Schur-complement tile of block column k: A_ij -= L_ik L_jk^T (rows i0, columns j0)
__device__ __forceinline__ void schur_tile(float* smem, float* Wb, int npad, int k, int i0, int j0) {
    float* Li = smem;
    float* Lj = smem + NB * LS;
    const int tid = threadIdx.x;
    const int k0 = k * NB;
    for (int e = tid; e < NB * NB; e += NT) {
        const int r = e >> 6, m = e & 63;
        Li[r * LS + m] = Wb[(size_t)(i0 + r) * npad + k0 + m];
        Lj[r * LS + m] = Wb[(size_t)(j0 + r) * npad + k0 + m];
    }
    __syncthreads();
    f4 acc[2][2];
    mfma_pqt(Li, Lj, acc);
    const int lane = tid & 63, w = tid >> 6;
#pragma unroll
    for (int a = 0; a < 2; ++a)
#pragma unroll
        for (int c = 0; c < 2; ++c)
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int r = (w >> 1) * 32 + a * 16 + 4 * (lane >> 4) + e;
                const int cc = (w & 1) * 32 + c * 16 + (lane & 15);
                Wb[(size_t)(i0 + r) * npad + j0 + cc] -= acc[a][c][e];
            }
}

// lower-triangle tile t -> (ib, jb), jb <= ib
__device__ __forceinline__ void tri_index(int t, int* ib, int* jb) {
    int i = (int)((sqrtf(8.0f * (float)t + 1.0f) - 1.0f) * 0.5f);
    while ((i + 1) * (i + 2) / 2 <= t) ++i;
    while (i * (i + 1) / 2 > t) --i;
    *ib = i;
    *jb = t - i * (i + 1) / 2;
}

// the first panel (block column 0)
__global__ __launch_bounds__(NT) void chol_panel(float* __restrict__ W, float* __restrict__ Xi, int npad, int T,
                                                 int* __restrict__ info, const unsigned* __restrict__ mask) {
    __shared__ float smem[kSmemFloats];
    __shared__ int first_bad;
    const int b = blockIdx.y;
    if (blockIdx.x > 0 && !mask_bit(mask + (size_t)b * T * kMaskWords, blockIdx.x, 0)) return;
    panel_body(smem, &first_bad, W + (size_t)b * npad * npad, Xi, b, npad, T, 0, blockIdx.x, info);
}

// update of block column k + 1 by block column k: A_i,k+1 -= L_ik L_k+1,k^T, i > k
__global__ __launch_bounds__(NT) void chol_schur_next(float* __restrict__ W, int npad, int k, int T,
                                                      const unsigned* __restrict__ mask) {
    __shared__ float smem[kSmemFloats];
    const int b = blockIdx.y;
    const unsigned* m = mask + (size_t)b * T * kMaskWords;
    const int ib = k + 1 + (int)blockIdx.x;
    if (!mask_bit(m, ib, k) || !mask_bit(m, k + 1, k)) return;
    schur_tile(smem, W + (size_t)b * npad * npad, npad, k, ib * NB, (k + 1) * NB);
}

// look-ahead: the panel of block column kp (T - kp workgroups) next to the rest of block
// column kp - 1's Schur update (tiles kp < j <= i), which touches no block of column kp
__global__ __launch_bounds__(NT) void chol_panel_schur(float* __restrict__ W, float* __restrict__ Xi, int npad, int T,
                                                       int kp, int* __restrict__ info,
                                                       const unsigned* __restrict__ mask) {
    __shared__ float smem[kSmemFloats];
    __shared__ int first_bad;
    const int b = blockIdx.y;
    float* Wb = W + (size_t)b * npad * npad;
    const unsigned* m = mask + (size_t)b * T * kMaskWords;
    const int npanel = T - kp;
    if ((int)blockIdx.x < npanel) {
        if (blockIdx.x > 0 && !mask_bit(m, kp + blockIdx.x, kp)) return;
        panel_body(smem, &first_bad, Wb, Xi, b, npad, T, kp, blockIdx.x, info);
        return;
    }
    int ib, jb;
    tri_index((int)blockIdx.x - npanel, &ib, &jb);
    ib += kp + 1;
    jb += kp + 1;
    if (!mask_bit(m, ib, kp - 1) || !mask_bit(m, jb, kp - 1)) return;
    schur_tile(smem, Wb, npad, kp - 1, ib * NB, jb * NB);
}

__global__ __launch_bounds__(NT) void chol_clear(int* __restrict__ flags, int count) {
    for (int e = blockIdx.x * NT + threadIdx.x; e < count; e += gridDim.x * NT) flags[e] = 0;
}

constexpr int kChunk = 16;  // published blocks a solve workgroup takes in one pass

// thread 0 waits for block j, then extends the run of published blocks (up to kChunk and
// below `limit` in the direction `dir`); every thread gets the run's end
__device__ __forceinline__ int ready_run(const int* fb, int j, int limit, int dir, int* timed_out, int* s_end) {
    if (threadIdx.x == 0) {
        int polls = 0;
        while (flag_load(fb + j) == 0) {
            if (++polls > kSpinLimit) {
                __hip_atomic_store(timed_out, 1, __ATOMIC_RELAXED, __HIP_MEMORY_SCOPE_AGENT);
                break;
            }
            __builtin_amdgcn_s_sleep(1);
        }
        int e = j + dir, cnt = 1;
        while (e != limit && cnt < kChunk && flag_load(fb + e) != 0) {
            e += dir;
            ++cnt;
        }
        *s_end = e;
    }
    __syncthreads();
    return *s_end;
}

// forward solve L y = r: workgroup i forms y_i = L_ii^-1 (r_i - sum_{j<i} L_ij y_j), taking
// the published y_j in runs; thread (r, q) owns row r and column quarter q of each block
__global__ __launch_bounds__(NT) void chol_fwd(const float* __restrict__ W, const float* __restrict__ Xi,
                                               const float* __restrict__ dsc, const double* __restrict__ rhs, int n,
                                               float* __restrict__ y, int* __restrict__ flags, int npad, int T,
                                               int* __restrict__ timed_out, const unsigned* __restrict__ mask) {
    __shared__ float Xs[NB * LS];
    __shared__ float part[4][NB];
    __shared__ float vbuf[kChunk * NB];
    __shared__ int s_end;
    const int b = blockIdx.y, i = blockIdx.x, tid = threadIdx.x;
    const int r = tid & 63, q = tid >> 6;
    const float* Wb = W + (size_t)b * npad * npad;
    const float* Xb = Xi + ((size_t)b * T + i) * NB * NB;
    float* yb = y + (size_t)b * npad;
    int* fb = flags + (size_t)b * T;
    for (int e = tid; e < NB * NB; e += NT) Xs[(e >> 6) * LS + (e & 63)] = Xb[e];
    const float* Lrow = Wb + (size_t)(i * NB + r) * npad + q * 16;
    float acc = 0.0f;
    // the L entries of the next block to wait for are loaded before the wait, so a block
    // published last costs one y load and the arithmetic
    // only the blocks L_ij of the block structure are waited for (the others are zero)
    const unsigned* mrow = mask + ((size_t)b * T + i) * kMaskWords;
    const int j0 = mask_next(mrow, 0, i);
    float4 pre[4];
    if (j0 < i) {
#pragma unroll
        for (int u = 0; u < 4; ++u) pre[u] = reinterpret_cast<const float4*>(Lrow + j0 * NB)[u];
    }
    for (int j = j0; j < i;) {
        const int je = ready_run(fb, j, i, 1, timed_out, &s_end);
        for (int e = tid; e < (je - j) * NB; e += NT) vbuf[e] = coherent_load(yb + j * NB + e);
        __syncthreads();
        for (int jj = j; jj < je; ++jj) {
            const float4* L4 = reinterpret_cast<const float4*>(Lrow + jj * NB);
            const float* v = vbuf + (jj - j) * NB + q * 16;
            float sacc = 0.0f;
#pragma unroll
            for (int u = 0; u < 4; ++u) {
                const float4 l = jj == j ? pre[u] : L4[u];
                sacc += l.x * v[4 * u] + l.y * v[4 * u + 1] + l.z * v[4 * u + 2] + l.w * v[4 * u + 3];
            }
            acc += sacc;
        }
        __syncthreads();
        j = mask_next(mrow, je, i);
        if (j < i) {
#pragma unroll
            for (int u = 0; u < 4; ++u) pre[u] = reinterpret_cast<const float4*>(Lrow + j * NB)[u];
        }
    }
    part[q][r] = acc;
    __syncthreads();
    if (tid < NB) {
        const int row = i * NB + tid;
        const float ri = row < n ? (float)(rhs[(size_t)b * n + row] * (double)dsc[(size_t)b * npad + row]) : 0.0f;
        vbuf[tid] = ri - (part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]);
    }
    __syncthreads();
    float sacc = 0.0f;
#pragma unroll
    for (int c = 0; c < 16; ++c) sacc += Xs[r * LS + q * 16 + c] * vbuf[q * 16 + c];
    part[q][r] = sacc;
    __syncthreads();
    if (tid < NB) yb[i * NB + tid] = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
    publish(fb + i);
}

// backward solve L^T x = y: workgroup w forms block i = T - 1 - w,
// x_i = L_ii^-T (y_i - sum_{j>i} L_ji^T x_j), taking the published x_j in runs; thread (c, q)
// owns column c and row quarter q of each block; x (float64, scaled back) is written, or added
// when accumulating, for rows < n
__global__ __launch_bounds__(NT) void chol_bwd(const float* __restrict__ W, const float* __restrict__ Xi,
                                               const float* __restrict__ dsc, const float* __restrict__ y,
                                               float* __restrict__ xs, double* __restrict__ x, int n, int accumulate,
                                               int* __restrict__ flags, int npad, int T, int* __restrict__ timed_out,
                                               const unsigned* __restrict__ mask) {
    __shared__ float part[4][NB];
    __shared__ float vbuf[kChunk * NB];
    __shared__ int s_end;
    const int b = blockIdx.y, i = T - 1 - (int)blockIdx.x, tid = threadIdx.x;
    const int c = tid & 63, q = tid >> 6;
    const float* Wb = W + (size_t)b * npad * npad;
    const float* Xb = Xi + ((size_t)b * T + i) * NB * NB;
    float* xb = xs + (size_t)b * npad;
    int* fb = flags + (size_t)b * T;
    float acc = 0.0f;
    // the L entries of the next block to wait for are loaded before the wait
    // only the blocks L_ji of the block structure are waited for (the others are zero)
    const unsigned* mb = mask + (size_t)b * T * kMaskWords;
    auto prev = [&](int from) {
        for (int jj = from; jj > i; --jj)
            if (mask_bit(mb, jj, i)) return jj;
        return i;
    };
    const int j0 = prev(T - 1);
    float pre[16];
    if (j0 > i) {
        const float* Lc = Wb + (size_t)(j0 * NB + q * 16) * npad + i * NB + c;
#pragma unroll
        for (int rr = 0; rr < 16; ++rr) pre[rr] = Lc[(size_t)rr * npad];
    }
    for (int j = j0; j > i;) {
        const int je = ready_run(fb, j, i, -1, timed_out, &s_end);  // blocks j, j-1, ..., je+1
        const int cnt = j - je;
        for (int e = tid; e < cnt * NB; e += NT) {
            const int blk = e >> 6;  // 0 -> block j
            vbuf[e] = coherent_load(xb + (j - blk) * NB + (e & 63));
        }
        __syncthreads();
        {
            const float* v = vbuf + q * 16;
            float sacc = 0.0f;
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) sacc += pre[rr] * v[rr];
            acc += sacc;
        }
        for (int t = 1; t < cnt; ++t) {
            const float* Lc = Wb + (size_t)((j - t) * NB + q * 16) * npad + i * NB + c;
            const float* v = vbuf + t * NB + q * 16;
            float sacc = 0.0f;
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) sacc += Lc[(size_t)rr * npad] * v[rr];
            acc += sacc;
        }
        __syncthreads();
        j = prev(je);
        if (j > i) {
            const float* Lc = Wb + (size_t)(j * NB + q * 16) * npad + i * NB + c;
#pragma unroll
            for (int rr = 0; rr < 16; ++rr) pre[rr] = Lc[(size_t)rr * npad];
        }
    }
    part[q][c] = acc;
    __syncthreads();
    if (tid < NB) vbuf[tid] = y[(size_t)b * npad + i * NB + tid] - (part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid]);
    __syncthreads();
    float sacc = 0.0f;
#pragma unroll
    for (int rr = 0; rr < 16; ++rr) sacc += Xb[(q * 16 + rr) * NB + c] * vbuf[q * 16 + rr];
    part[q][c] = sacc;
    __syncthreads();
    if (tid < NB) {
        const float xv = part[0][tid] + part[1][tid] + part[2][tid] + part[3][tid];
        xb[i * NB + tid] = xv;
        const int row = i * NB + tid;
        if (row < n) {
            double* o = x + (size_t)b * n + row;
            const double xr = (double)xv * (double)dsc[(size_t)b * npad + row];
            *o = accumulate ? *o + xr : xr;
        }
    }
    publish(fb + i);
}

// r = rhs - A0 x in float64, one wave per row
__global__ __launch_bounds__(NT) void chol_residual(const float* __restrict__ A0, const double* __restrict__ rhs,
                                                    const double* __restrict__ x, double* __restrict__ r, int n,
                                                    int npad) {
    const int b = blockIdx.y;
    const int row = blockIdx.x * (NT / 64) + (threadIdx.x >> 6);
    const int lane = threadIdx.x & 63;
    if (row >= n) return;
    const float* a = A0 + (size_t)b * npad * npad + (size_t)row * npad;
    const double* xb = x + (size_t)b * n;
    double s = 0.0;
    for (int c = lane; c < n; c += 64) s += (double)a[c] * xb[c];
#pragma unroll
    for (int off = 32; off > 0; off >>= 1) s += __shfl_xor(s, off, 64);
    if (lane == 0) r[(size_t)b * n + row] = rhs[(size_t)b * n + row] - s;
}

}  // namespace

struct SapxChol {
    int max_n = 0, npad_max = 0, batch = 0;
    int n = 0, npad = 0, T = 0;
    float* W = nullptr;
    float* A0 = nullptr;
    float* Xi = nullptr;
    float* y = nullptr;
    float* xs = nullptr;
    float* dsc = nullptr;
    double* r = nullptr;
    int* flags = nullptr;  // 2 x batch x T: forward, backward
    unsigned* mask = nullptr;  // batch x T x kMaskWords: block structure of the factor
    int* info = nullptr;
    int* timed_out = nullptr;
};

extern "C" {

int32_t sapx_chol_create(int32_t max_num_equations, int32_t batch_size, SapxChol** out) {
    if (!out) return cfail(SAPX_ERR_INVALID, "sapx_chol_create: null output");
    *out = nullptr;
    if (max_num_equations <= 0 || batch_size <= 0)
        return cfail(SAPX_ERR_INVALID, "sapx_chol_create: max_num_equations and batch_size must be positive");
    const long long npad = ((long long)max_num_equations + NB - 1) / NB * NB;
    const long long T = npad / NB;
    if (T * batch_size > kMaxResident)
        return cfail(SAPX_ERR_CAPACITY, "sapx_chol_create: (padded order / 64) x batch_size must be at most " +
                                            std::to_string(kMaxResident));
    SapxChol* s = new SapxChol();
    s->max_n = max_num_equations;
    s->npad_max = (int)npad;
    s->batch = batch_size;
    const size_t mat = (size_t)batch_size * npad * npad;
    bool ok = hipMalloc(&s->W, mat * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->A0, mat * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->Xi, (size_t)batch_size * T * NB * NB * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->y, (size_t)batch_size * npad * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->xs, (size_t)batch_size * npad * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->dsc, (size_t)batch_size * npad * sizeof(float)) == hipSuccess &&
              hipMalloc(&s->r, (size_t)batch_size * npad * sizeof(double)) == hipSuccess &&
              hipMalloc(&s->flags, (size_t)2 * batch_size * T * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->mask, (size_t)batch_size * T * kMaskWords * sizeof(unsigned)) == hipSuccess &&
              hipMalloc(&s->info, (size_t)batch_size * sizeof(int)) == hipSuccess &&
              hipMalloc(&s->timed_out, sizeof(int)) == hipSuccess &&
              hipMemset(s->info, 0, (size_t)batch_size * sizeof(int)) == hipSuccess &&
              hipMemset(s->timed_out, 0, sizeof(int)) == hipSuccess && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        sapx_chol_destroy(s);
        return cfail(SAPX_ERR_HIP, "sapx_chol_create: device allocation failed");
    }
    *out = s;
    return SAPX_OK;
}

void sapx_chol_destroy(SapxChol* s) {
    if (!s) return;
    (void)hipFree(s->W);
    (void)hipFree(s->A0);
    (void)hipFree(s->Xi);
    (void)hipFree(s->y);
    (void)hipFree(s->xs);
    (void)hipFree(s->dsc);
    (void)hipFree(s->r);
    (void)hipFree(s->flags);
    (void)hipFree(s->mask);
    (void)hipFree(s->info);
    (void)hipFree(s->timed_out);
    delete s;
}

int32_t sapx_chol_factorize(SapxChol* s, const float* A, int32_t n, int32_t lda, int64_t a_stride, void* stream) {
    if (!s || !A) return cfail(SAPX_ERR_INVALID, "sapx_chol_factorize: null argument");
    if (n <= 0 || n > s->max_n)
        return cfail(SAPX_ERR_INVALID, "sapx_chol_factorize: n must be in 1.." + std::to_string(s->max_n));
    if (lda < n || a_stride < (int64_t)lda * n) return cfail(SAPX_ERR_INVALID, "sapx_chol_factorize: lda < n or stride too small");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    s->n = n;
    s->npad = (n + NB - 1) / NB * NB;
    s->T = s->npad / NB;
    const int npad = s->npad, T = s->T, B = s->batch;
    const long long total = (long long)npad * npad;
    const int lgrid = (int)std::min<long long>((total + NT - 1) / NT, 4096);
    hipLaunchKernelGGL(chol_scale, dim3(T, B), dim3(NT), 0, st, A, lda, (long long)a_stride, n, s->dsc, npad);
    hipLaunchKernelGGL(chol_load, dim3(lgrid, B), dim3(NT), 0, st, A, lda, (long long)a_stride, n, s->dsc, s->W, s->A0,
                       npad, s->info);
    // block structure of the factor (cleared by a kernel: the sequence is graph-capturable)
    const int mwords = B * T * kMaskWords;
    hipLaunchKernelGGL(chol_clear, dim3((mwords + NT - 1) / NT), dim3(NT), 0, st, reinterpret_cast<int*>(s->mask),
                       mwords);
    hipLaunchKernelGGL(chol_mask, dim3(T * (T + 1) / 2, B), dim3(NT), 0, st, s->W, npad, T, s->mask);
    hipLaunchKernelGGL(chol_symbolic, dim3(B), dim3(1024), 0, st, T, s->mask);
    // panel 0; then per block column k: update column k + 1, then factor its panel while the
    // rest of column k's update runs in the same launch
    hipLaunchKernelGGL(chol_panel, dim3(T, B), dim3(NT), 0, st, s->W, s->Xi, npad, T, s->info, s->mask);
    for (int k = 0; k + 1 < T; ++k) {
        hipLaunchKernelGGL(chol_schur_next, dim3(T - k - 1, B), dim3(NT), 0, st, s->W, npad, k, T, s->mask);
        const int m2 = T - k - 2;
        hipLaunchKernelGGL(chol_panel_schur, dim3((T - k - 1) + m2 * (m2 + 1) / 2, B), dim3(NT), 0, st, s->W, s->Xi,
                           npad, T, k + 1, s->info, s->mask);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfail(SAPX_ERR_HIP, std::string("sapx_chol_factorize: ") + hipGetErrorString(e));
    return SAPX_OK;
}

int32_t sapx_chol_solve(SapxChol* s, const double* rhs, double* result, int32_t refine, void* stream) {
    if (!s || !rhs || !result) return cfail(SAPX_ERR_INVALID, "sapx_chol_solve: null argument");
    if (s->n <= 0) return cfail(SAPX_ERR_INVALID, "sapx_chol_solve: factorize first");
    if (refine < 0 || refine > 8) return cfail(SAPX_ERR_INVALID, "sapx_chol_solve: refine must be 0..8");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int n = s->n, npad = s->npad, T = s->T, B = s->batch;
    int* ff = s->flags;
    int* fbk = s->flags + (size_t)B * T;
    const int nflags = 2 * B * T;
    for (int pass = 0; pass <= refine; ++pass) {
        const double* b = rhs;
        if (pass > 0) {
            hipLaunchKernelGGL(chol_residual, dim3((n + 3) / 4, B), dim3(NT), 0, st, s->A0, rhs, result, s->r, n, npad);
            b = s->r;
        }
        hipLaunchKernelGGL(chol_clear, dim3((nflags + NT - 1) / NT), dim3(NT), 0, st, s->flags, nflags);
        hipLaunchKernelGGL(chol_fwd, dim3(T, B), dim3(NT), 0, st, s->W, s->Xi, s->dsc, b, n, s->y, ff, npad, T,
                           s->timed_out, s->mask);
        hipLaunchKernelGGL(chol_bwd, dim3(T, B), dim3(NT), 0, st, s->W, s->Xi, s->dsc, s->y, s->xs, result, n,
                           pass > 0 ? 1 : 0, fbk, npad, T, s->timed_out, s->mask);
    }
    const hipError_t e = hipGetLastError();
    if (e != hipSuccess) return cfail(SAPX_ERR_HIP, std::string("sapx_chol_solve: ") + hipGetErrorString(e));
    return SAPX_OK;
}

int32_t sapx_chol_info(SapxChol* s, int32_t* info) {
    if (!s || !info) return cfail(SAPX_ERR_INVALID, "sapx_chol_info: null argument");
    int to = 0;
    if (hipDeviceSynchronize() != hipSuccess ||
        hipMemcpy(info, s->info, (size_t)s->batch * sizeof(int), hipMemcpyDeviceToHost) != hipSuccess ||
        hipMemcpy(&to, s->timed_out, sizeof(int), hipMemcpyDeviceToHost) != hipSuccess)
        return cfail(SAPX_ERR_HIP, "sapx_chol_info: reading the status failed");
    if (to) return cfail(SAPX_ERR_HIP, "sapx_chol_info: a solve's wait for a block timed out");
    return SAPX_OK;
}

}  // extern "C"

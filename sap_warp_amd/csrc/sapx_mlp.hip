// Learned warm start: a small ReLU MLP trained and evaluated on the GPU (gfx950).
//
// No counterpart in the reference: this is the network that predicts a correction
// to SolverSAP's plain initial velocity.  Everything runs through one GEMM form,
// C[M x N] = A[M x K] . B[N x K]^T with both operands bfloat16 and contiguous
// along K, on v_mfma_f32_16x16x16_bf16 with float32 accumulation.  One wave owns a
// 64 x 64 tile of C (4 x 4 MFMA tiles) and reads its operand fragments straight
// from global memory (8 bytes per lane, K-contiguous), so no LDS staging is
// needed.  The three products of training use it as
//   forward      h_{l+1} = relu(h_l W_l^T + b_l)        A = h_l,        B = W_l
//   input grad   dz_l    = (dz_{l+1} W_l) * [h_l > 0]    A = dz_{l+1},   B = W_l^T
//   weight grad  dW_l    = dz_{l+1}^T h_l                A = dz_{l+1}^T, B = h_l^T
// and the epilogues write each activation / gradient both row-major and
// transposed, so every operand is already K-contiguous when it is read.
// bfloat16 rounding points: normalised inputs, weights, hidden activations and
// back-propagated gradients; biases, the output and all accumulations are float32.
#include <hip/hip_runtime.h>

#include <cmath>
#include <cstdint>
#include <cstring>
#include <string>
#include <vector>

#include "../../include/sapx.h"

namespace sapx {
namespace mlp {

typedef short s4 __attribute__((ext_vector_type(4)));
typedef float f4 __attribute__((ext_vector_type(4)));

constexpr int kMaxLayers = 8;

__device__ __forceinline__ uint16_t f2bf(float f) {
    uint32_t u = __float_as_uint(f);
    if ((u & 0x7fffffffu) > 0x7f800000u) return 0x7fc0;  // NaN
    u += 0x7fffu + ((u >> 16) & 1u);                     // round to nearest, ties to even
    return (uint16_t)(u >> 16);
}
__device__ __forceinline__ float bf2f(uint16_t h) { return __uint_as_float((uint32_t)h << 16); }

enum Epilogue : int {
    EPI_F32 = 0,         // Cf = acc (+ bias)
    EPI_RELU_BF16 = 1,   // Cb = bf16(relu(acc + bias)), optional transposed copy Ct
    EPI_MASK_BF16 = 2,   // Cb = bf16(acc * [H > 0]), transposed copy Ct, column sums of acc * [H > 0] into colsum
};

struct GemmArgs {
    const uint16_t* A;
    const uint16_t* B;
    int lda, ldb;
    int K;
    const float* bias;
    float* Cf;
    uint16_t* Cb;
    uint16_t* Ct;
    int ldc, ldt;
    const uint16_t* H;
    int ldh;
    float* colsum;
};

// one wave: a 64 x 64 tile of C at (blockIdx.y, blockIdx.x); M, N multiples of 64,
// K a multiple of 16, operand rows 8-byte aligned
template <int EPI>
__global__ __launch_bounds__(64) void gemm_nt_kernel(GemmArgs g) {
    const int lane = threadIdx.x;
    const int r = lane & 15, kq = lane >> 4;
    const int m0 = blockIdx.y * 64, n0 = blockIdx.x * 64;
    f4 acc[4][4];
#pragma unroll
    for (int i = 0; i < 4; ++i)
#pragma unroll
        for (int j = 0; j < 4; ++j) acc[i][j] = f4{0.0f, 0.0f, 0.0f, 0.0f};
    const uint16_t* Ap = g.A + (size_t)(m0 + r) * g.lda + 4 * kq;
    const uint16_t* Bp = g.B + (size_t)(n0 + r) * g.ldb + 4 * kq;
    const size_t as = (size_t)16 * g.lda, bs = (size_t)16 * g.ldb;
    for (int k0 = 0; k0 < g.K; k0 += 16) {
        s4 a[4], b[4];
#pragma unroll
        for (int i = 0; i < 4; ++i) a[i] = *reinterpret_cast<const s4*>(Ap + i * as + k0);
#pragma unroll
        for (int j = 0; j < 4; ++j) b[j] = *reinterpret_cast<const s4*>(Bp + j * bs + k0);
#pragma unroll
        for (int i = 0; i < 4; ++i)
#pragma unroll
            for (int j = 0; j < 4; ++j) acc[i][j] = __builtin_amdgcn_mfma_f32_16x16x16bf16_1k(a[i], b[j], acc[i][j], 0, 0, 0);
    }
    // D layout: acc[i][j][e] = C[m0 + 16 i + 4 kq + e][n0 + 16 j + r]
#pragma unroll
    for (int j = 0; j < 4; ++j) {
        const int n = n0 + 16 * j + r;
        const float bn = g.bias ? g.bias[n] : 0.0f;
        float csum = 0.0f;
#pragma unroll
        for (int i = 0; i < 4; ++i) {
            const int mb = m0 + 16 * i + 4 * kq;
            uint16_t tb[4];
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int m = mb + e;
                float v = acc[i][j][e];
                if (EPI == EPI_F32) {
                    g.Cf[(size_t)m * g.ldc + n] = v + bn;
                } else if (EPI == EPI_RELU_BF16) {
                    v = fmaxf(v + bn, 0.0f);
                    const uint16_t h = f2bf(v);
                    g.Cb[(size_t)m * g.ldc + n] = h;
                    tb[e] = h;
                } else {
                    v = bf2f(g.H[(size_t)m * g.ldh + n]) > 0.0f ? v : 0.0f;
                    const uint16_t h = f2bf(v);
                    g.Cb[(size_t)m * g.ldc + n] = h;
                    tb[e] = h;
                    csum += v;
                }
            }
            if (EPI != EPI_F32 && g.Ct) {
                // four consecutive rows of column n: one 8-byte store into the transposed copy
                const uint64_t packed = (uint64_t)tb[0] | ((uint64_t)tb[1] << 16) | ((uint64_t)tb[2] << 32) |
                                        ((uint64_t)tb[3] << 48);
                *reinterpret_cast<uint64_t*>(g.Ct + (size_t)n * g.ldt + mb) = packed;
            }
        }
        if (EPI == EPI_MASK_BF16) {
            // lanes r, r + 16, r + 32, r + 48 hold the same column
            csum += __shfl_xor(csum, 16, 64);
            csum += __shfl_xor(csum, 32, 64);
            if (kq == 0) atomicAdd(g.colsum + n, csum);
        }
    }
}

// x_raw [B x F] float32 -> normalised bfloat16 rows [Bp x d0] and the transposed copy
// [d0 x Bp]; rows >= B and columns >= F are zero
__global__ void normalize_input_kernel(const float* x, int B, int F, int Bp, int d0, const float* mean, const float* stdv,
                                       uint16_t* xb, uint16_t* xt) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= (size_t)Bp * d0) return;
    const int b = (int)(t / d0), f = (int)(t % d0);
    float v = 0.0f;
    if (b < B && f < F) v = (x[(size_t)b * F + f] - mean[f]) / stdv[f];
    const uint16_t h = f2bf(v);
    xb[t] = h;
    xt[(size_t)f * Bp + b] = h;
}

// normalised output y [Bp x dL] -> y_raw [B x O] = y * std + mean
__global__ void unnormalize_output_kernel(const float* y, int B, int O, int dL, const float* mean, const float* stdv,
                                          float* out) {
    const int t = blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= B * O) return;
    const int b = t / O, o = t % O;
    out[t] = y[(size_t)b * dL + o] * stdv[o] + mean[o];
}

// mean squared error over the real outputs in normalised units, its gradient dz
// (bfloat16, row-major and transposed) and the output bias gradient (column sums);
// blocks of 256 threads = 4 row lanes x 64 columns, 64 rows per block
__global__ __launch_bounds__(256) void loss_grad_kernel(const float* y, const float* target, int B, int O, int Bp,
                                                        int dL, const float* mean, const float* stdv, uint16_t* dz,
                                                        uint16_t* dzt, float* dbias, float* loss) {
    __shared__ float part[4][64];
    __shared__ float lpart[256];
    const int c = threadIdx.x & 63, rl = threadIdx.x >> 6;
    const float scale = 2.0f / ((float)B * (float)O);
    float cs = 0.0f, ls = 0.0f;
    for (int cc = c; cc < dL; cc += 64) {
        float csc = 0.0f;
        for (int k = 0; k < 16; ++k) {
            const int b = blockIdx.x * 64 + rl * 16 + k;
            if (b >= Bp) break;
            float d = 0.0f;
            if (b < B && cc < O) {
                const float t = (target[(size_t)b * O + cc] - mean[cc]) / stdv[cc];
                d = y[(size_t)b * dL + cc] - t;
                ls += d * d;
            }
            const float gz = scale * d;
            const uint16_t h = f2bf(gz);
            dz[(size_t)b * dL + cc] = h;
            dzt[(size_t)cc * Bp + b] = h;
            csc += gz;
        }
        if (cc == c) cs = csc;
        else atomicAdd(dbias + cc, csc);  // only when dL > 64
    }
    part[rl][c] = cs;
    lpart[threadIdx.x] = ls;
    __syncthreads();
    if (rl == 0 && c < dL) atomicAdd(dbias + c, part[0][c] + part[1][c] + part[2][c] + part[3][c]);
    for (int s = 128; s > 0; s >>= 1) {
        if ((int)threadIdx.x < s) lpart[threadIdx.x] += lpart[threadIdx.x + s];
        __syncthreads();
    }
    if (threadIdx.x == 0) atomicAdd(loss, lpart[0] / ((float)B * (float)O));
}

__global__ void fill_kernel(float* p, size_t n, float v) {
    const size_t t = (size_t)blockIdx.x * blockDim.x + threadIdx.x;
    if (t < n) p[t] = v;
}

struct LayerOffsets {
    int L;
    int dims[kMaxLayers + 1];
    long long w_off[kMaxLayers];   // flat float offset of W_l (b_l follows it)
    long long wb_off[kMaxLayers];  // offset of W_l inside the bfloat16 weight arrays
    long long total;
};

// Adam over the flat parameters (bias-corrected moments), then the bfloat16 weight
// copies W_l and W_l^T are refreshed from the float32 master values
__global__ void adam_kernel(float* p, const float* grad, float* m, float* v, LayerOffsets lo, float lr, float b1,
                            float b2, float eps, float c1, float c2, uint16_t* wb, uint16_t* wt, int update) {
    const long long t = (long long)blockIdx.x * blockDim.x + threadIdx.x;
    if (t >= lo.total) return;
    float x = p[t];
    if (update) {
        const float g = grad[t];
        const float mm = b1 * m[t] + (1.0f - b1) * g;
        const float vv = b2 * v[t] + (1.0f - b2) * g * g;
        m[t] = mm;
        v[t] = vv;
        x -= lr * (mm / c1) / (sqrtf(vv / c2) + eps);
        p[t] = x;
    }
    for (int l = 0; l < lo.L; ++l) {
        const long long nin = lo.dims[l], nout = lo.dims[l + 1];
        const long long rel = t - lo.w_off[l];
        if (rel >= 0 && rel < nin * nout) {
            const long long o = rel / nin, i = rel % nin;
            const uint16_t h = f2bf(x);
            wb[lo.wb_off[l] + o * nin + i] = h;
            wt[lo.wb_off[l] + i * nout + o] = h;
        }
    }
}

}  // namespace mlp
}  // namespace sapx

using namespace sapx::mlp;

namespace sapx {
void set_last_error(const char* msg);
}

namespace {
int32_t mfail(int32_t code, const std::string& msg) {
    sapx::set_last_error(msg.c_str());
    return code;
}
}  // namespace

#define MLP_HIP(expr)                                                                        \
    do {                                                                                     \
        hipError_t _e = (expr);                                                              \
        if (_e != hipSuccess) return mfail(SAPX_ERR_HIP, std::string(#expr) + ": " + hipGetErrorString(_e)); \
    } while (0)

struct SapxMlp {
    LayerOffsets lo{};
    int in_features = 0, out_features = 0;
    long long wb_total = 0;
    float* params = nullptr;
    float* grads = nullptr;
    float* m = nullptr;
    float* v = nullptr;
    uint16_t* wb = nullptr;
    uint16_t* wt = nullptr;
    float* norm = nullptr;  // x_mean [F], x_std [F], y_mean [O], y_std [O]
    float* loss = nullptr;
    int adam_t = 0;
    // workspace for batches up to cap rows
    int cap = 0;
    uint16_t* xb = nullptr;
    uint16_t* xt = nullptr;
    uint16_t* h[kMaxLayers] = {nullptr};
    uint16_t* ht[kMaxLayers] = {nullptr};
    float* y = nullptr;
    uint16_t* dz[2] = {nullptr, nullptr};
    uint16_t* dzt[2] = {nullptr, nullptr};
    void free_ws() {
        void* ps[] = {xb, xt, y, dz[0], dz[1], dzt[0], dzt[1]};
        for (void* p : ps)
            if (p) (void)hipFree(p);
        for (int l = 0; l < kMaxLayers; ++l) {
            if (h[l]) (void)hipFree(h[l]);
            if (ht[l]) (void)hipFree(ht[l]);
            h[l] = ht[l] = nullptr;
        }
        xb = xt = nullptr;
        y = nullptr;
        dz[0] = dz[1] = dzt[0] = dzt[1] = nullptr;
        cap = 0;
    }
    ~SapxMlp() {
        free_ws();
        void* ps[] = {params, grads, m, v, wb, wt, norm, loss};
        for (void* p : ps)
            if (p) (void)hipFree(p);
    }
    int maxw() const {
        int w = 0;
        for (int l = 0; l <= lo.L; ++l) w = lo.dims[l] > w ? lo.dims[l] : w;
        return w;
    }
};

static int32_t reserve(SapxMlp* M, int batch) {
    const int Bp = (batch + 63) / 64 * 64;
    if (Bp <= M->cap) return SAPX_OK;
    M->free_ws();
    const int L = M->lo.L;
    const size_t mw = (size_t)M->maxw();
    MLP_HIP(hipMalloc(&M->xb, (size_t)Bp * M->lo.dims[0] * 2));
    MLP_HIP(hipMalloc(&M->xt, (size_t)Bp * M->lo.dims[0] * 2));
    for (int l = 1; l < L; ++l) {
        MLP_HIP(hipMalloc(&M->h[l], (size_t)Bp * M->lo.dims[l] * 2));
        MLP_HIP(hipMalloc(&M->ht[l], (size_t)Bp * M->lo.dims[l] * 2));
    }
    MLP_HIP(hipMalloc(&M->y, (size_t)Bp * M->lo.dims[L] * 4));
    for (int k = 0; k < 2; ++k) {
        MLP_HIP(hipMalloc(&M->dz[k], (size_t)Bp * mw * 2));
        MLP_HIP(hipMalloc(&M->dzt[k], (size_t)Bp * mw * 2));
    }
    M->cap = Bp;
    return SAPX_OK;
}

static hipError_t gemm(int epi, const GemmArgs& g, int M, int N, hipStream_t st) {
    const dim3 grid(N / 64, M / 64), block(64);
    switch (epi) {
        case EPI_F32: hipLaunchKernelGGL(gemm_nt_kernel<EPI_F32>, grid, block, 0, st, g); break;
        case EPI_RELU_BF16: hipLaunchKernelGGL(gemm_nt_kernel<EPI_RELU_BF16>, grid, block, 0, st, g); break;
        default: hipLaunchKernelGGL(gemm_nt_kernel<EPI_MASK_BF16>, grid, block, 0, st, g); break;
    }
    return hipGetLastError();
}

static hipError_t refresh_bf16(SapxMlp* M, hipStream_t st, int update, float lr, float b1, float b2, float eps) {
    float c1 = 1.0f, c2 = 1.0f;
    if (update) {
        c1 = 1.0f - powf(b1, (float)M->adam_t);
        c2 = 1.0f - powf(b2, (float)M->adam_t);
    }
    const int threads = 256;
    const unsigned blocks = (unsigned)((M->lo.total + threads - 1) / threads);
    hipLaunchKernelGGL(adam_kernel, dim3(blocks), dim3(threads), 0, st, M->params, M->grads, M->m, M->v, M->lo, lr, b1,
                       b2, eps, c1, c2, M->wb, M->wt, update);
    return hipGetLastError();
}

// forward pass of `batch` rows already normalised into M->xb / M->xt; output (normalised) in M->y
static int32_t forward_normalised(SapxMlp* M, int Bp, hipStream_t st) {
    const int L = M->lo.L;
    for (int l = 0; l < L; ++l) {
        GemmArgs g{};
        g.A = l == 0 ? M->xb : M->h[l];
        g.lda = M->lo.dims[l];
        g.B = M->wb + M->lo.wb_off[l];
        g.ldb = M->lo.dims[l];
        g.K = M->lo.dims[l];
        g.bias = M->params + M->lo.w_off[l] + (long long)M->lo.dims[l] * M->lo.dims[l + 1];
        if (l < L - 1) {
            g.Cb = M->h[l + 1];
            g.Ct = M->ht[l + 1];
            g.ldc = M->lo.dims[l + 1];
            g.ldt = Bp;
            MLP_HIP(gemm(EPI_RELU_BF16, g, Bp, M->lo.dims[l + 1], st));
        } else {
            g.Cf = M->y;
            g.ldc = M->lo.dims[L];
            MLP_HIP(gemm(EPI_F32, g, Bp, M->lo.dims[L], st));
        }
    }
    return SAPX_OK;
}

static int32_t normalise(SapxMlp* M, const float* x, int batch, int Bp, hipStream_t st) {
    const int d0 = M->lo.dims[0];
    const size_t tot = (size_t)Bp * d0;
    hipLaunchKernelGGL(normalize_input_kernel, dim3((unsigned)((tot + 255) / 256)), dim3(256), 0, st, x, batch,
                       M->in_features, Bp, d0, M->norm, M->norm + M->in_features, M->xb, M->xt);
    MLP_HIP(hipGetLastError());
    return SAPX_OK;
}

extern "C" {

int32_t sapx_mlp_create(int32_t layers, const int32_t* dims, int32_t in_features, int32_t out_features,
                        SapxMlp** out) {
    if (!out || !dims || layers < 1 || layers > kMaxLayers) return mfail(SAPX_ERR_INVALID, "sapx_mlp_create: bad layer count");
    for (int l = 0; l <= layers; ++l)
        if (dims[l] <= 0 || dims[l] % 64 != 0) return mfail(SAPX_ERR_INVALID, "sapx_mlp_create: widths must be positive multiples of 64");
    if (in_features <= 0 || in_features > dims[0] || out_features <= 0 || out_features > dims[layers])
        return mfail(SAPX_ERR_INVALID, "sapx_mlp_create: real widths exceed the padded widths");
    SapxMlp* M = new SapxMlp();
    M->lo.L = layers;
    long long off = 0, woff = 0;
    for (int l = 0; l <= layers; ++l) M->lo.dims[l] = dims[l];
    for (int l = 0; l < layers; ++l) {
        M->lo.w_off[l] = off;
        M->lo.wb_off[l] = woff;
        off += (long long)dims[l] * dims[l + 1] + dims[l + 1];
        woff += (long long)dims[l] * dims[l + 1];
    }
    M->lo.total = off;
    M->wb_total = woff;
    M->in_features = in_features;
    M->out_features = out_features;
    const size_t pb = (size_t)off * 4;
    if (hipMalloc(&M->params, pb) != hipSuccess || hipMalloc(&M->grads, pb) != hipSuccess ||
        hipMalloc(&M->m, pb) != hipSuccess || hipMalloc(&M->v, pb) != hipSuccess ||
        hipMalloc(&M->wb, (size_t)woff * 2) != hipSuccess || hipMalloc(&M->wt, (size_t)woff * 2) != hipSuccess ||
        hipMalloc(&M->norm, (size_t)(2 * in_features + 2 * out_features) * 4) != hipSuccess ||
        hipMalloc(&M->loss, 4) != hipSuccess) {
        delete M;
        return mfail(SAPX_ERR_HIP, "sapx_mlp_create: hipMalloc failed");
    }
    (void)hipMemset(M->params, 0, pb);
    (void)hipMemset(M->grads, 0, pb);
    (void)hipMemset(M->m, 0, pb);
    (void)hipMemset(M->v, 0, pb);
    (void)hipMemset(M->wb, 0, (size_t)woff * 2);
    (void)hipMemset(M->wt, 0, (size_t)woff * 2);
    std::vector<float> nrm(2 * in_features + 2 * out_features, 0.0f);
    for (int f = 0; f < in_features; ++f) nrm[in_features + f] = 1.0f;
    for (int o = 0; o < out_features; ++o) nrm[2 * in_features + out_features + o] = 1.0f;
    (void)hipMemcpy(M->norm, nrm.data(), nrm.size() * 4, hipMemcpyHostToDevice);
    *out = M;
    return SAPX_OK;
}

void sapx_mlp_destroy(SapxMlp* mlp) { delete mlp; }

int64_t sapx_mlp_param_count(const SapxMlp* mlp) { return mlp ? (int64_t)mlp->lo.total : 0; }

int32_t sapx_mlp_write(SapxMlp* M, int32_t which, const float* host, int64_t count) {
    if (!M || !host) return mfail(SAPX_ERR_INVALID, "sapx_mlp_write: null argument");
    const int F = M->in_features, O = M->out_features;
    if (which == SAPX_MLP_PARAMS) {
        if (count != M->lo.total) return mfail(SAPX_ERR_INVALID, "sapx_mlp_write: parameter count mismatch");
        MLP_HIP(hipMemcpy(M->params, host, (size_t)count * 4, hipMemcpyHostToDevice));
        MLP_HIP(hipMemset(M->m, 0, (size_t)count * 4));
        MLP_HIP(hipMemset(M->v, 0, (size_t)count * 4));
        M->adam_t = 0;
        MLP_HIP(refresh_bf16(M, nullptr, 0, 0.0f, 0.0f, 0.0f, 0.0f));
        MLP_HIP(hipDeviceSynchronize());
        return SAPX_OK;
    }
    const int off[4] = {0, F, 2 * F, 2 * F + O};
    const int len[4] = {F, F, O, O};
    if (which < SAPX_MLP_X_MEAN || which > SAPX_MLP_Y_STD) return mfail(SAPX_ERR_INVALID, "sapx_mlp_write: bad buffer id");
    const int k = which - SAPX_MLP_X_MEAN;
    if (count != len[k]) return mfail(SAPX_ERR_INVALID, "sapx_mlp_write: normalisation length mismatch");
    MLP_HIP(hipMemcpy(M->norm + off[k], host, (size_t)count * 4, hipMemcpyHostToDevice));
    return SAPX_OK;
}

int32_t sapx_mlp_read(const SapxMlp* M, int32_t which, float* host, int64_t count) {
    if (!M || !host) return mfail(SAPX_ERR_INVALID, "sapx_mlp_read: null argument");
    const float* src = which == SAPX_MLP_PARAMS ? M->params : which == SAPX_MLP_GRADS ? M->grads : nullptr;
    if (!src) return mfail(SAPX_ERR_INVALID, "sapx_mlp_read: bad buffer id");
    if (count != M->lo.total) return mfail(SAPX_ERR_INVALID, "sapx_mlp_read: parameter count mismatch");
    MLP_HIP(hipDeviceSynchronize());
    MLP_HIP(hipMemcpy(host, src, (size_t)count * 4, hipMemcpyDeviceToHost));
    return SAPX_OK;
}

int32_t sapx_mlp_reserve(SapxMlp* M, int32_t batch) {
    if (!M || batch <= 0) return mfail(SAPX_ERR_INVALID, "sapx_mlp_reserve: bad batch");
    return reserve(M, batch);
}

int32_t sapx_mlp_forward(SapxMlp* M, const float* x, int32_t batch, float* y, void* stream) {
    if (!M || !x || !y || batch <= 0) return mfail(SAPX_ERR_INVALID, "sapx_mlp_forward: bad argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const int32_t rc = reserve(M, batch);
    if (rc != SAPX_OK) return rc;
    const int Bp = (batch + 63) / 64 * 64;
    int32_t r = normalise(M, x, batch, Bp, st);
    if (r != SAPX_OK) return r;
    r = forward_normalised(M, Bp, st);
    if (r != SAPX_OK) return r;
    const int O = M->out_features, F = M->in_features;
    hipLaunchKernelGGL(unnormalize_output_kernel, dim3((batch * O + 255) / 256), dim3(256), 0, st, M->y, batch, O,
                       M->lo.dims[M->lo.L], M->norm + 2 * F, M->norm + 2 * F + O, y);
    MLP_HIP(hipGetLastError());
    return SAPX_OK;
}

int32_t sapx_mlp_gradients(SapxMlp* M, const float* x, const float* target, int32_t batch, float* loss, void* stream) {
    if (!M || !x || !target || batch <= 0) return mfail(SAPX_ERR_INVALID, "sapx_mlp_gradients: bad argument");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    int32_t r = reserve(M, batch);
    if (r != SAPX_OK) return r;
    const int Bp = (batch + 63) / 64 * 64;
    const int L = M->lo.L, F = M->in_features, O = M->out_features;
    r = normalise(M, x, batch, Bp, st);
    if (r != SAPX_OK) return r;
    r = forward_normalised(M, Bp, st);
    if (r != SAPX_OK) return r;
    // bias gradients and the loss accumulate atomically: clear them with a fill kernel
    hipLaunchKernelGGL(fill_kernel, dim3((unsigned)((M->lo.total + 255) / 256)), dim3(256), 0, st, M->grads,
                       (size_t)M->lo.total, 0.0f);
    hipLaunchKernelGGL(fill_kernel, dim3(1), dim3(64), 0, st, M->loss, (size_t)1, 0.0f);
    MLP_HIP(hipGetLastError());
    const int dL = M->lo.dims[L];
    hipLaunchKernelGGL(loss_grad_kernel, dim3(Bp / 64), dim3(256), 0, st, M->y, target, batch, O, Bp, dL,
                       M->norm + 2 * F, M->norm + 2 * F + O, M->dz[0], M->dzt[0],
                       M->grads + M->lo.w_off[L - 1] + (long long)M->lo.dims[L - 1] * dL, M->loss);
    MLP_HIP(hipGetLastError());
    int cur = 0;
    for (int l = L - 1; l >= 0; --l) {
        const int din = M->lo.dims[l], dout = M->lo.dims[l + 1];
        // dW_l = dz_{l+1}^T h_l
        GemmArgs gw{};
        gw.A = M->dzt[cur];
        gw.lda = Bp;
        gw.B = l == 0 ? M->xt : M->ht[l];
        gw.ldb = Bp;
        gw.K = Bp;
        gw.Cf = M->grads + M->lo.w_off[l];
        gw.ldc = din;
        MLP_HIP(gemm(EPI_F32, gw, dout, din, st));
        if (l > 0) {
            // dz_l = (dz_{l+1} W_l) * [h_l > 0]; the bias gradient of layer l-1 is its column sum
            GemmArgs gx{};
            gx.A = M->dz[cur];
            gx.lda = dout;
            gx.B = M->wt + M->lo.wb_off[l];
            gx.ldb = dout;
            gx.K = dout;
            gx.Cb = M->dz[cur ^ 1];
            gx.Ct = M->dzt[cur ^ 1];
            gx.ldc = din;
            gx.ldt = Bp;
            gx.H = M->h[l];
            gx.ldh = din;
            gx.colsum = M->grads + M->lo.w_off[l - 1] + (long long)M->lo.dims[l - 1] * din;
            MLP_HIP(gemm(EPI_MASK_BF16, gx, Bp, din, st));
            cur ^= 1;
        }
    }
    if (loss) MLP_HIP(hipMemcpyAsync(loss, M->loss, 4, hipMemcpyDeviceToDevice, st));
    return SAPX_OK;
}

int32_t sapx_mlp_adam(SapxMlp* M, float lr, float beta1, float beta2, float eps, void* stream) {
    if (!M) return mfail(SAPX_ERR_INVALID, "sapx_mlp_adam: null model");
    M->adam_t += 1;
    MLP_HIP(refresh_bf16(M, reinterpret_cast<hipStream_t>(stream), 1, lr, beta1, beta2, eps));
    return SAPX_OK;
}

}  // extern "C"

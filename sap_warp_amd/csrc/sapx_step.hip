// Fused SAP timestep, one workgroup per world (gfx950).
//
// A world's whole step runs inside one workgroup with every intermediate in
// LDS: public->SAP boundary conversion, forward kinematics, bias forces,
// composite-rigid-body mass matrix, free motion, contact frames / Jacobian /
// weights, PD and limit terms, the SAP Newton loop (projection onto the
// regularised friction cone, gradient, Hessian, blocked Cholesky with MFMA
// Schur updates, exact-root or Armijo line search, convergence test),
// position integration and public forward kinematics.  HBM traffic is the
// compulsory state/control/contact reads and state writes only.
//
// Precision: kinematics, dynamics, the contact Jacobian, the Hessian and its
// Cholesky factor are float32 (the reference's approx32 preset solves its
// linear systems in float32, sim/solver_sap.py:78-87); velocities, the SAP
// objective, gradient, impulses and the line search are float64 accumulated
// over the float32 data, as in the preset's fp64 contact solve.
//
// Reference algorithm (file:line) for each phase is cited inline.
#include "sapx_kcommon.h"

namespace sapx {

// ---------------------------------------------------------------------------
// LDS view
// ---------------------------------------------------------------------------
struct Smem {
    float* A;       // n x ast  dynamics matrix
    float* H;       // npad x (npad+1)  Hessian / Cholesky
    float* J;       // 3C x js  contact Jacobian (kinematics scratch before it)
    float* bq;      // NB x 8  body poses (free-motion FK)
    float* S;       // N x 6  motion subspace (body origin, world)
    double* ck;     // C x kCk  rt, rn, vhat, mu, 1/rt, 1/rn, mu_hat, f
    double* cv;     // C x 9  vc, dvc, gamma
    float* cG;      // C x 9
    int* cb;        // C x 2 local bodies
    double* d64;    // V_NUM x npad f64 vectors
    float* d32;     // F_NUM x npad f32 vectors
    int* dflag;     // 4 x npad
    double* red;    // 8 x 8
    float* q;       // 2 x NQ coords (in, out)
    float* mp;      // per-world model parameters (prefetched)
    double* bqd;    // NB x 8  body poses in float64 (free-motion FK)
};

enum { V_V = 0, V_V0, V_VS, V_DV, V_G, V_JI, V_DP, V_AV, V_T1, V_T2, V_AVS, V_AR, V_NUM };
enum { F_HD = 0, F_AINV, F_DSC, F_PDA, F_PDG, F_PDL, F_LOV, F_UPV, F_LOR, F_UPR, F_TAU, F_JF, F_DINV, F_NUM };
enum { I_PART = 0, I_PD, I_LO, I_UP, I_NUM };

template <bool BIG>
__device__ __forceinline__ Smem smem_view(char* base, const StepParams& P) {
    Smem s;
    s.A = reinterpret_cast<float*>(base + P.o_A);
    s.H = reinterpret_cast<float*>(base + P.o_H);
    s.J = reinterpret_cast<float*>(base + P.o_J);
    s.bq = reinterpret_cast<float*>(base + P.o_bq);
    s.S = reinterpret_cast<float*>(base + P.o_S);
    s.ck = reinterpret_cast<double*>(base + P.o_ck);
    s.cv = reinterpret_cast<double*>(base + P.o_cv);
    s.cG = reinterpret_cast<float*>(base + P.o_cG);
    s.cb = reinterpret_cast<int*>(base + P.o_cb);
    s.d64 = reinterpret_cast<double*>(base + P.o_dv64);
    s.d32 = reinterpret_cast<float*>(base + P.o_dv32);
    s.dflag = reinterpret_cast<int*>(base + P.o_dflag);
    s.red = reinterpret_cast<double*>(base + P.o_red);
    s.q = reinterpret_cast<float*>(base + P.o_q);
    s.mp = reinterpret_cast<float*>(base + P.o_misc);
    s.bqd = reinterpret_cast<double*>(base + P.o_bqd);
    if constexpr (BIG) {
        float* g = P.g_work + (size_t)blockIdx.x * P.g_stride;
        s.A = g + P.g_A;
        s.H = g + P.g_H;
        s.J = g + P.g_J;
    }
    return s;
}

// ---------------------------------------------------------------------------
// Blocked Cholesky of H (npad x npad, row stride hs), in place, lower.
// Diagonal 16x16 blocks: wave-synchronous right-looking factor held in
// registers (one row per lane, readlane broadcasts), stopping at row n;
// panel TRSM: one lane per row using the stored inverse diagonal; trailing
// Schur updates: MFMA f32 16x16x4.  dinv[i] = 1 / L_ii.  Rows/cols >= n hold
// the identity (identity padding as in blocked_cholesky.py:725-901).
// ---------------------------------------------------------------------------
template <int NT>
__device__ __forceinline__ void cholesky_blocked(float* H, int npad, int hs, int n, float* dinv) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nb = npad >> 4;
    for (int kb = 0; kb < nb; ++kb) {
        const int k0 = kb * 16;
        const int jmax = min(16, n - k0);
        // 1. diagonal block factor (wave 0; lanes 0..15 own one row each)
        if (wave == 0) {
            float row[16];
            const int r = lane & 15;
#pragma unroll
            for (int c = 0; c < 16; ++c) row[c] = H[(k0 + r) * hs + k0 + c];
            float my_dinv = 1.0f;
#pragma unroll
            for (int j = 0; j < 16; ++j) {
                if (j < jmax) {
                    const float djj = fmaxf(rl(row[j], j), 1.0e-30f);
                    const float di = __builtin_amdgcn_rsqf(djj);
                    const float d = djj * di;
                    float lij = row[j];
                    if (r == j) { lij = d; my_dinv = di; }
                    else if (r > j) lij = row[j] * di;
                    row[j] = lij;
#pragma unroll
                    for (int c = j + 1; c < 16; ++c) {
                        const float lcj = rl(row[j], c);  // L[c][j]
                        if (r > j && c <= r) row[c] -= lij * lcj;
                    }
                }
            }
            if (lane < 16) {
#pragma unroll
                for (int c = 0; c < 16; ++c) H[(k0 + r) * hs + k0 + c] = (c <= r) ? row[c] : 0.0f;
                dinv[k0 + r] = my_dinv;
            }
        }
        __syncthreads();
        if (kb == nb - 1) break;
        // 2. panel TRSM: rows below the diagonal block that hold real dofs
        const int rows_below = min(npad, n) - k0 - 16;
        for (int t = threadIdx.x; t < rows_below; t += NT) {
            const int i = k0 + 16 + t;
            float x[16];
#pragma unroll
            for (int c = 0; c < 16; ++c) x[c] = H[i * hs + k0 + c];
#pragma unroll
            for (int c = 0; c < 16; ++c) {
                float sacc = x[c];
#pragma unroll
                for (int k = 0; k < c; ++k) sacc -= x[k] * H[(k0 + c) * hs + k0 + k];
                x[c] = sacc * dinv[k0 + c];
            }
#pragma unroll
            for (int c = 0; c < 16; ++c) H[i * hs + k0 + c] = x[c];
        }
        __syncthreads();
        // 3. trailing update with MFMA: tiles (ib, jb), kb < jb <= ib
        const int rem = nb - kb - 1;
        const int ntiles = rem * (rem + 1) / 2;
        for (int t = wave; t < ntiles; t += NW) {
            int ib = 0, acc_t = t;
            while (acc_t > ib) { acc_t -= ib + 1; ++ib; }
            const int jb = acc_t;
            const int i0 = (kb + 1 + ib) * 16;
            const int j0 = (kb + 1 + jb) * 16;
            typedef float f4 __attribute__((ext_vector_type(4)));
            f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
#pragma unroll
            for (int kk = 0; kk < 4; ++kk) {
                const int kcol = k0 + kk * 4 + (lane >> 4);
                float av = H[(i0 + (lane & 15)) * hs + kcol];
                float bv = H[(j0 + (lane & 15)) * hs + kcol];
                acc = __builtin_amdgcn_mfma_f32_16x16x4f32(av, bv, acc, 0, 0, 0);
            }
#pragma unroll
            for (int e = 0; e < 4; ++e) {
                const int i = i0 + 4 * (lane >> 4) + e;
                const int j = j0 + (lane & 15);
                H[i * hs + j] -= acc[e];
            }
        }
        __syncthreads();
    }
}

// Solve L L^T x = b (float32 factor), wave 0 only; b/x in LDS as double.
// Lane i owns x_i.  Serial substitution in 16-column blocks: each block's L
// entries are loaded into registers before its 16 dependent steps, so the
// chain is readlane -> multiply -> FMA without LDS latency.  Rows >= n are
// identity.  Requires n <= 64.
__device__ __forceinline__ void chol_solve_wave0(const float* H, int hs, int n, const float* dinv, double* b) {
    const int lane = threadIdx.x & 63;
    if ((threadIdx.x >> 6) != 0) return;
    float y = (lane < n) ? (float)b[lane] : 0.0f;
    const float di = (lane < n) ? dinv[lane] : 1.0f;
    const int nblk = (n + 15) >> 4;
    // forward: L y = b; lane i uses row i of L
    for (int jb = 0; jb < nblk; ++jb) {
        const int j0 = jb * 16;
        float L[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int j = j0 + k;
            L[k] = (lane > j && lane < n && j < n) ? H[lane * hs + j] : 0.0f;
        }
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int j = j0 + k;
            if (j < n) {
                const float yj = rl(y, j) * rl(di, j);
                if (lane == j) y = yj;
                y -= L[k] * yj;
            }
        }
    }
    // backward: L^T x = y; lane i uses column i of L
    for (int jb = nblk - 1; jb >= 0; --jb) {
        const int j0 = jb * 16;
        float L[16];
#pragma unroll
        for (int k = 0; k < 16; ++k) {
            const int j = j0 + k;
            L[k] = (lane < j && j < n) ? H[j * hs + lane] : 0.0f;
        }
#pragma unroll
        for (int k = 15; k >= 0; --k) {
            const int j = j0 + k;
            if (j < n) {
                const float xj = rl(y, j) * rl(di, j);
                if (lane == j) y = xj;
                y -= L[k] * xj;
            }
        }
    }
    if (lane < n) b[lane] = (double)y;
}

// Inverses of the diagonal 16x16 blocks of the Cholesky factor, one block per
// wave: lane c of the block's wave forward-substitutes column c of inv(L_kk)
// (all L reads are LDS broadcasts, no cross-lane chain).  The strictly lower
// part of inv(L_kk) is stored transposed in the block's strictly upper half,
// which the factorization leaves unused; its diagonal is dinv.
template <int NT>
__device__ __forceinline__ void chol_block_inverses(float* H, int npad, int hs, const float* dinv) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    const int nblk = npad >> 4;
    for (int kb = wave; kb < nblk; kb += NW) {
        const int k0 = kb * 16;
        const int c = lane & 15;
        float x[16];
#pragma unroll
        for (int i = 0; i < 16; ++i) {
            float acc = 0.0f;
#pragma unroll
            for (int m = 0; m < i; ++m) acc += H[(k0 + i) * hs + k0 + m] * x[m];
            const float di = dinv[k0 + i];
            x[i] = (i == c) ? di : ((i > c) ? -di * acc : 0.0f);
        }
        if (lane < 16) {
#pragma unroll
            for (int i = 1; i < 16; ++i)
                if (i > c) H[(k0 + c) * hs + k0 + i] = x[i];
        }
    }
    __syncthreads();
}

// Solve L L^T x = b with the blocked factor and its inverted diagonal blocks
// (chol_block_inverses), wave 0 only.  Lane i owns x_i.  Per 16-row block: the
// off-block contribution is a dot product over finished blocks (independent
// readlanes), then the block is one 16x16 triangular mat-vec.  Rows >= n are
// identity.  Requires npad <= 64.
__device__ __forceinline__ void chol_solve_blocks(const float* H, int npad, int hs, int n, const float* dinv,
                                                  double* b) {
    const int lane = threadIdx.x & 63;
    if ((threadIdx.x >> 6) != 0) return;
    const int nblk = npad >> 4;
    const int row = lane < npad ? lane : npad - 1;
    const float di = lane < n ? dinv[lane] : 1.0f;
    float y = (lane < n) ? (float)b[lane] : 0.0f;
    // forward: L y = b
    for (int kb = 0; kb < nblk; ++kb) {
        const int k0 = kb * 16;
        float r = y;
        for (int mb = 0; mb < kb; ++mb) {
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int m = mb * 16 + t;
                r -= H[row * hs + m] * rl(y, m);
            }
        }
        const int li = lane - k0;  // position inside the block (valid for 0..15)
        float yk = di * r;
#pragma unroll
        for (int t = 0; t < 15; ++t) {
            const float rt = rl(r, k0 + t);
            const float dk = (li > t && li < 16) ? H[(k0 + t) * hs + k0 + (li & 15)] : 0.0f;
            yk += dk * rt;
        }
        if (li >= 0 && li < 16) y = yk;
    }
    // backward: L^T x = y
    float x = y;
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int k0 = kb * 16;
        float r = x;
        for (int mb = kb + 1; mb < nblk; ++mb) {
#pragma unroll
            for (int t = 0; t < 16; ++t) {
                const int m = mb * 16 + t;
                r -= H[m * hs + row] * rl(x, m);
            }
        }
        const int li = lane - k0;
        float xk = di * r;
#pragma unroll
        for (int t = 1; t < 16; ++t) {
            const float rt = rl(r, k0 + t);
            const float dk = (li >= 0 && li < t) ? H[(k0 + (li & 15)) * hs + k0 + t] : 0.0f;
            xk += dk * rt;
        }
        if (li >= 0 && li < 16) x = xk;
    }
    if (lane < n) b[lane] = (double)x;
}

// Solve L L^T x = b for any npad (a multiple of 16) with the blocked factor and its inverted
// diagonal blocks (chol_block_inverses), wave 0 only; b is the float64 working vector in LDS or
// device memory and holds x on return.  Forward: per 16-row block, lanes 0..15 apply the block's
// inverse to its finished right-hand side, then the wave subtracts L_ik y_k from every row below.
// Backward: the transposed sweep from the last block.  Rows >= n are identity (b_i = 0 there).
__device__ __forceinline__ void chol_solve_wave_big(const float* H, int npad, int hs, int n, const float* dinv,
                                                    double* b) {
    const int lane = threadIdx.x & 63;
    if ((threadIdx.x >> 6) != 0) return;
    const int nblk = npad >> 4;
    for (int i = lane; i < npad; i += 64)
        if (i >= n) b[i] = 0.0;
    wave_sync();
    for (int kb = 0; kb < nblk; ++kb) {
        const int k0 = kb * 16;
        float yk = 0.0f;
        if (lane < 16) {
            float acc = dinv[k0 + lane] * (float)b[k0 + lane];
            for (int c = 0; c < lane; ++c) acc += H[(k0 + c) * hs + k0 + lane] * (float)b[k0 + c];
            yk = acc;
        }
        wave_sync();
        if (lane < 16) b[k0 + lane] = (double)yk;
        wave_sync();
        for (int i = k0 + 16 + lane; i < npad; i += 64) {
            float acc = (float)b[i];
#pragma unroll
            for (int t = 0; t < 16; ++t) acc -= H[i * hs + k0 + t] * (float)b[k0 + t];
            b[i] = (double)acc;
        }
        wave_sync();
    }
    for (int kb = nblk - 1; kb >= 0; --kb) {
        const int k0 = kb * 16;
        float xk = 0.0f;
        if (lane < 16) {
            float acc = dinv[k0 + lane] * (float)b[k0 + lane];
            for (int i = lane + 1; i < 16; ++i) acc += H[(k0 + lane) * hs + k0 + i] * (float)b[k0 + i];
            xk = acc;
        }
        wave_sync();
        if (lane < 16) b[k0 + lane] = (double)xk;
        wave_sync();
        for (int i = lane; i < k0; i += 64) {
            float acc = (float)b[i];
#pragma unroll
            for (int t = 0; t < 16; ++t) acc -= H[(k0 + t) * hs + i] * (float)b[k0 + t];
            b[i] = (double)acc;
        }
        wave_sync();
    }
    for (int i = lane; i < npad; i += 64)
        if (i >= n) b[i] = 0.0;
}

// the factor solve for the world's padded size: register rows up to 64, blocked beyond
__device__ __forceinline__ void chol_solve(const float* H, int npad, int hs, int n, const float* dinv, double* b) {
    if (npad <= 64) chol_solve_blocks(H, npad, hs, n, dinv, b);
    else chol_solve_wave_big(H, npad, hs, n, dinv, b);
}

__device__ __forceinline__ double model_terms(const Smem& s, int npad, int i, double vi, double dt, double* imp,
                                              double* hd) {
    double cost = 0.0;
    *imp = 0.0;
    *hd = 0.0;
    const float* d32 = s.d32;
    if (s.dflag[I_PD * npad + i]) {
        const double gain = d32[F_PDG * npad + i];
        const double lim = d32[F_PDL * npad + i];
        const double y = (double)d32[F_PDA * npad + i] - gain * vi;
        *imp += dt * clamp_s(y, lim);
        *hd += dt * gain * clamp_d(y, lim);
        if (gain > 0.0) cost += (dt / gain) * clamp_a(y, lim);
    }
    if (s.dflag[I_LO * npad + i]) {
        const double r = d32[F_LOR * npad + i];
        const double g = (1.0 / r) * ((double)d32[F_LOV * npad + i] - vi);
        if (g > 0.0) { *imp += g; *hd += 1.0 / r; cost += 0.5 * r * g * g; }
    }
    if (s.dflag[I_UP * npad + i]) {
        const double r = d32[F_UPR * npad + i];
        const double g = (1.0 / r) * ((double)d32[F_UPV * npad + i] + vi);
        if (g > 0.0) { *imp -= g; *hd += 1.0 / r; cost += 0.5 * r * g * g; }
    }
    return cost;
}

// ---------------------------------------------------------------------------
// The fused step kernel
// ---------------------------------------------------------------------------
// BIG: A, H and J in the world's block of device memory (P.g_work), the rest in LDS
template <int NT, bool BIG>
__global__ __launch_bounds__(NT, 2) void sapx_step_kernel(StepParams P) {
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    const DevModel& M = P.m;
    const int w = blockIdx.x;
    const int tid = threadIdx.x;
    const int n = M.N;
    const int nb = M.NB;
    const int nj = M.J;
    const int npad = P.npad;
    const int hs = npad + 1;
    const int js = P.js;
    const int ast = P.ast;  // row stride of A (odd: conflict-free row-per-thread access)
    const int C = P.C;
    Smem s = smem_view<BIG>(smem_raw, P);
    const Topo T = topo_view(reinterpret_cast<const int*>(smem_raw + P.o_topo), M);
    double* d64 = s.d64;
    float* d32 = s.d32;
    const double dt = P.dt_world ? P.dt_world[w] : P.dt;
    const float h = (float)dt;

    // scratch aliases inside the J region (valid until the Jacobian is built)
    float* kV = s.J;                 // nb x 6 body velocity
    float* kA = kV + nb * 6;          // nb x 6 bias acceleration
    float* kF = kA + nb * 6;          // nb x 6 bias force (incl. external)
    float* kI = kF + nb * 6;          // nb x 10 inertia about O: I(6 sym), m c(3), m
    float* kIc = kI + nb * 10;        // nb x 10 composite
    float* kFc = kIc + nb * 10;       // nb x 6 composite force at body origin
    float* kSO = kFc + nb * 6;        // n x 6 motion subspace about O
    float* kFi = kSO + n * 6;         // n x 6 composite inertia * SO
    float* kExt = kFi + n * 6;        // nb x 6 external wrench (SAP, negated)
    float* kQdPub = kExt + nb * 6;    // n public velocity
    float* q_in = s.q;
    float* q_out = s.q + M.NQ;

    // ------------------------------------------------------------------ 0/1
    // loads + public -> SAP boundary (sap_helpers.py:1482-1522, 1811-1848, 2058-2072)
    // per-world model parameters -> LDS (one coalesced pass; reused by every phase)
    float* mXp = s.mp;                 // nj x 8
    float* mXc = mXp + nj * 8;         // nj x 8
    float* mAx = mXc + nj * 8;         // n x 4
    float* mCom = mAx + n * 4;         // nb x 4
    float* mI = mCom + nb * 4;         // nb x 9
    float* mM = mI + nb * 9;           // nb
    for (int t = tid; t < nj * 7; t += NT) {
        const int j = t / 7, k = t % 7;
        mXp[j * 8 + k] = M.joint_X_p[(size_t)w * nj * 7 + t];
        mXc[j * 8 + k] = M.joint_X_c[(size_t)w * nj * 7 + t];
    }
    for (int t = tid; t < n * 3; t += NT) mAx[(t / 3) * 4 + t % 3] = M.joint_axis[(size_t)w * n * 3 + t];
    for (int t = tid; t < nb * 3; t += NT) mCom[(t / 3) * 4 + t % 3] = M.body_com[(size_t)w * nb * 3 + t];
    for (int t = tid; t < nb * 9; t += NT) mI[t] = M.body_inertia[(size_t)w * nb * 9 + t];
    for (int t = tid; t < nb; t += NT) mM[t] = M.body_mass[(size_t)w * nb + t];
    int* topo_lds = reinterpret_cast<int*>(smem_raw + P.o_topo);
    for (int k = tid; k < M.topo_ints; k += NT) topo_lds[k] = M.topo[k];
    for (int k = tid; k < M.NQ; k += NT) q_in[k] = P.jq_in[(size_t)w * M.NQ + k];
    __syncthreads();
    for (int d = tid; d < n; d += NT) {
        const size_t g = (size_t)w * n + d;
        kQdPub[d] = P.jqd_in[g];
        d64[V_V0 * npad + d] = (double)P.jqd_in[g];
        d32[F_JF * npad + d] = P.c_jf ? P.c_jf[g] : 0.0f;
    }
    for (int b = tid; b < nb; b += NT) {
        float* e = kExt + b * 6;
        for (int k = 0; k < 6; ++k) e[k] = 0.0f;
        if (P.bf_in) {
            const size_t gb = (size_t)w * nb + b;
            const float* f = P.bf_in + gb * 6;
            V3 fc = v3(f[0], f[1], f[2]);
            V3 tc = v3(f[3], f[4], f[5]);
            TF x = load_tf(P.bq_in + gb * 7);
            const float* com = mCom + b * 4;
            V3 c = qrot(qnormalize(x.q), v3(com[0], com[1], com[2]));
            V3 to = tc + cross(c, fc);
            e[0] = -to.x; e[1] = -to.y; e[2] = -to.z;
            e[3] = -fc.x; e[4] = -fc.y; e[5] = -fc.z;
        }
    }
    __syncthreads();
    for (int j = tid; j < nj; j += NT) {
        const int jt = T.jtype[j];
        if (jt == kJointFree || jt == kJointDistance) {
            const int ds = T.jds[j];
            const int child = T.jchild[j];
            const size_t gb = (size_t)w * nb + child;
            TF x = load_tf(P.bq_in + gb * 7);
            const float* com = mCom + child * 4;
            V3 r = qrot(qnormalize(x.q), v3(com[0], com[1], com[2]));
            const size_t g = (size_t)w * n + ds;
            double vc0 = P.jqd_in[g + 0], vc1 = P.jqd_in[g + 1], vc2 = P.jqd_in[g + 2];
            double w0 = P.jqd_in[g + 3], w1 = P.jqd_in[g + 4], w2 = P.jqd_in[g + 5];
            double rx = r.x, ry = r.y, rz = r.z;
            d64[V_V0 * npad + ds + 0] = w0;
            d64[V_V0 * npad + ds + 1] = w1;
            d64[V_V0 * npad + ds + 2] = w2;
            d64[V_V0 * npad + ds + 3] = vc0 - (w1 * rz - w2 * ry);
            d64[V_V0 * npad + ds + 4] = vc1 - (w2 * rx - w0 * rz);
            d64[V_V0 * npad + ds + 5] = vc2 - (w0 * ry - w1 * rx);
            if (P.c_jf) {
                V3 fc = v3(P.c_jf[g], P.c_jf[g + 1], P.c_jf[g + 2]);
                V3 tc = v3(P.c_jf[g + 3], P.c_jf[g + 4], P.c_jf[g + 5]);
                V3 to = tc + cross(r, fc);
                d32[F_JF * npad + ds + 0] = to.x;
                d32[F_JF * npad + ds + 1] = to.y;
                d32[F_JF * npad + ds + 2] = to.z;
                d32[F_JF * npad + ds + 3] = fc.x;
                d32[F_JF * npad + ds + 4] = fc.y;
                d32[F_JF * npad + ds + 5] = fc.z;
            }
        }
    }
    __syncthreads();

    // ------------------------------------------------------------------ 2
    // forward kinematics + body velocity / bias acceleration (free_motion.py:412-530,
    // sap_helpers.py:704-1433).  (a) every joint's local part in parallel: mobilizer
    // transform (float64) composed with X_c^-1, motion subspace / velocity / bias in the
    // parent joint frame; (b) wave 0 walks the tree level by level and only composes with
    // the parent and rotates the local parts, with wave-level ordering instead of block
    // barriers between levels.
    double* kXl = reinterpret_cast<double*>(kQdPub + ((n + 1) & ~1));  // nj x 8 local transforms
    for (int j = tid; j < nj; j += NT) {
        const int jt = T.jtype[j];
        const int child = T.jchild[j];
        const int qs = T.jqs[j];
        const int ds = T.jds[j];
        const TF Xc = load_tf(mXc + j * 8);
        const bool xcid = tf_is_identity(mXc + j * 8);
        const float* axp = mAx + ds * 4;
        const V3 axis = v3(axp[0], axp[1], axp[2]);
        const D3 axd = d3(axis.x, axis.y, axis.z);
        DTF Xl = DTF{d3(0.0, 0.0, 0.0), DQ{0.0, 0.0, 0.0, 1.0}};
        double plain = 0.0;  // 1: X_wc = [X_wpj.p, X_wpj.q * q_FM] without normalisation
        SV vl = sv(v3(0, 0, 0), v3(0, 0, 0));
        SV al = vl;
        D6Ang d6;
        if (jt == kJointRevolute && xcid) {
            // _sap_revolute_identity_child_kinematics (sap_helpers.py:704-720)
            Xl.q = dq_axis_angle(axd, (double)q_in[qs]);
            plain = 1.0;
            const SV S = sv(axis, v3(0, 0, 0));
            store_sv(s.S + ds * 6, S);
            vl = sv_scale(S, (float)d64[V_V0 * npad + ds]);
        } else {
            DTF Xj = Xl;
            if (jt == kJointPrismatic) {
                const double qv = q_in[qs];
                Xj.p = d3(axis.x * qv, axis.y * qv, axis.z * qv);
            } else if (jt == kJointRevolute) {
                Xj.q = dq_axis_angle(axd, (double)q_in[qs]);
            } else if (jt == kJointFree || jt == kJointDistance) {
                Xj = DTF{d3(q_in[qs], q_in[qs + 1], q_in[qs + 2]), DQ{q_in[qs + 3], q_in[qs + 4], q_in[qs + 5], q_in[qs + 6]}};
            } else if (jt == kJointBall) {
                Xj.q = DQ{q_in[qs], q_in[qs + 1], q_in[qs + 2], q_in[qs + 3]};
            } else if (jt == kJointD6) {
                // linear coordinates along their axes, then the angular rotation (indices clamped
                // into the joint's own dofs; validated at model creation: 1 <= nd <= 6)
                const int nl = T.jlin[j], last = T.jds[j + 1] - ds - 1;
                D3 c[3];
                double qa[3];
#pragma unroll
                for (int k = 0; k < 3; ++k) {
                    const int kl = min(k, last), ka = min(nl + k, last);
                    const float* al = mAx + (ds + kl) * 4;
                    const double ql = k < nl ? (double)q_in[qs + kl] : 0.0;
                    Xj.p = dadd(Xj.p, dscale(d3(al[0], al[1], al[2]), ql));
                    const float* aa = mAx + (ds + ka) * 4;
                    c[k] = d3(aa[0], aa[1], aa[2]);
                    qa[k] = (double)q_in[qs + ka];
                }
                d6 = d6_angular(last + 1 - nl, c[0], c[1], c[2], qa[0], qa[1], qa[2]);
                Xj.q = d6.rot;
            }
            Xl = xcid ? Xj : dtf_compose_n(Xj, dtf_inverse_n(dtf_from(Xc)));
            // mobilizer velocity / bias in the parent joint frame (sap_helpers.py:1289-1433)
            const TF Xmb = tf_inverse_n(Xc);
            const Q4 qj = Q4{(float)Xj.q.x, (float)Xj.q.y, (float)Xj.q.z, (float)Xj.q.w};
            const V3 pmb = qrot(qj, Xmb.p);
            if (jt == kJointPrismatic) {
                const SV S = sv(v3(0, 0, 0), axis);
                store_sv(s.S + ds * 6, S);
                vl = sv_scale(S, (float)d64[V_V0 * npad + ds]);
            } else if (jt == kJointRevolute) {
                const SV S = sv(axis, cross(axis, pmb));
                store_sv(s.S + ds * 6, S);
                const float qd = (float)d64[V_V0 * npad + ds];
                const V3 wf = axis * qd;
                al = sv(v3(0, 0, 0), cross(wf, cross(wf, pmb)));
                vl = sv_scale(S, qd);
            } else if (jt == kJointD6) {
                // linear axes fixed in the joint frame, angular axes at the current angles
                // (sap_helpers.py:1387-1432); the bias is w x (w x p) of the summed angular rate
                const int nl = T.jlin[j], nd = T.jds[j + 1] - ds;
                V3 wv = v3(0, 0, 0), vv = v3(0, 0, 0);
                for (int k = 0; k < nd; ++k) {
                    const bool lin = k < nl;
                    const int m = k - nl;
                    const D3 ad = dsel(m <= 0, d6.a0, dsel(m == 1, d6.a1, d6.a2));
                    const float* ap = mAx + (ds + k) * 4;
                    const V3 a = vsel(lin, v3(ap[0], ap[1], ap[2]), v3((float)ad.x, (float)ad.y, (float)ad.z));
                    const V3 z = v3(0, 0, 0);
                    store_sv(s.S + (ds + k) * 6, sv(vsel(lin, z, a), vsel(lin, a, cross(a, pmb))));
                    const float qd = (float)d64[V_V0 * npad + ds + k];
                    wv = wv + vsel(lin, z, a * qd);
                    vv = vv + vsel(lin, a * qd, z);
                }
                vl = sv(wv, vv + cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
            } else if (jt == kJointFree || jt == kJointDistance) {
                const double* v0 = d64 + V_V0 * npad + ds;
                const V3 wv = v3((float)v0[0], (float)v0[1], (float)v0[2]);
                const V3 vv = v3((float)v0[3], (float)v0[4], (float)v0[5]);
                vl = sv(wv, vv + cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
                for (int k = 0; k < 3; ++k) {
                    const V3 e = v3(k == 0, k == 1, k == 2);
                    store_sv(s.S + (ds + k) * 6, sv(e, cross(e, pmb)));
                    store_sv(s.S + (ds + 3 + k) * 6, sv(v3(0, 0, 0), e));
                }
            } else if (jt == kJointBall) {
                // three angular dofs in the joint frame (solver_sap.py:316-330)
                const double* v0 = d64 + V_V0 * npad + ds;
                const V3 wv = v3((float)v0[0], (float)v0[1], (float)v0[2]);
                vl = sv(wv, cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
                for (int k = 0; k < 3; ++k) {
                    const V3 e = v3(k == 0, k == 1, k == 2);
                    store_sv(s.S + (ds + k) * 6, sv(e, cross(e, pmb)));
                }
            }
        }
        store_dtf(kXl + j * 8, Xl);
        kXl[j * 8 + 7] = plain;
        store_sv(kV + child * 6, vl);
        store_sv(kA + child * 6, al);
    }
    __syncthreads();
    if (tid < 64) {
        const int lane = tid;
        for (int L = 0; L < M.num_levels; ++L) {
            const int l0 = T.level_start[L], l1 = T.level_start[L + 1];
            for (int t = l0 + lane; t < l1; t += 64) {
                const int j = T.level_joint[t];
                const int parent = T.jparent[j];
                const int child = T.jchild[j];
                const int ds = T.jds[j], nd = T.jds[j + 1] - ds;
                const DTF Xp_d = dtf_from(load_tf(mXp + j * 8));
                const DTF Xwpj_d = parent >= 0 ? dtf_compose_n(load_dtf(s.bqd + parent * 8), Xp_d) : Xp_d;
                const DTF Xl = load_dtf(kXl + j * 8);
                const DTF Xwc_d = (kXl[j * 8 + 7] != 0.0) ? DTF{Xwpj_d.p, dqmul(Xwpj_d.q, Xl.q)}
                                                          : dtf_compose_n(Xwpj_d, Xl);
                const Q4 R = tf_from(Xwpj_d).q;
                for (int k = 0; k < nd; ++k) {
                    const SV Sl = load_sv(s.S + (ds + k) * 6);
                    store_sv(s.S + (ds + k) * 6, sv(qrot(R, Sl.w), qrot(R, Sl.v)));
                }
                const SV vl = load_sv(kV + child * 6);
                const SV al = load_sv(kA + child * 6);
                const SV vPB = sv(qrot(R, vl.w), qrot(R, vl.v));
                const SV aPB = sv(qrot(R, al.w), qrot(R, al.v));
                store_dtf(s.bqd + child * 8, Xwc_d);
                const TF Xwc = tf_from(Xwc_d);
                float* bqc = s.bq + child * 8;
                bqc[0] = Xwc.p.x; bqc[1] = Xwc.p.y; bqc[2] = Xwc.p.z;
                bqc[3] = Xwc.q.x; bqc[4] = Xwc.q.y; bqc[5] = Xwc.q.z; bqc[6] = Xwc.q.w;
                SV V = vPB, A = aPB;
                if (parent >= 0) {
                    const V3 pPB = Xwc.p - v3(s.bq[parent * 8], s.bq[parent * 8 + 1], s.bq[parent * 8 + 2]);
                    const SV Vp = load_sv(kV + parent * 6);
                    const SV Ap = load_sv(kA + parent * 6);
                    V = compose_velocity(Vp, pPB, vPB);
                    A = compose_acceleration(Ap, Vp, pPB, vPB, aPB);
                }
                store_sv(kV + child * 6, V);
                store_sv(kA + child * 6, A);
            }
            // the next level reads this level's LDS writes from other lanes of this wave
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();

    // ------------------------------------------------------------------ 3
    // body dynamics (free_motion.py:680-719) about each body origin. The composite inertias,
    // the bias forces and the dynamics matrix are taken about the origin of the child body of
    // each dof's joint, with lever arms between bodies from the float64 poses (about one point
    // for a whole tree, float32 loses a light link's small inertia to the parallel-axis terms
    // of the bodies between); the motion subspace about O = the origin of the body's tree root
    // serves the contact Jacobian. p_b - O, formed from the float64 poses:
    auto tree_offset = [&](int b) {
        const int r = T.body_root[b];
        return v3((float)(s.bqd[b * 8] - s.bqd[r * 8]), (float)(s.bqd[b * 8 + 1] - s.bqd[r * 8 + 1]),
                  (float)(s.bqd[b * 8 + 2] - s.bqd[r * 8 + 2]));
    };
    for (int b = tid; b < nb; b += NT) {
        const size_t gb = (size_t)w * nb + b;
        const float m = mM[b];
        const float* Ip = mI + b * 9;
        const float* cp = mCom + b * 4;
        const float* g = M.gravity + (size_t)w * 3;
        const float* bqb = s.bq + b * 8;
        Q4 q = Q4{bqb[3], bqb[4], bqb[5], bqb[6]};
        V3 c0 = qrot(q, v3(1, 0, 0)), c1 = qrot(q, v3(0, 1, 0)), c2 = qrot(q, v3(0, 0, 1));
        // I_W = R I R^T
        float Rm[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
        float T[3][3], IW[3][3];
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
                T[r][cc] = Rm[r][0] * Ip[0 * 3 + cc] + Rm[r][1] * Ip[1 * 3 + cc] + Rm[r][2] * Ip[2 * 3 + cc];
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
                IW[r][cc] = T[r][0] * Rm[cc][0] + T[r][1] * Rm[cc][1] + T[r][2] * Rm[cc][2];
        V3 c = qrot(q, v3(cp[0], cp[1], cp[2]));
        float cc2 = dot(c, c);
        float IB[3][3];
        float cv_[3] = {c.x, c.y, c.z};
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) IB[r][k] = IW[r][k] + m * ((r == k ? cc2 : 0.0f) - cv_[r] * cv_[k]);
        SV V = load_sv(kV + b * 6);
        SV A = load_sv(kA + b * 6);
        V3 Iw = v3(IB[0][0] * V.w.x + IB[0][1] * V.w.y + IB[0][2] * V.w.z,
                   IB[1][0] * V.w.x + IB[1][1] * V.w.y + IB[1][2] * V.w.z,
                   IB[2][0] * V.w.x + IB[2][1] * V.w.y + IB[2][2] * V.w.z);
        V3 Ia = v3(IB[0][0] * A.w.x + IB[0][1] * A.w.y + IB[0][2] * A.w.z,
                   IB[1][0] * A.w.x + IB[1][1] * A.w.y + IB[1][2] * A.w.z,
                   IB[2][0] * A.w.x + IB[2][1] * A.w.y + IB[2][2] * A.w.z);
        // I_s * A = [I_B alpha + m c x a, -m c x alpha + m a]
        V3 fw = Ia + cross(c, A.v) * m;
        V3 fv = (A.v - cross(c, A.w)) * m;
        // bias [w x I_B w, m w x (w x c)]
        fw = fw + cross(V.w, Iw);
        fv = fv + cross(V.w, cross(V.w, c)) * m;
        // gravity [c x m g, m g]
        V3 mg = v3(g[0], g[1], g[2]) * m;
        fw = fw - cross(c, mg);
        fv = fv - mg;
        const float* e = kExt + b * 6;
        float* F = kF + b * 6;
        F[0] = fw.x + e[0]; F[1] = fw.y + e[1]; F[2] = fw.z + e[2];
        F[3] = fv.x + e[3]; F[4] = fv.y + e[4]; F[5] = fv.z + e[5];
        // inertia about the body origin, first moment m c, mass
        float* I = kI + b * 10;
        int e6 = 0;
        for (int r = 0; r < 3; ++r)
            for (int k = r; k < 3; ++k) I[e6++] = IB[r][k];
        I[6] = m * c.x; I[7] = m * c.y; I[8] = m * c.z; I[9] = m;
    }
    __syncthreads();

    // ------------------------------------------------------------------ 4
    // composites over subtrees about each body origin: reverse level sweep on wave 0, each
    // body pulls its children's finished composites shifted by d = p_ch - p_c (float64
    // poses): I += I_ch + m (|d|^2 1 - d d^T) + 2 (d.h) 1 - d h^T - h d^T, h += h_ch + m d,
    // torque += torque_ch + d x f (deterministic order; wave-level ordering between levels)
    if (tid < 64) {
        const int lane = tid;
        for (int L = M.num_levels - 1; L >= 0; --L) {
            const int l0 = T.level_start[L], l1 = T.level_start[L + 1];
            for (int t = l0 + lane; t < l1; t += 64) {
                const int c = T.jchild[T.level_joint[t]];
                float Ic[10];
#pragma unroll
                for (int k = 0; k < 10; ++k) Ic[k] = kI[c * 10 + k];
                const float* F = kF + c * 6;
                V3 tw = v3(F[0], F[1], F[2]), fv = v3(F[3], F[4], F[5]);
                const double pcx = s.bqd[c * 8], pcy = s.bqd[c * 8 + 1], pcz = s.bqd[c * 8 + 2];
                for (int e = T.child_start[c]; e < T.child_start[c + 1]; ++e) {
                    const int ch = T.child[e];
                    const float* Ich = kIc + ch * 10;
                    const V3 dl = v3((float)(s.bqd[ch * 8] - pcx), (float)(s.bqd[ch * 8 + 1] - pcy),
                                     (float)(s.bqd[ch * 8 + 2] - pcz));
                    const float md = Ich[9];
                    const V3 hd = v3(Ich[6], Ich[7], Ich[8]);
                    const float dd = dot(dl, dl), dh2 = 2.0f * dot(dl, hd);
                    Ic[0] += Ich[0] + md * (dd - dl.x * dl.x) + dh2 - 2.0f * dl.x * hd.x;
                    Ic[1] += Ich[1] - md * dl.x * dl.y - dl.x * hd.y - hd.x * dl.y;
                    Ic[2] += Ich[2] - md * dl.x * dl.z - dl.x * hd.z - hd.x * dl.z;
                    Ic[3] += Ich[3] + md * (dd - dl.y * dl.y) + dh2 - 2.0f * dl.y * hd.y;
                    Ic[4] += Ich[4] - md * dl.y * dl.z - dl.y * hd.z - hd.y * dl.z;
                    Ic[5] += Ich[5] + md * (dd - dl.z * dl.z) + dh2 - 2.0f * dl.z * hd.z;
                    Ic[6] += hd.x + md * dl.x; Ic[7] += hd.y + md * dl.y; Ic[8] += hd.z + md * dl.z;
                    Ic[9] += md;
                    const float* Fch = kFc + ch * 6;
                    const V3 f = v3(Fch[3], Fch[4], Fch[5]);
                    // shift_force(F_ch, p_c - p_ch) = [tau + (p_ch - p_c) x f, f] (sap_helpers.py:962-966)
                    tw = tw + v3(Fch[0], Fch[1], Fch[2]) + cross(dl, f);
                    fv = fv + f;
                }
#pragma unroll
                for (int k = 0; k < 10; ++k) kIc[c * 10 + k] = Ic[k];
                float* Fc = kFc + c * 6;
                Fc[0] = tw.x; Fc[1] = tw.y; Fc[2] = tw.z; Fc[3] = fv.x; Fc[4] = fv.y; Fc[5] = fv.z;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();
    for (int d = tid; d < n; d += NT) {
        const int j = T.dof_joint[d];
        const int c = T.jchild[j];
        SV S = load_sv(s.S + d * 6);
        // tau = joint_f - S . f_s  (free_motion.py:93-147, sap_helpers.py:2090-2118)
        d32[F_TAU * npad + d] = d32[F_JF * npad + d] - sv_dot(S, load_sv(kFc + c * 6));
        // S about O (contact Jacobian)
        SV SO = shift_velocity(S, -tree_offset(c));
        store_sv(kSO + d * 6, SO);
        // composite inertia of the joint's child subtree times S, both about p_c
        const float* I = kIc + c * 10;
        V3 hv = v3(I[6], I[7], I[8]);
        V3 Iw = v3(I[0] * S.w.x + I[1] * S.w.y + I[2] * S.w.z, I[1] * S.w.x + I[3] * S.w.y + I[4] * S.w.z,
                   I[2] * S.w.x + I[4] * S.w.y + I[5] * S.w.z);
        SV Fi = sv(Iw + cross(hv, S.v), S.v * I[9] - cross(hv, S.w));
        store_sv(kFi + d * 6, Fi);
    }
    __syncthreads();
    // mass matrix by composite rigid bodies: H_ij = (Ic S_i) . S_j about p_ci for joint(j) on
    // joint(i)'s root path (equals J^T M J of free_motion.py:843-1179); S_j moves to p_ci by
    // v + w x (p_ci - p_cj), the difference of the two float32 tree offsets
    auto crb = [&](int a, int b) {  // (Ic S_a) . S_b about p_ca
        const V3 dl = tree_offset(T.jchild[T.dof_joint[a]]) - tree_offset(T.jchild[T.dof_joint[b]]);
        const SV Sb = load_sv(s.S + b * 6);
        return sv_dot(load_sv(kFi + a * 6), sv(Sb.w, Sb.v + cross(Sb.w, dl)));
    };
    for (int t = tid; t < npad * npad; t += NT) {
        const int i = t / npad, j = t % npad;
        float val;
        if (i < n && j < n) {
            const int ji = T.dof_joint[i], jj = T.dof_joint[j];
            const unsigned long long anc_i = T.body_anc[T.jchild[ji]];
            const unsigned long long anc_j = T.body_anc[T.jchild[jj]];
            // two dofs of one joint lie on each other's path: both triangles take
            // (Ic S_max(i,j)) . S_min(i,j), so A is exactly symmetric in float32 (the two
            // products differ by rounding, and the solve's gradient and line search read A
            // from opposite sides)
            const bool up_i = bit(anc_i, jj), up_j = bit(anc_j, ji);
            val = 0.0f;
            if (up_i && (!up_j || j <= i)) val = crb(i, j);
            else if (up_j) val = crb(j, i);
            if (i == j) val += M.armature[(size_t)w * n + i];
            s.A[i * ast + j] = val;
        } else {
            val = (i == j) ? 1.0f : 0.0f;
        }
        s.H[i * hs + j] = val;
    }
    __syncthreads();

    // ------------------------------------------------------------------ 5
    // free motion v* = v0 + dt A^-1 tau with one refinement step (free_motion.py:1970-2178)
    cholesky_blocked<NT>(s.H, npad, hs, n, d32 + F_DINV * npad);
    chol_block_inverses<NT>(s.H, npad, hs, d32 + F_DINV * npad);
    for (int d = tid; d < npad; d += NT) d64[V_T1 * npad + d] = (d < n) ? (double)d32[F_TAU * npad + d] : 0.0;
    __syncthreads();
    chol_solve(s.H, npad, hs, n, d32 + F_DINV * npad, d64 + V_T1 * npad);
    __syncthreads();
    for (int i = tid; i < n; i += NT) {
        d64[V_T2 * npad + i] = (double)d32[F_TAU * npad + i] - dot_fd(s.A + i * ast, d64 + V_T1 * npad, n);
    }
    __syncthreads();
    chol_solve(s.H, npad, hs, n, d32 + F_DINV * npad, d64 + V_T2 * npad);
    __syncthreads();
    for (int d = tid; d < npad; d += NT) {
        double qdd = (d < n) ? d64[V_T1 * npad + d] + d64[V_T2 * npad + d] : 0.0;
        double v0 = (d < n) ? d64[V_V0 * npad + d] : 0.0;
        d64[V_VS * npad + d] = v0 + dt * qdd;
    }
    __syncthreads();

    // ------------------------------------------------------------------ 6
    // contacts of this world (contact_jacobian.py:404-562, sap_helpers.py:148-503)
    int nc = 0;
    int c_begin = 0;
    if (P.cworld_start) {
        c_begin = P.cworld_start[w];
        nc = P.cworld_start[w + 1] - c_begin;
    }
    if (nc > C) nc = C;
    if (nc < 0) nc = 0;
    float* cx0 = s.H;            // C x 3 witness 0 (world)
    float* cx1 = cx0 + C * 3;    // witness 1
    float* cpc = cx1 + C * 3;    // contact point
    float* cR = cpc + C * 3;     // C x 9 frame columns cx, cy, n
    double* cphi = reinterpret_cast<double*>(cR + C * 9);  // C x 4: phi0, k, tau, mu (raw)
    for (int c = tid; c < nc; c += NT) {
        const int idx = c_begin + c;
        const int sh0 = P.cshape0[idx], sh1 = P.cshape1[idx];
        const int b0g = sh0 >= 0 ? M.shape_body[sh0] : -1;
        const int b1g = sh1 >= 0 ? M.shape_body[sh1] : -1;
        const int lb0 = b0g >= 0 ? b0g - w * nb : -1;
        const int lb1 = b1g >= 0 ? b1g - w * nb : -1;
        s.cb[c * 2] = lb0;
        s.cb[c * 2 + 1] = lb1;
        // world witness points and the signed distance in float64 from the float64 poses
        D3 x0d = d3(P.cpoint0[idx * 3], P.cpoint0[idx * 3 + 1], P.cpoint0[idx * 3 + 2]);
        D3 x1d = d3(P.cpoint1[idx * 3], P.cpoint1[idx * 3 + 1], P.cpoint1[idx * 3 + 2]);
        if (lb0 >= 0) x0d = dtf_point(load_dtf(s.bqd + lb0 * 8), x0d);
        if (lb1 >= 0) x1d = dtf_point(load_dtf(s.bqd + lb1 * 8), x1d);
        D3 nd = d3(P.cnormal[idx * 3], P.cnormal[idx * 3 + 1], P.cnormal[idx * 3 + 2]);
        const double nd2 = nd.x * nd.x + nd.y * nd.y + nd.z * nd.z;
        nd = (nd2 <= 1.0e-32) ? d3(0.0, 0.0, 0.0) : dscale(nd, 1.0 / sqrt(nd2));
        const V3 x0 = v3((float)x0d.x, (float)x0d.y, (float)x0d.z);
        const V3 x1 = v3((float)x1d.x, (float)x1d.y, (float)x1d.z);
        V3 nrm = v3((float)nd.x, (float)nd.y, (float)nd.z);
        // frame: MakeFromOneUnitVector (sap_helpers.py:467-503)
        float ax = fabsf(nrm.x), ay = fabsf(nrm.y), az = fabsf(nrm.z);
        int im = 0;
        float mn = ax;
        if (ay < mn) { im = 1; mn = ay; }
        if (az < mn) { im = 2; mn = az; }
        float mag = sqrtf(1.0f - mn * mn);
        float rr = 1.0f / mag;
        float u = im == 0 ? nrm.x : (im == 1 ? nrm.y : nrm.z);
        float sg = -rr * u;
        V3 tx, ty;
        if (im == 0) { tx = v3(0.0f, -rr * nrm.z, rr * nrm.y); ty = v3(mag, sg * nrm.y, sg * nrm.z); }
        else if (im == 1) { tx = v3(rr * nrm.z, 0.0f, -rr * nrm.x); ty = v3(sg * nrm.x, mag, sg * nrm.z); }
        else { tx = v3(-rr * nrm.y, rr * nrm.x, 0.0f); ty = v3(sg * nrm.x, sg * nrm.y, mag); }
        // materials (sap_helpers.py:148-237)
        double k0 = P.fallback_k, k1 = P.fallback_k;
        if (sh0 >= 0) k0 = M.shape_ke[sh0];
        if (!isfinite(k0) || k0 <= 0.0) k0 = P.fallback_k;
        if (!isfinite(k0) || k0 <= 0.0) k0 = 1.0;
        if (sh1 >= 0) k1 = M.shape_ke[sh1];
        if (!isfinite(k1) || k1 <= 0.0) k1 = P.fallback_k;
        if (!isfinite(k1) || k1 <= 0.0) k1 = 1.0;
        double kden = k0 + k1;
        double w0 = 0.5, w1 = 0.5;
        if (kden != 0.0) { w0 = k0 / kden; w1 = k1 / kden; }
        const D3 pcd = dadd(dscale(x0d, w0), dscale(x1d, w1));
        V3 pc = v3((float)pcd.x, (float)pcd.y, (float)pcd.z);
        double kc;
        if (!isfinite(k0)) kc = k1;
        else if (!isfinite(k1)) kc = k0;
        else if (kden <= 1.0e-12) kc = fmax(k0, k1);
        else kc = k0 * k1 / kden;
        double mu0 = P.fallback_mu, mu1 = P.fallback_mu;
        if (sh0 >= 0) mu0 = M.shape_mu[sh0];
        if (sh1 >= 0) mu1 = M.shape_mu[sh1];
        mu0 = fmax(mu0, 0.0);
        mu1 = fmax(mu1, 0.0);
        double mud = mu0 + mu1;
        double muc = (mud <= 1.0e-12) ? 0.0 : 2.0 * mu0 * mu1 / mud;
        double t0 = P.fallback_tau, t1 = P.fallback_tau;
        if (sh0 >= 0) t0 = M.shape_tau[sh0];
        if (!isfinite(t0) || t0 < 0.0) t0 = P.fallback_tau;
        if (!isfinite(t0) || t0 < 0.0) t0 = 0.0;
        if (sh1 >= 0) t1 = M.shape_tau[sh1];
        if (!isfinite(t1) || t1 < 0.0) t1 = P.fallback_tau;
        if (!isfinite(t1) || t1 < 0.0) t1 = 0.0;
        const D3 gapv = dsub(x1d, x0d);
        double phi = gapv.x * nd.x + gapv.y * nd.y + gapv.z * nd.z - (double)P.cmargin0[idx] - (double)P.cmargin1[idx];
        cx0[c * 3] = x0.x; cx0[c * 3 + 1] = x0.y; cx0[c * 3 + 2] = x0.z;
        cx1[c * 3] = x1.x; cx1[c * 3 + 1] = x1.y; cx1[c * 3 + 2] = x1.z;
        cpc[c * 3] = pc.x; cpc[c * 3 + 1] = pc.y; cpc[c * 3 + 2] = pc.z;
        float* R = cR + c * 9;
        R[0] = tx.x; R[1] = tx.y; R[2] = tx.z; R[3] = ty.x; R[4] = ty.y; R[5] = ty.z;
        R[6] = nrm.x; R[7] = nrm.y; R[8] = nrm.z;
        cphi[c * 4] = phi;
        cphi[c * 4 + 1] = kc;
        cphi[c * 4 + 2] = t0 + t1;
        cphi[c * 4 + 3] = muc;
    }
    __syncthreads();
    // contact Jacobian: J_c[:, d] = R^T (u1 - u0),  u_s = S_v + S_w x (x_s - p_child(d))
    // for d on body s's root path (contact_jacobian.py:519-562)
    for (int t = tid; t < nc * js; t += NT) {
        const int c = t / js, d = t % js;
        float j0 = 0.0f, j1 = 0.0f, j2 = 0.0f;
        if (d < n) {
            const int jd = T.dof_joint[d];
            const int cd = T.jchild[jd];
            SV S = load_sv(s.S + d * 6);
            V3 pcd = v3(s.bq[cd * 8], s.bq[cd * 8 + 1], s.bq[cd * 8 + 2]);
            V3 rel = v3(0, 0, 0);
            const int lb0 = s.cb[c * 2], lb1 = s.cb[c * 2 + 1];
            const bool wit = P.point_mode == 0;
            if (lb1 >= 0 && bit(T.body_anc[lb1], jd)) {
                V3 xs = wit ? v3(cx1[c * 3], cx1[c * 3 + 1], cx1[c * 3 + 2]) : v3(cpc[c * 3], cpc[c * 3 + 1], cpc[c * 3 + 2]);
                rel = rel + S.v + cross(S.w, xs - pcd);
            }
            if (lb0 >= 0 && bit(T.body_anc[lb0], jd)) {
                V3 xs = wit ? v3(cx0[c * 3], cx0[c * 3 + 1], cx0[c * 3 + 2]) : v3(cpc[c * 3], cpc[c * 3 + 1], cpc[c * 3 + 2]);
                rel = rel - (S.v + cross(S.w, xs - pcd));
            }
            const float* R = cR + c * 9;
            j0 = R[0] * rel.x + R[1] * rel.y + R[2] * rel.z;
            j1 = R[3] * rel.x + R[4] * rel.y + R[5] * rel.z;
            j2 = R[6] * rel.x + R[7] * rel.y + R[8] * rel.z;
        }
        s.J[(c * 3 + 0) * js + d] = j0;
        s.J[(c * 3 + 1) * js + d] = j1;
        s.J[(c * 3 + 2) * js + d] = j2;
    }
    // diag(A) data (contact_solve.py:405-419)
    for (int d = tid; d < npad; d += NT) {
        double a = (d < n) ? (double)s.A[d * ast + d] : 1.0;
        if (a < 1.0e-12 || !isfinite(a)) a = 1.0e-12;
        d32[F_AINV * npad + d] = (float)(1.0 / a);
        d32[F_DSC * npad + d] = (float)(1.0 / sqrt(a));
    }
    __syncthreads();
    // effective contact weights + regularisation constants (contact_jacobian.py:153-251,
    // contact_solve.py:791-822)
    const double bf = P.beta * P.beta / (4.0 * kPi * kPi);
    for (int c = tid; c < nc; c += NT) {
        double wv;
        if (P.weight_mode == 0) {
            const float* R = cR + c * 9;
            V3 dirs[3] = {v3(R[0], R[1], R[2]), v3(R[3], R[4], R[5]), v3(R[6], R[7], R[8])};
            double acc = 0.0;
            for (int side = 0; side < 2; ++side) {
                const int lb = s.cb[c * 2 + side];
                if (lb < 0) continue;
                const size_t gb = (size_t)w * nb + lb;
                const double m = mM[lb];
                if (m <= 1.0e-12) continue;
                const float* bb = s.bq + lb * 8;
                Q4 q = qnormalize(Q4{bb[3], bb[4], bb[5], bb[6]});
                const float* com = mCom + lb * 4;
                V3 comw = v3(bb[0], bb[1], bb[2]) + qrot(q, v3(com[0], com[1], com[2]));
                const float* xs = side == 0 ? cx0 + c * 3 : cx1 + c * 3;
                V3 r = v3(xs[0], xs[1], xs[2]) - comw;
                const float* I = mI + lb * 9;
                double ix = fmax((double)I[0], 1.0e-12), iy = fmax((double)I[4], 1.0e-12), iz = fmax((double)I[8], 1.0e-12);
                for (int k = 0; k < 3; ++k) {
                    V3 dn = dirs[k];
                    float dl = sqrtf(dot(dn, dn));
                    dn = (dl * dl <= 1.0e-32f) ? v3(0, 0, 0) : dn * (1.0f / dl);
                    V3 cb = qrot(qconj(q), cross(r, dn));
                    acc += 1.0 / m + (double)cb.x * cb.x / ix + (double)cb.y * cb.y / iy + (double)cb.z * cb.z / iz;
                }
            }
            wv = acc / 3.0;
        } else {
            double W00 = 0, W01 = 0, W02 = 0, W11 = 0, W12 = 0, W22 = 0;
            for (int d = 0; d < n; ++d) {
                const double ai = d32[F_AINV * npad + d];
                const double a0 = s.J[(c * 3) * js + d], a1 = s.J[(c * 3 + 1) * js + d], a2 = s.J[(c * 3 + 2) * js + d];
                W00 += ai * a0 * a0; W01 += ai * a0 * a1; W02 += ai * a0 * a2;
                W11 += ai * a1 * a1; W12 += ai * a1 * a2; W22 += ai * a2 * a2;
            }
            wv = sqrt(W00 * W00 + W11 * W11 + W22 * W22 + 2.0 * (W01 * W01 + W02 * W02 + W12 * W12)) / 3.0;
        }
        if (wv < 1.0e-12 || !isfinite(wv)) wv = 1.0e-12;
        double kc = cphi[c * 4 + 1];
        if (kc <= 0.0 || !isfinite(kc)) kc = 1.0;
        const double tau_c = fmax(cphi[c * 4 + 2], 0.0);
        const double rn_soft = 1.0 / (dt * kc * (dt + tau_c));
        const double rn = fmax(fmax(bf * wv, rn_soft), 1.0e-30);
        const double rt = fmax(P.sigma * wv, 1.0e-30);
        double mu = cphi[c * 4 + 3];
        if (mu < 0.0 || !isfinite(mu)) mu = 0.0;
        // per-contact constants of the projection, hoisted out of every evaluation
        double* kc8 = s.ck + c * kCk;
        kc8[0] = rt;
        kc8[1] = rn;
        kc8[2] = -cphi[c * 4] / (dt + tau_c);
        kc8[3] = mu;
        kc8[4] = 1.0 / rt;
        kc8[5] = 1.0 / rn;
        const double mu_tilde = mu * sqrt(rt / rn);
        kc8[6] = mu * rt / rn;                          // mu_hat
        kc8[7] = 1.0 / (1.0 + mu_tilde * mu_tilde);     // f
        if (P.dbg_phi0) P.dbg_phi0[(size_t)w * C + c] = cphi[c * 4];
        if (P.dbg_w) P.dbg_w[(size_t)w * C + c] = wv;
    }
    __syncthreads();

    // ------------------------------------------------------------------ 7
    // PD + limit terms (contact_solve.py:490-683) and participating dofs (422-487)
    for (int i = tid; i < npad; i += NT) {
        int pd = 0, lo = 0, up = 0, part = 0;
        float pda = 0, pdg = 0, pdl = 0, lov = 0, upv = 0, lor = 1, upr = 1;
        if (i < n) {
            const int t = T.dof_target[i];
            const size_t gt = (size_t)w * n + t;
            const size_t gi = (size_t)w * n + i;
            const int coord = T.dof_coord[i];
            const double ainv = d32[F_AINV * npad + i];
            if (M.target_mode[gt] != 0) {
                const double kp = M.target_ke[gt], kd = M.target_kd[gt];
                if (!(kp <= 0.0 && kd <= 0.0) && !(coord < 0 && kp > 0.0)) {
                    // effective gains (sap_helpers.py:2305-2340)
                    const double pbf = kPdBeta * kPdBeta / (4.0 * kPi * kPi);
                    const double r_nr = pbf * fmax(ainv, 1.0e-12);
                    const double common = dt * kp + kd;
                    double kpe = 0.0, kde = 0.0;
                    if (common > 0.0) {
                        const double r = 1.0 / (dt * common);
                        if (r >= r_nr) { kpe = kp; kde = kd; }
                        else if (dt * kp > kd && kp > 0.0) {
                            const double tau = kd / kp;
                            kpe = 1.0 / (dt * (dt + tau) * r_nr);
                            kde = tau * kpe;
                        } else if (kd > 0.0) {
                            const double tinv = kp / kd;
                            kde = 1.0 / (dt * (1.0 + dt * tinv) * r_nr);
                            kpe = tinv * kde;
                        } else {
                            kpe = 1.0 / (dt * dt * r_nr);
                            kde = 0.0;
                        }
                    }
                    const double gain = dt * kpe + kde;
                    if (gain > 0.0) {
                        double q0 = 0.0, qt = 0.0;
                        if (coord >= 0) {
                            q0 = q_in[coord];
                            qt = P.c_tpos ? P.c_tpos[gt] : 0.0f;
                        }
                        const double vd = P.c_tvel ? P.c_tvel[gt] : 0.0f;
                        const double u0 = P.c_act ? P.c_act[gt] : 0.0f;
                        pd = 1;
                        pda = (float)(kpe * (qt - q0) + kde * vd + u0);
                        pdg = (float)gain;
                        pdl = M.effort_limit[gt];
                    }
                }
            }
            if (coord >= 0 && T.dof_limit_ok[i]) {
                const double ke = M.limit_ke[gi];
                if (ke > 0.0) {
                    const double q0 = q_in[coord];
                    const double lower = M.limit_lower[gi], upper = M.limit_upper[gi];
                    const double window = kLimitWindow * dt *
                                          fmax(fabs(d64[V_V0 * npad + i]), fabs(d64[V_VS * npad + i]));
                    const double kd = M.limit_kd[gi];
                    const double tau_d = kd > 0.0 ? kd / ke : 0.0;
                    const double lbf = kLimitBeta * kLimitBeta / (4.0 * kPi * kPi);
                    const double r = fmax(lbf * fmax(ainv, 1.0e-12), 1.0 / (dt * ke * (dt + tau_d)));
                    if (isfinite(lower) && q0 - lower <= window) {
                        lo = 1; lov = (float)(-(q0 - lower) / (dt + tau_d)); lor = (float)r;
                    }
                    if (isfinite(upper) && upper - q0 <= window) {
                        up = 1; upv = (float)(-(upper - q0) / (dt + tau_d)); upr = (float)r;
                    }
                }
            }
            part = pd | lo | up;
            for (int c = 0; c < nc && !part; ++c) {
                for (int side = 0; side < 2; ++side) {
                    const int lb = s.cb[c * 2 + side];
                    if (lb >= 0) {
                        const int st = T.body_art_start[lb], ct = T.body_art_count[lb];
                        if (st >= 0 && ct > 0 && i >= st && i < st + ct) part = 1;
                    }
                }
                if (s.J[(c * 3) * js + i] != 0.0f || s.J[(c * 3 + 1) * js + i] != 0.0f ||
                    s.J[(c * 3 + 2) * js + i] != 0.0f)
                    part = 1;
            }
        }
        s.dflag[I_PART * npad + i] = part;
        s.dflag[I_PD * npad + i] = pd;
        s.dflag[I_LO * npad + i] = lo;
        s.dflag[I_UP * npad + i] = up;
        d32[F_PDA * npad + i] = pda;
        d32[F_PDG * npad + i] = pdg;
        d32[F_PDL * npad + i] = pdl;
        d32[F_LOV * npad + i] = lov;
        d32[F_UPV * npad + i] = upv;
        d32[F_LOR * npad + i] = lor;
        d32[F_UPR * npad + i] = upr;
    }
    __syncthreads();

    // ------------------------------------------------------------------ 8
    // SAP Newton solve (contact_solve.py:6129-6285, 5474-5655).  The mat-vecs (J v, A v, J^T
    // gamma, the refinement residual, the line-search data), the Hessian assembly and its
    // factorization use the whole block; projections, reductions, the triangular solves and
    // the line search run on wave 0 with wave-level ordering.  ctl[0] carries wave 0's
    // continue / stop decision to the other waves, ctl[1] the unconstrained flag.
    const int lane = tid & 63, wave = tid >> 6;
    int* ctl = reinterpret_cast<int*>(s.red);
    double* v = d64 + V_V * npad;
    double* vs = d64 + V_VS * npad;
    double* dvv = d64 + V_DV * npad;
    double* grad = d64 + V_G * npad;
    double* jimp = d64 + V_JI * npad;
    double* dp = d64 + V_DP * npad;
    double* av = d64 + V_AV * npad;
    int iteration = 0, ls_total = 0, flags = 0, ls_status = 0;  // meaningful on wave 0
    bool unconstrained = false;
    if (P.dbg_A) {
        for (int t = tid; t < n * n; t += NT) P.dbg_A[(size_t)w * n * n + t] = s.A[(t / n) * ast + (t % n)];
    }
    if (P.dbg_J) {
        const size_t jb = (size_t)w * C * 3 * n;
        for (int t = tid; t < C * 3 * n; t += NT) {
            const int r = t / n;
            P.dbg_J[jb + t] = (r < 3 * nc) ? (double)s.J[r * js + (t % n)] : 0.0;
        }
    }
    if (wave == 0) {
        int am = 0;
        for (int i = lane; i < n; i += 64)
            am |= s.dflag[I_PD * npad + i] | s.dflag[I_LO * npad + i] | s.dflag[I_UP * npad + i];
        unconstrained = (nc == 0) && !__any(am);
        for (int i = lane; i < npad; i += 64) {
            double vi = 0.0;
            if (i < n) {
                vi = P.use_guess ? P.v_guess[(size_t)w * n + i] : d64[V_V0 * npad + i];
                if (unconstrained || !s.dflag[I_PART * npad + i]) vi = vs[i];
            }
            v[i] = vi;
            dvv[i] = 0.0;
            d64[V_AVS * npad + i] = (i < n) ? dot_fd(s.A + i * ast, vs, n) : 0.0;
        }
        wave_sync();
    }
    double alpha = 1.0, prev_cost = 0.0, cost = 0.0;
    const double cost_min_alpha = 0.5;
    // the mat-vecs of each stage run on the whole workgroup (one row / dof / contact per thread,
    // each value computed as before); projections, reductions and the line search stay on wave 0
    if (wave == 0 && lane == 0) ctl[1] = unconstrained ? 1 : 0;
    __syncthreads();
    const bool free_world = ctl[1] != 0;
    while (true) {
        // ---- evaluate at v: vc = J v and A v, projection + model terms, then
        //      impulses / gradient / norms (contact_solve.py:743-921, 1493-1627)
        if (!free_world) {
            for (int t = tid; t < 3 * nc + n; t += NT) {
                if (t < 3 * nc) {
                    s.cv[(t / 3) * 9 + (t % 3)] = dot_fd(s.J + t * js, v, n);
                } else {
                    const int i = t - 3 * nc;
                    const double avi = dot_fd(s.A + i * ast, v, n);
                    av[i] = avi;
                    d64[V_AR * npad + i] = avi - d64[V_AVS * npad + i];
                }
            }
        }
        __syncthreads();
        double part3[4] = {0.0, 0.0, 0.0, 0.0};  // cost, |D g|^2, |D A v|^2, |D j|^2 (wave 0)
        if (wave == 0 && !free_world) {
            for (int t = lane; t < nc + n; t += 64) {
                if (t < nc) {
                    const int c = t;
                    double vc[3] = {s.cv[c * 9], s.cv[c * 9 + 1], s.cv[c * 9 + 2]};
                    double g[3];
                    part3[0] += project_contact(s.ck + c * kCk, vc, g, s.cG + c * 9, true);
                    s.cv[c * 9 + 6] = g[0]; s.cv[c * 9 + 7] = g[1]; s.cv[c * 9 + 8] = g[2];
                } else {
                    const int i = t - nc;
                    double imp, hd;
                    part3[0] += model_terms(s, npad, i, v[i], dt, &imp, &hd);
                    d32[F_HD * npad + i] = (float)hd;
                    dp[i] = imp;  // temporarily holds the model impulse
                }
            }
        }
        __syncthreads();
        if (!free_world) {
            for (int i = tid; i < n; i += NT) {
                double j0a = 0.0, j1a = 0.0, j2a = 0.0;
                for (int c = 0; c < nc; ++c) {
                    j0a += (double)s.J[(c * 3) * js + i] * s.cv[c * 9 + 6];
                    j1a += (double)s.J[(c * 3 + 1) * js + i] * s.cv[c * 9 + 7];
                    j2a += (double)s.J[(c * 3 + 2) * js + i] * s.cv[c * 9 + 8];
                }
                const double jc = dp[i] + ((j0a + j1a) + j2a);
                jimp[i] = jc;
                grad[i] = d64[V_AR * npad + i] - jc;
            }
        }
        __syncthreads();
        if (wave == 0) {
            bool stop = false;
            if (free_world) {
                flags = 1;
                stop = true;
            } else {
                for (int i = lane; i < n; i += 64) {
                    const double jc = jimp[i];
                    const double ar = d64[V_AR * npad + i];
                    const double avv = av[i];
                    part3[0] += 0.5 * (v[i] - vs[i]) * ar;
                    if (s.dflag[I_PART * npad + i]) {
                        const double sc = d32[F_DSC * npad + i];
                        part3[1] += (sc * grad[i]) * (sc * grad[i]);
                        part3[2] += (sc * avv) * (sc * avv);
                        part3[3] += (sc * jc) * (sc * jc);
                    }
                }
                wave_sync();
#pragma unroll
                for (int k = 0; k < 4; ++k) part3[k] = wave_sum(part3[k]);
                cost = part3[0];
                const double gnorm = sqrt(fmax(part3[1], 0.0));
                const double pnorm = sqrt(fmax(part3[2], 0.0));
                const double jnorm = sqrt(fmax(part3[3], 0.0));
                // ---- convergence (contact_solve.py:1970-2087)
                const double opt_tol = P.opt_abs + P.opt_rel * fmax(pnorm, jnorm);
                const bool opt = gnorm <= opt_tol;
                bool costok = false;
                if (iteration > 0 && alpha > cost_min_alpha) {
                    const double scale = 0.5 * (fabs(cost) + fabs(prev_cost));
                    if (fabs(cost - prev_cost) < P.cost_abs + P.cost_rel * scale) costok = true;
                }
                if (opt || costok) {
                    flags = 1 | (opt ? 2 : 0) | (costok ? 4 : 0);
                    stop = true;
                } else if (iteration >= P.max_iter) {
                    flags = 8;
                    stop = true;
                }
            }
            if (lane == 0) ctl[0] = stop ? 0 : 1;
        }
        __syncthreads();
        if (ctl[0] == 0) break;
        // ---- Hessian H = A + J^T G J + diag(hd) with MFMA (contact_solve.py:1334-1654)
        {
            constexpr int NWV = NT / 64;
            const int hl = tid & 63, hw = tid >> 6;
            const int tb = npad >> 4;
            const int ntile = tb * (tb + 1) / 2;
            const int K = 3 * nc;
            for (int tt = hw; tt < ntile; tt += NWV) {
                int ib = 0, rem = tt;
                while (rem > ib) { rem -= ib + 1; ++ib; }
                const int jb = rem;
                const int i0 = ib * 16, j0 = jb * 16;
                typedef float f4 __attribute__((ext_vector_type(4)));
                f4 acc = {0.0f, 0.0f, 0.0f, 0.0f};
                const int ia = i0 + (hl & 15);
                const int jbb = j0 + (hl & 15);
                // four k-steps' operands are loaded before their four MFMAs, so the loads (device
                // memory for large worlds) overlap instead of each MFMA waiting on its own
                for (int k0 = 0; k0 < K; k0 += 16) {
                    float a4[4], b4[4];
#pragma unroll
                    for (int u = 0; u < 4; ++u) {
                        const int k = k0 + 4 * u + (hl >> 4);
                        float a = 0.0f, b = 0.0f;
                        if (k < K) {
                            const int c = k / 3, r = k % 3;
                            if (ia < n) a = s.J[k * js + ia];
                            if (jbb < n) {
                                const float* G = s.cG + c * 9 + r * 3;
                                const float* Jc = s.J + (c * 3) * js + jbb;
                                b = G[0] * Jc[0] + G[1] * Jc[js] + G[2] * Jc[2 * js];
                            }
                        }
                        a4[u] = a;
                        b4[u] = b;
                    }
#pragma unroll
                    for (int u = 0; u < 4; ++u) acc = __builtin_amdgcn_mfma_f32_16x16x4f32(a4[u], b4[u], acc, 0, 0, 0);
                }
#pragma unroll
                for (int e = 0; e < 4; ++e) {
                    const int i = i0 + 4 * (hl >> 4) + e;
                    const int j = j0 + (hl & 15);
                    float val;
                    if (i < n && j < n) {
                        val = s.A[i * ast + j] + acc[e];
                        if (i == j) val += d32[F_HD * npad + i] + (float)P.diag_shift;
                    } else {
                        val = (i == j) ? 1.0f : 0.0f;
                    }
                    s.H[i * hs + j] = val;
                }
            }
        }
        __syncthreads();
        cholesky_blocked<NT>(s.H, npad, hs, n, d32 + F_DINV * npad);
        chol_block_inverses<NT>(s.H, npad, hs, d32 + F_DINV * npad);
        if (wave == 0) {
            for (int i = lane; i < npad; i += 64) dvv[i] = (i < n) ? -grad[i] : 0.0;
            wave_sync();
            chol_solve(s.H, npad, hs, n, d32 + F_DINV * npad, dvv);
        }
        __syncthreads();
        // ---- mixed-precision refinement of the Newton direction: residual
        //      r = -grad - (A + J^T G J + diag(hd + shift)) dv in float64 from the stored
        //      float32 operands, correction solved with the float32 factor
        for (int rep = 0; rep < kNewtonRefine; ++rep) {
            for (int t = tid; t < 3 * nc + n; t += NT) {
                if (t < 3 * nc) {
                    s.cv[(t / 3) * 9 + 3 + (t % 3)] = dot_fd(s.J + t * js, dvv, n);
                } else {
                    const int i = t - 3 * nc;
                    dp[i] = dot_fd(s.A + i * ast, dvv, n);
                }
            }
            __syncthreads();
            // G_c (J_c dv) once per contact, in place of J_c dv (recomputed before it is read again)
            for (int c = tid; c < nc; c += NT) {
                const float* G = s.cG + c * 9;
                const double t0 = s.cv[c * 9 + 3], t1 = s.cv[c * 9 + 4], t2 = s.cv[c * 9 + 5];
                s.cv[c * 9 + 3] = (double)G[0] * t0 + (double)G[1] * t1 + (double)G[2] * t2;
                s.cv[c * 9 + 4] = (double)G[3] * t0 + (double)G[4] * t1 + (double)G[5] * t2;
                s.cv[c * 9 + 5] = (double)G[6] * t0 + (double)G[7] * t1 + (double)G[8] * t2;
            }
            __syncthreads();
            for (int i = tid; i < npad; i += NT) {
                double r = 0.0;
                if (i < n) {
                    double acc = dp[i] + ((double)d32[F_HD * npad + i] + P.diag_shift) * dvv[i];
                    for (int c = 0; c < nc; ++c) {
                        const double u0 = s.cv[c * 9 + 3], u1 = s.cv[c * 9 + 4], u2 = s.cv[c * 9 + 5];
                        acc += (double)s.J[(c * 3) * js + i] * u0 + (double)s.J[(c * 3 + 1) * js + i] * u1 +
                               (double)s.J[(c * 3 + 2) * js + i] * u2;
                    }
                    r = -grad[i] - acc;
                }
                d64[V_T1 * npad + i] = r;
            }
            __syncthreads();
            if (wave == 0) {
                chol_solve(s.H, npad, hs, n, d32 + F_DINV * npad, d64 + V_T1 * npad);
                wave_sync();
                for (int i = lane; i < n; i += 64) dvv[i] += d64[V_T1 * npad + i];
            }
            __syncthreads();
        }
        // ---- line search data: dvc = J dv and dp = A dv (contact_solve.py:2234-2263)
        for (int t = tid; t < 3 * nc + n; t += NT) {
            if (t < 3 * nc) {
                s.cv[(t / 3) * 9 + 3 + (t % 3)] = dot_fd(s.J + t * js, dvv, n);
            } else {
                const int i = t - 3 * nc;
                dp[i] = dot_fd(s.A + i * ast, dvv, n);
            }
        }
        __syncthreads();
        if (wave == 0) {
            double ld[4] = {0, 0, 0, 0};  // dell0, dell_a0, d2ell_a, m0
            for (int t = lane; t < 3 * nc + n; t += 64) {
                if (t >= 3 * nc) {
                    const int i = t - 3 * nc;
                    const double acc = dp[i];
                    const double ri = v[i] - vs[i];
                    ld[0] += grad[i] * dvv[i];
                    ld[1] += acc * ri;
                    ld[2] += dvv[i] * acc;
                    ld[3] += 0.5 * ri * d64[V_AR * npad + i];
                }
            }
            wave_sync();
#pragma unroll
            for (int k = 0; k < 4; ++k) ld[k] = wave_sum(ld[k]);
            const double dell0 = ld[0], dell_a0 = ld[1], d2ell_a = ld[2], m0 = ld[3];
            // trial evaluation of phi(a), phi'(a), phi''(a)
            auto trial = [&](double a, bool second, double* d1, double* d2) -> double {
                double tv[3] = {0.0, 0.0, 0.0};
                for (int c = lane; c < nc; c += 64) {
                    double vc[3], dvc[3], g[3], cc;
                    for (int r = 0; r < 3; ++r) {
                        dvc[r] = s.cv[c * 9 + 3 + r];
                        vc[r] = s.cv[c * 9 + r] + a * dvc[r];
                    }
                    if (second) {
                        tv[2] += project_curvature(s.ck + c * kCk, vc, dvc, g, &cc);
                    } else {
                        float Gd[9];
                        cc = project_contact(s.ck + c * kCk, vc, g, Gd, false);
                    }
                    tv[0] += cc;
                    tv[1] -= dvc[0] * g[0] + dvc[1] * g[1] + dvc[2] * g[2];
                }
                for (int i = lane; i < n; i += 64) {
                    double imp, hd;
                    tv[0] += model_terms(s, npad, i, v[i] + a * dvv[i], dt, &imp, &hd);
                    tv[1] -= dvv[i] * imp;
                    tv[2] += hd * dvv[i] * dvv[i];
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) tv[k] = wave_sum(tv[k]);
                *d1 = dell_a0 + a * d2ell_a + tv[1];
                *d2 = d2ell_a + tv[2];
                return m0 + a * dell_a0 + 0.5 * a * a * d2ell_a + tv[0];
            };
            bool accepted = false;
            double a_acc = 1.0;
            int ls_it = 0;
            ls_status = 0;
            if (!(dell0 < 0.0) || !isfinite(dell0)) {
                ls_status = -1;
            } else if (P.ls_mode == 0) {
                // Armijo backtracking (contact_solve.py:2670-2905, 5108-5305)
                const double amax = 1.0 / P.rho;
                double a = amax, d1, d2;
                double ell = trial(a, false, &d1, &d2);
                const double slop = (P.rel_slop / 10.0) * fmax(1.0, 0.5 * (fabs(ell) + fabs(cost)));
                if (d1 < 0.0 || d1 < slop) {
                    accepted = true;
                    a_acc = a;
                } else if (P.ls_max_iter > 1) {
                    double a_prev = a, e_prev = ell;
                    int it = 1;
                    while (true) {
                        a = a * P.rho;
                        ell = trial(a, false, &d1, &d2);
                        const double den = a - a_prev;
                        const double dapprox = fabs(den) > 0.0 ? (ell - e_prev) / den : 0.0;
                        if (fabs(dapprox) < slop) { accepted = true; a_acc = a; ls_it = it; break; }
                        if (ell > e_prev && ell < cost + P.armijo_c * a * dell0) {
                            accepted = true;
                            ls_it = it;
                            a_acc = (e_prev < cost + P.armijo_c * a_prev * dell0) ? a_prev : a;
                            break;
                        }
                        if (it >= P.ls_max_iter - 1) {
                            if (ell < cost + P.armijo_c * a * dell0) { accepted = true; a_acc = a; ls_it = P.ls_max_iter; }
                            else ls_status = -2;
                            break;
                        }
                        a_prev = a;
                        e_prev = ell;
                        ++it;
                    }
                } else {
                    ls_status = -2;
                }
            } else {
                // exact root (contact_solve.py:2913-3140, 5700-5908), cubic Hermite seed
                const double amax = kExactAlphaMax;
                double d1, d2;
                double ell = trial(amax, true, &d1, &d2);
                if (d1 <= 0.0) {
                    accepted = true; a_acc = amax;
                } else if (-dell0 < P.cost_abs + P.cost_rel * ell) {
                    accepted = true; a_acc = 1.0;
                } else {
                    const double scale = -dell0;
                    const double f_upper0 = d1 / scale;
                    if (!isfinite(f_upper0) || f_upper0 <= 0.0) {
                        ls_status = -3;
                    } else {
                        double guess = -dell0 / d2;
                        const double hh = amax;
                        const double theta = dell0 + d1 - 3.0 * (ell - cost) / hh;
                        const double disc = theta * theta - dell0 * d1;
                        if (isfinite(disc) && disc >= 0.0) {
                            const double gm = sqrt(disc);
                            const double den = d1 - dell0 + 2.0 * gm;
                            if (den != 0.0) {
                                const double cubic = hh - hh * (d1 + gm - theta) / den;
                                if (isfinite(cubic) && cubic > 0.0 && cubic <= hh) guess = cubic;
                            }
                        }
                        if (guess > amax) guess = amax;
                        if (!isfinite(guess) || guess < 0.0 || guess > amax || kExactFTol * guess <= 0.0) {
                            ls_status = -4;
                        } else {
                            double x_lo = 0.0, x_hi = amax, f_hi = f_upper0, root = guess;
                            double mdx = -amax, mdx_prev = -amax;
                            const double x_tol = kExactFTol * guess;
                            int it = 1;
                            while (true) {
                                trial(root, true, &d1, &d2);
                                const double f = d1 / scale, df = d2 / scale;
                                if ((f < 0.0 && f_hi >= 0.0) || (f >= 0.0 && f_hi < 0.0)) {
                                    x_lo = root;
                                } else {
                                    x_hi = root;
                                    f_hi = f;
                                }
                                if (fabs(f) < kExactFTol) { accepted = true; a_acc = root; ls_it = it; break; }
                                if (it >= P.ls_max_iter) { ls_status = -5; break; }
                                const bool slow = 2.0 * fabs(f) > fabs(mdx_prev * df);
                                mdx_prev = mdx;
                                const double bdx = 0.5 * (x_lo - x_hi);
                                const double broot = x_lo - bdx;
                                double ndx = 0.0, nroot = root;
                                if (fabs(df) > 0.0) { ndx = f / df; nroot = root - ndx; }
                                const bool bis = slow || nroot < x_lo || nroot > x_hi || !isfinite(nroot) || !isfinite(ndx);
                                const double nxt = bis ? broot : nroot;
                                const double nd = bis ? bdx : ndx;
                                if (fabs(nd) < x_tol) { accepted = true; a_acc = nxt; ls_it = it; break; }
                                root = nxt;
                                mdx = nd;
                                ++it;
                            }
                        }
                    }
                }
            }
            // ---- commit (contact_solve.py:3154-3176)
            if (accepted) {
                alpha = a_acc;
                prev_cost = cost;
                ls_total += ls_it;
                for (int i = lane; i < n; i += 64) v[i] += alpha * dvv[i];
            }
            wave_sync();
            ++iteration;
        }
        __syncthreads();  // the committed v before the next evaluation's mat-vecs
    }
    __syncthreads();

    // ------------------------------------------------------------------ 9
    // integration (solver_sap.py:288-519) + SAP -> public velocity
    float* qd_pub = kQdPub;  // reuse: public output velocity
    for (int k = tid; k < M.NQ; k += NT) q_out[k] = q_in[k];
    __syncthreads();
    for (int j = tid; j < nj; j += NT) {
        const int jt = T.jtype[j];
        const int qs = T.jqs[j], ds = T.jds[j], nd = T.jds[j + 1] - T.jds[j];
        if (jt == kJointFree || jt == kJointDistance || jt == kJointBall) {
            // free: (p, quaternion) with [w, v_O]; ball: quaternion with w (solver_sap.py:288-519)
            const bool ball = jt == kJointBall;
            const int qr = ball ? qs : qs + 3;
            Q4 q = Q4{q_in[qr], q_in[qr + 1], q_in[qr + 2], q_in[qr + 3]};
            V3 wn = v3((float)v[ds], (float)v[ds + 1], (float)v[ds + 2]);
            V3 von = ball ? v3(0, 0, 0) : v3((float)v[ds + 3], (float)v[ds + 4], (float)v[ds + 5]);
            const double* v0 = d64 + V_V0 * npad + ds;
            V3 wo = v3((float)v0[0], (float)v0[1], (float)v0[2]);
            V3 voo = ball ? v3(0, 0, 0) : v3((float)v0[3], (float)v0[4], (float)v0[5]);
            Q4 qn;
            V3 vmid;
            if (P.integ_mode == 0) {
                V3 wm = (wo + wn) * 0.5f;
                float ang = length(wm) * h;
                qn = q;
                if (ang > 1.0e-12f) {
                    V3 ax = wm * (1.0f / length(wm));
                    float sa = sinf(0.5f * ang), ca = cosf(0.5f * ang);
                    Q4 r = Q4{ax.x * sa, ax.y * sa, ax.z * sa, ca};
                    Q4 o = qmul(r, q);
                    float nn = sqrtf(o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w);
                    if (nn > 0.0f) { float inv = 1.0f / nn; o = Q4{o.x * inv, o.y * inv, o.z * inv, o.w * inv}; }
                    qn = o;
                }
                vmid = (voo + von) * 0.5f;
            } else {
                float hd = 0.5f * h;
                float x = q.x + hd * (q.w * wn.x + wn.y * q.z - wn.z * q.y);
                float y = q.y + hd * (q.w * wn.y + wn.z * q.x - wn.x * q.z);
                float z = q.z + hd * (q.w * wn.z + wn.x * q.y - wn.y * q.x);
                float sw = q.w - hd * (wn.x * q.x + wn.y * q.y + wn.z * q.z);
                float nn = sqrtf(x * x + y * y + z * z + sw * sw);
                if (nn > 0.0f) { float inv = 1.0f / nn; x *= inv; y *= inv; z *= inv; sw *= inv; }
                qn = Q4{x, y, z, sw};
                vmid = von;
            }
            if (ball) {
                q_out[qs] = qn.x; q_out[qs + 1] = qn.y; q_out[qs + 2] = qn.z; q_out[qs + 3] = qn.w;
                qd_pub[ds] = wn.x; qd_pub[ds + 1] = wn.y; qd_pub[ds + 2] = wn.z;
                continue;
            }
            q_out[qs] = q_in[qs] + h * vmid.x;
            q_out[qs + 1] = q_in[qs + 1] + h * vmid.y;
            q_out[qs + 2] = q_in[qs + 2] + h * vmid.z;
            q_out[qs + 3] = qn.x; q_out[qs + 4] = qn.y; q_out[qs + 5] = qn.z; q_out[qs + 6] = qn.w;
            const size_t gb = (size_t)w * nb + T.jchild[j];
            const float* com = mCom + T.jchild[j] * 4;
            V3 rc = qrot(qnormalize(qn), v3(com[0], com[1], com[2]));
            V3 vcom = von + cross(wn, rc);
            qd_pub[ds] = vcom.x; qd_pub[ds + 1] = vcom.y; qd_pub[ds + 2] = vcom.z;
            qd_pub[ds + 3] = wn.x; qd_pub[ds + 4] = wn.y; qd_pub[ds + 5] = wn.z;
        } else if (jt == kJointFixed) {
            for (int a = 0; a < nd; ++a) qd_pub[ds + a] = (float)v[ds + a];
        } else {
            for (int a = 0; a < nd; ++a) {
                const float qd = (float)v[ds + a];
                const float qm = P.integ_mode == 0 ? 0.5f * ((float)d64[V_V0 * npad + ds + a] + qd) : qd;
                q_out[qs + a] = q_in[qs + a] + h * qm;
                qd_pub[ds + a] = qd;
            }
        }
    }
    __syncthreads();
    // public forward kinematics (sap_kinematics.py:84-229): joint-local transform
    // X_j X_c^-1 and joint velocity in parallel, then the wave-0 tree walk
    float* pbq = s.bq;     // reuse: public poses
    float* pbqd = kV;      // reuse: public velocities (nb x 6, [v, w])
    float* kPl = reinterpret_cast<float*>(kXl);  // nj x 16: X_j X_c^-1 (7), joint lin (3), ang (3)
    for (int j = tid; j < nj; j += NT) {
        const int jt = T.jtype[j];
        const int qs = T.jqs[j], ds = T.jds[j];
        TF Xj = TF{v3(0, 0, 0), Q4{0, 0, 0, 1}};
        V3 jl = v3(0, 0, 0), ja = v3(0, 0, 0);
        const float* axp = mAx + ds * 4;
        const V3 axis = v3(axp[0], axp[1], axp[2]);
        if (jt == kJointPrismatic) {
            Xj.p = axis * q_out[qs];
            jl = axis * qd_pub[ds];
        } else if (jt == kJointRevolute) {
            Xj.q = q_axis_angle(axis, q_out[qs]);
            ja = axis * qd_pub[ds];
        } else if (jt == kJointFree || jt == kJointDistance) {
            Xj = TF{v3(q_out[qs], q_out[qs + 1], q_out[qs + 2]), Q4{q_out[qs + 3], q_out[qs + 4], q_out[qs + 5], q_out[qs + 6]}};
            jl = v3(qd_pub[ds], qd_pub[ds + 1], qd_pub[ds + 2]);
            ja = v3(qd_pub[ds + 3], qd_pub[ds + 4], qd_pub[ds + 5]);
        } else if (jt == kJointBall) {
            Xj.q = Q4{q_out[qs], q_out[qs + 1], q_out[qs + 2], q_out[qs + 3]};
            ja = v3(qd_pub[ds], qd_pub[ds + 1], qd_pub[ds + 2]);
        } else if (jt == kJointD6) {
            // sap_kinematics.py:170-218 (the rotation in float64, rounded once)
            const int nl = T.jlin[j], last = T.jds[j + 1] - ds - 1;
            D3 c[3];
            double qa[3];
#pragma unroll
            for (int k = 0; k < 3; ++k) {
                const int kl = min(k, last), ka = min(nl + k, last);
                const float* al = mAx + (ds + kl) * 4;
                const float on = k < nl ? 1.0f : 0.0f;
                Xj.p = Xj.p + v3(al[0], al[1], al[2]) * (on * q_out[qs + kl]);
                jl = jl + v3(al[0], al[1], al[2]) * (on * qd_pub[ds + kl]);
                const float* aa = mAx + (ds + ka) * 4;
                c[k] = d3(aa[0], aa[1], aa[2]);
                qa[k] = (double)q_out[qs + ka];
            }
            const int na = last + 1 - nl;
            const D6Ang r = d6_angular(na, c[0], c[1], c[2], qa[0], qa[1], qa[2]);
            Xj.q = Q4{(float)r.rot.x, (float)r.rot.y, (float)r.rot.z, (float)r.rot.w};
            const float w0 = na > 0 ? qd_pub[ds + min(nl, last)] : 0.0f;
            const float w1 = na > 1 ? qd_pub[ds + min(nl + 1, last)] : 0.0f;
            const float w2 = na > 2 ? qd_pub[ds + min(nl + 2, last)] : 0.0f;
            ja = v3((float)r.a0.x, (float)r.a0.y, (float)r.a0.z) * w0 + v3((float)r.a1.x, (float)r.a1.y, (float)r.a1.z) * w1 +
                 v3((float)r.a2.x, (float)r.a2.y, (float)r.a2.z) * w2;
        }
        const TF Xjc = tf_mul(Xj, tf_inverse(load_tf(mXc + j * 8)));
        float* o = kPl + j * 16;
        store_tf(o, Xjc);
        o[7] = jl.x; o[8] = jl.y; o[9] = jl.z; o[10] = ja.x; o[11] = ja.y; o[12] = ja.z;
    }
    __syncthreads();
    if (tid < 64) {
        const int lane = tid;
        for (int L = 0; L < M.num_levels; ++L) {
            const int l0 = T.level_start[L], l1 = T.level_start[L + 1];
            for (int t = l0 + lane; t < l1; t += 64) {
                const int j = T.level_joint[t];
                const int parent = T.jparent[j], child = T.jchild[j];
                const TF Xp = load_tf(mXp + j * 8);
                TF Xwpj = Xp;
                V3 vlin = v3(0, 0, 0), vang = v3(0, 0, 0);
                if (parent >= 0) {
                    const TF Xwp = load_tf(pbq + parent * 8);
                    Xwpj = tf_mul(Xwp, Xp);
                    const float* com = mCom + parent * 4;
                    const V3 rp = Xwpj.p - (Xwp.p + qrot(Xwp.q, v3(com[0], com[1], com[2])));
                    const float* vp = pbqd + parent * 6;
                    vang = v3(vp[3], vp[4], vp[5]);
                    vlin = v3(vp[0], vp[1], vp[2]) + cross(vang, rp);
                }
                const float* pl = kPl + j * 16;
                const TF Xwc = tf_mul(Xwpj, load_tf(pl));
                float* o = pbq + child * 8;
                o[0] = Xwc.p.x; o[1] = Xwc.p.y; o[2] = Xwc.p.z; o[3] = Xwc.q.x; o[4] = Xwc.q.y; o[5] = Xwc.q.z; o[6] = Xwc.q.w;
                const V3 lv = vlin + qrot(Xwpj.q, v3(pl[7], pl[8], pl[9]));
                const V3 av3 = vang + qrot(Xwpj.q, v3(pl[10], pl[11], pl[12]));
                float* ov = pbqd + child * 6;
                ov[0] = lv.x; ov[1] = lv.y; ov[2] = lv.z; ov[3] = av3.x; ov[4] = av3.y; ov[5] = av3.z;
            }
            __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
            __builtin_amdgcn_wave_barrier();
            __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
        }
    }
    __syncthreads();
    // ------------------------------------------------------------------ outputs
    for (int k = tid; k < M.NQ; k += NT) P.jq_out[(size_t)w * M.NQ + k] = q_out[k];
    for (int d = tid; d < n; d += NT) {
        const size_t g = (size_t)w * n + d;
        P.jqd_out[g] = qd_pub[d];
        P.v_guess[g] = v[d];
        if (P.dbg_v) P.dbg_v[g] = v[d];
        if (P.dbg_vstar) P.dbg_vstar[g] = vs[d];
        if (P.dbg_v0) P.dbg_v0[g] = d64[V_V0 * npad + d];
        if (P.dbg_adiag) P.dbg_adiag[g] = s.A[d * ast + d];
    }
    for (int t = tid; t < nb * 7; t += NT) {
        const int b = t / 7, k = t % 7;
        P.bq_out[((size_t)w * nb + b) * 7 + k] = pbq[b * 8 + k];
    }
    if (P.bqd_out)
        for (int t = tid; t < nb * 6; t += NT) P.bqd_out[(size_t)w * nb * 6 + t] = pbqd[t];
    if (P.bf_out && P.bf_in && P.bf_out != P.bf_in)
        for (int t = tid; t < nb * 6; t += NT) P.bf_out[(size_t)w * nb * 6 + t] = P.bf_in[(size_t)w * nb * 6 + t];
    if (P.dbg_gamma)
        for (int t = tid; t < C * 3; t += NT) {
            const int c = t / 3, r = t % 3;
            P.dbg_gamma[(size_t)w * C * 3 + t] = (c < nc) ? s.cv[c * 9 + 6 + r] : 0.0;
        }
    if (tid == 0) {
        P.st_iters[w] = iteration;
        P.st_ls[w] = ls_total;
        P.st_nc[w] = nc;
        P.st_flags[w] = flags | ((ls_status & 0xff) << 8);
        if (P.st_hist) {
            atomicAdd(&P.st_hist[iteration < 255 ? iteration : 255], 1ull);
            atomicAdd(&P.st_hist[256], (unsigned long long)ls_total);
            atomicAdd(&P.st_hist[257], 1ull);
        }
    }
}

}  // namespace sapx

// host-side launcher (C++ linkage inside the library)
namespace sapx {
hipError_t launch_step(const StepParams& P, hipStream_t stream) {
    constexpr int NT = 256;
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)sapx_step_kernel<NT, false>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        (void)hipFuncSetAttribute((const void*)sapx_step_kernel<NT, true>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr_set = true;
    }
    if (P.g_work)
        hipLaunchKernelGGL((sapx_step_kernel<NT, true>), dim3(P.m.W), dim3(NT), P.smem_bytes, stream, P);
    else
        hipLaunchKernelGGL((sapx_step_kernel<NT, false>), dim3(P.m.W), dim3(NT), P.smem_bytes, stream, P);
    return hipGetLastError();
}
}  // namespace sapx

// Device helpers shared by the step kernels (gfx950): SAP constants, wave-level
// reductions and broadcasts, spatial-vector algebra, the topology template view,
// the SAP projection onto the regularised friction cone and scalar clamps.
#ifndef SAPX_KCOMMON_H
#define SAPX_KCOMMON_H

#include "sapx_device.h"
#include "sapx_params.h"

namespace sapx {

constexpr double kPi = 3.141592653589793;
constexpr double kSoftNormTol = 1.0e-7;   // contact_solve.py:47
constexpr double kPdBeta = 0.1;           // contact_solve.py:48
constexpr double kLimitBeta = 0.1;        // contact_solve.py:49
constexpr double kLimitWindow = 2.0;      // contact_solve.py:51
constexpr double kExactAlphaMax = 1.5;    // contact_solve.py:52
constexpr double kExactFTol = 1.0e-8;     // contact_solve.py:54
constexpr int kLsMonotone = 2;            // SAPX_LS_MONOTONE_DECAY
constexpr double kMonoDecay = 0.5;        // monotone_decay constants (contact_solve.py:5431-5443)
constexpr double kMonoMinAlpha = 1.0e-8;
constexpr double kMonoRelaxR = 1.0e-12;
constexpr double kMonoRelaxA = 1.0e-14;
constexpr int kNewtonRefine = 1;          // float64-residual refinement steps per Newton direction
constexpr int kCk = 8;                    // doubles of per-contact projection constants


// ---------------------------------------------------------------------------
// block utilities
// ---------------------------------------------------------------------------
// double-precision DPP move (two 32-bit halves); lanes outside ROW_MASK keep `old`
template <int CTRL, int ROW_MASK = 0xF>
__device__ __forceinline__ double dpp_d(double old, double x) {
    const long long xi = __double_as_longlong(x), oi = __double_as_longlong(old);
    const int lo = __builtin_amdgcn_update_dpp((int)oi, (int)xi, CTRL, ROW_MASK, 0xF, false);
    const int hi = __builtin_amdgcn_update_dpp((int)(oi >> 32), (int)(xi >> 32), CTRL, ROW_MASK, 0xF, false);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
__device__ __forceinline__ double readlane_d(double x, int l) {
    const long long xi = __double_as_longlong(x);
    const int lo = __builtin_amdgcn_readlane((int)xi, l);
    const int hi = __builtin_amdgcn_readlane((int)(xi >> 32), l);
    return __longlong_as_double(((long long)hi << 32) | (unsigned int)lo);
}
// sum over the 64 lanes of a fully active wave, returned uniformly: quad / half-row /
// row butterflies, then row_bcast:15 and row_bcast:31 carry row sums up to lane 63
__device__ __forceinline__ double wave_sum(double x) {
    x += dpp_d<0xB1>(0.0, x);        // quad_perm [1,0,3,2]
    x += dpp_d<0x4E>(0.0, x);        // quad_perm [2,3,0,1]
    x += dpp_d<0x141>(0.0, x);       // row_half_mirror
    x += dpp_d<0x140>(0.0, x);       // row_mirror
    x += dpp_d<0x142, 0xA>(0.0, x);  // row_bcast:15 -> rows 1, 3
    x += dpp_d<0x143, 0xC>(0.0, x);  // row_bcast:31 -> rows 2, 3
    return readlane_d(x, 63);
}

// order this wave's LDS writes before its later LDS reads from other lanes
__device__ __forceinline__ void wave_sync() {
    __builtin_amdgcn_fence(__ATOMIC_RELEASE, "wavefront");
    __builtin_amdgcn_wave_barrier();
    __builtin_amdgcn_fence(__ATOMIC_ACQUIRE, "wavefront");
}

// Block-wide sums of K <= 8 values.  Wave partials go to one half of `red`
// (alternating with `parity`), so a single barrier suffices: the half written
// by call k is rewritten only at call k + 2, after every thread has passed the
// barrier of call k + 1 and therefore finished reading call k's partials.
template <int NT>
__device__ __forceinline__ void block_sum(double* vals, int K, double* red, int& parity) {
    constexpr int NW = NT / 64;
    const int lane = threadIdx.x & 63;
    const int wave = threadIdx.x >> 6;
    double* half = red + parity * (NW * 8);
    for (int k = 0; k < K; ++k) {
        const double x = wave_sum(vals[k]);
        if (lane == 0) half[wave * 8 + k] = x;
    }
    __syncthreads();
    for (int k = 0; k < K; ++k) {
        double sum = 0.0;
#pragma unroll
        for (int wv = 0; wv < NW; ++wv) sum += half[wv * 8 + k];
        vals[k] = sum;
    }
    parity ^= 1;
}

template <int NT>
__device__ __forceinline__ int block_or(int x, double* red, int& parity) {
    double v[1] = {x ? 1.0 : 0.0};
    block_sum<NT>(v, 1, red, parity);
    return v[0] > 0.0;
}

__device__ __forceinline__ bool bit(unsigned long long mask, int i) { return (mask >> i) & 1ull; }

// SAP spatial helpers (angular-first), float32
struct SV {
    V3 w, v;
};
__device__ __forceinline__ SV sv(V3 w, V3 v) { return SV{w, v}; }
__device__ __forceinline__ SV load_sv(const float* a) { return SV{v3(a[0], a[1], a[2]), v3(a[3], a[4], a[5])}; }
__device__ __forceinline__ void store_sv(float* a, SV s) {
    a[0] = s.w.x; a[1] = s.w.y; a[2] = s.w.z; a[3] = s.v.x; a[4] = s.v.y; a[5] = s.v.z;
}
__device__ __forceinline__ SV sv_scale(SV a, float s) { return SV{a.w * s, a.v * s}; }
__device__ __forceinline__ SV sv_add(SV a, SV b) { return SV{a.w + b.w, a.v + b.v}; }
__device__ __forceinline__ float sv_dot(SV a, SV b) { return dot(a.w, b.w) + dot(a.v, b.v); }
// sap_helpers.py:948-994
__device__ __forceinline__ SV shift_velocity(SV V, V3 p) { return SV{V.w, V.v + cross(V.w, p)}; }
__device__ __forceinline__ SV compose_velocity(SV Vp, V3 p, SV Vpb) {
    return SV{Vp.w + Vpb.w, Vp.v + cross(Vp.w, p) + Vpb.v};
}
__device__ __forceinline__ SV compose_acceleration(SV Ap, SV Vp, V3 p, SV Vpb, SV Apb) {
    V3 w = Vp.w;
    V3 a_shift = Ap.v + cross(Ap.w, p) + cross(w, cross(w, p));
    return SV{Ap.w + Apb.w + cross(w, Vpb.w), a_shift + Apb.v + cross(w, Vpb.v) * 2.0f};
}
// mobilizer velocity -> body origin (sap_helpers.py:997-1024)
__device__ __forceinline__ SV mob_to_body(Q4 R, V3 p, SV Vf) {
    return SV{qrot(R, Vf.w), qrot(R, Vf.v + cross(Vf.w, p))};
}
__device__ __forceinline__ SV mob_bias_to_body(Q4 R, V3 p, SV Vf) {
    V3 w = Vf.w;
    return SV{v3(0.0f, 0.0f, 0.0f), qrot(R, cross(w, cross(w, p)))};
}

// ---------------------------------------------------------------------------
// wave-level broadcast of lane l's value (v_readlane: VALU -> SGPR, no LDS trip)
// ---------------------------------------------------------------------------
__device__ __forceinline__ float rl(float x, int l) {
    return __int_as_float(__builtin_amdgcn_readlane(__float_as_int(x), l));
}

// LDS copy of the per-world topology template (identical for every world)
struct Topo {
    const int* jtype;
    const int* jparent;
    const int* jchild;
    const int* jqs;
    const int* jds;
    const int* level_start;
    const int* level_joint;
    const int* dof_joint;
    const int* dof_coord;
    const int* dof_target;
    const int* dof_limit_ok;
    const int* body_art_start;
    const int* body_art_count;
    const int* child_start;
    const int* child;
    const unsigned long long* body_anc;      // NB: bit j = joint j on the body's root path
    const unsigned long long* body_anc_dof;  // NB: bit d = dof d's joint on the body's root path
    const unsigned long long* dof_anc;       // N: body_anc_dof of the child body of dof d's joint
    const int* body_up;                      // fk_rounds x NB: ancestor body 2^r levels up, -1 past the root
    const int* body_root;                    // NB: root body of the body's tree
    const int* jlin;                         // J: a D6 joint's linear axis count (0 otherwise)
};

__device__ __forceinline__ Topo topo_view(const int* t, const DevModel& M) {
    Topo T;
    T.jtype = t + M.topo_off[TP_JTYPE];
    T.jparent = t + M.topo_off[TP_JPARENT];
    T.jchild = t + M.topo_off[TP_JCHILD];
    T.jqs = t + M.topo_off[TP_JQS];
    T.jds = t + M.topo_off[TP_JDS];
    T.level_start = t + M.topo_off[TP_LEVEL_START];
    T.level_joint = t + M.topo_off[TP_LEVEL_JOINT];
    T.dof_joint = t + M.topo_off[TP_DOF_JOINT];
    T.dof_coord = t + M.topo_off[TP_DOF_COORD];
    T.dof_target = t + M.topo_off[TP_DOF_TARGET];
    T.dof_limit_ok = t + M.topo_off[TP_DOF_LIMIT_OK];
    T.body_art_start = t + M.topo_off[TP_BODY_ART_START];
    T.body_art_count = t + M.topo_off[TP_BODY_ART_COUNT];
    T.child_start = t + M.topo_off[TP_CHILD_START];
    T.child = t + M.topo_off[TP_CHILD];
    T.body_anc = reinterpret_cast<const unsigned long long*>(t + M.topo_off[TP_BODY_ANC]);
    T.body_anc_dof = reinterpret_cast<const unsigned long long*>(t + M.topo_off[TP_BODY_ANC_DOF]);
    T.dof_anc = reinterpret_cast<const unsigned long long*>(t + M.topo_off[TP_DOF_ANC]);
    T.body_up = t + M.topo_off[TP_BODY_UP];
    T.jlin = t + M.topo_off[TP_JLIN];
    T.body_root = t + M.topo_off[TP_BODY_ROOT];
    return T;
}

// ---------------------------------------------------------------------------
// SAP projection (contact_solve.py:743-921); returns cost, writes gamma and G
// ---------------------------------------------------------------------------
__device__ __forceinline__ double project_contact(const double* k, const double vc[3], double g[3], float* G,
                                                  bool want_G) {
    const double rt = k[0], rn = k[1], vhat = k[2], mu = k[3];
    const double rt_inv = k[4], rn_inv = k[5];
    const double y0 = -rt_inv * vc[0], y1 = -rt_inv * vc[1], y2 = rn_inv * (vhat - vc[2]);
    // the soft norm keeps yr >= 1e-7 > 0; one reciprocal square root gives both yr and 1 / yr
    const double yr2 = y0 * y0 + y1 * y1 + kSoftNormTol * kSoftNormTol;
    const double yr_inv = rsqrt(yr2);
    const double yr = yr2 * yr_inv;
    const double tx = y0 * yr_inv, ty = y1 * yr_inv;
    g[0] = g[1] = g[2] = 0.0;
    float Gm[9] = {0, 0, 0, 0, 0, 0, 0, 0, 0};
    if (mu <= 1.0e-12) {
        if (y2 > 0.0) {
            g[2] = y2;
            Gm[8] = (float)rn_inv;
        }
    } else {
        const double mu_hat = k[6], f = k[7];
        if (yr <= mu * y2) {
            g[0] = y0; g[1] = y1; g[2] = y2;
            Gm[0] = (float)rt_inv; Gm[4] = (float)rt_inv; Gm[8] = (float)rn_inv;
        } else if ((-mu_hat * yr < y2) && (y2 * mu < yr)) {
            const double gn = (y2 + mu_hat * yr) * f;
            g[0] = mu * gn * tx; g[1] = mu * gn * ty; g[2] = gn;
            if (want_G) {
                const double p00 = tx * tx, p01 = tx * ty, p11 = ty * ty;
                const double gy = gn * yr_inv;
                const double mf = mu_hat * f;
                Gm[0] = (float)(mu * (gy * (1.0 - p00) + mf * p00) * rt_inv);
                Gm[1] = (float)(mu * (gy * (-p01) + mf * p01) * rt_inv);
                Gm[2] = (float)(mu * f * tx * rn_inv);
                Gm[3] = Gm[1];
                Gm[4] = (float)(mu * (gy * (1.0 - p11) + mf * p11) * rt_inv);
                Gm[5] = (float)(mu * f * ty * rn_inv);
                Gm[6] = (float)(mf * tx * rt_inv);
                Gm[7] = (float)(mf * ty * rt_inv);
                Gm[8] = (float)(f * rn_inv);
            }
        }
    }
    if (want_G) {
#pragma unroll
        for (int e = 0; e < 9; ++e) G[e] = Gm[e];
    }
    return 0.5 * (rt * (g[0] * g[0] + g[1] * g[1]) + rn * g[2] * g[2]);
}

// contact-space curvature dvc^T G dvc with G recomputed at vc (exact line search)
__device__ __forceinline__ double project_curvature(const double* k, const double vc[3], const double dvc[3],
                                                    double g[3], double* cost) {
    float G[9];
    *cost = project_contact(k, vc, g, G, true);
    double Gd0 = (double)G[0] * dvc[0] + (double)G[1] * dvc[1] + (double)G[2] * dvc[2];
    double Gd1 = (double)G[3] * dvc[0] + (double)G[4] * dvc[1] + (double)G[5] * dvc[2];
    double Gd2 = (double)G[6] * dvc[0] + (double)G[7] * dvc[1] + (double)G[8] * dvc[2];
    return dvc[0] * Gd0 + dvc[1] * Gd1 + dvc[2] * Gd2;
}

// sum_d a[d] * x[d] for float32 row a and float64 vector x (4 independent accumulators)
__device__ __forceinline__ double dot_fd(const float* a, const double* x, int n) {
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
    int d = 0;
    for (; d + 4 <= n; d += 4) {
        s0 += (double)a[d] * x[d];
        s1 += (double)a[d + 1] * x[d + 1];
        s2 += (double)a[d + 2] * x[d + 2];
        s3 += (double)a[d + 3] * x[d + 3];
    }
    for (; d < n; ++d) s0 += (double)a[d] * x[d];
    return (s0 + s1) + (s2 + s3);
}

// PD / limit helpers (sap_helpers.py:2251-2340)
__device__ __forceinline__ double clamp_s(double x, double lim) {
    if (lim <= 0.0 || !isfinite(lim)) return x;
    return fmax(-lim, fmin(lim, x));
}
__device__ __forceinline__ double clamp_d(double x, double lim) {
    if (lim <= 0.0 || !isfinite(lim)) return 1.0;
    return (-lim <= x && x <= lim) ? 1.0 : 0.0;
}
__device__ __forceinline__ double clamp_a(double x, double lim) {
    if (lim <= 0.0 || !isfinite(lim)) return 0.5 * x * x;
    if (x < -lim) return -lim * (x + 0.5 * lim);
    if (x <= lim) return 0.5 * x * x;
    return lim * (x - 0.5 * lim);
}

// model (PD + limit) terms of dof i at velocity vi: returns cost, adds impulse and hdiag
}  // namespace sapx

#endif  // SAPX_KCOMMON_H

// Input features of the learned warm start (gfx950): one wave per world.
//
// No counterpart in the reference.  Per world, in this order (all float32):
//   free / distance joints: height, gravity direction, angular and linear velocity,
//     the last three in the joint's body frame (yaw- and translation-invariant);
//   other joints, per dof: q, qd, target_pos - q, target_vel - qd;
//   bodies: contact count, min signed distance (clipped to +-kPhiClip, +kPhiClip when
//     the body has no contact), summed contact normal (pointing to the partner) in the
//     body frame of the world's first free joint;
//   the previous step's correction (converged v minus the plain start, as the step
//     kernel writes it for training), zeros on the first step.
// Signed distances use the public body poses: phi = (X1 p1 - X0 p0) . n - margin0 - margin1,
// the contact writer's convention (sim/collision/pipeline.py:51-118).
#include <hip/hip_runtime.h>

#include "sapx_params.h"

namespace sapx {

namespace {

constexpr float kPhiClip = 0.05f;
constexpr int kFeatContacts = 256;  // contacts staged in LDS per pass

struct R3 {
    float m[9];
};

__device__ __forceinline__ R3 quat_rot(float x, float y, float z, float w) {
    const float s = rsqrtf(fmaxf(x * x + y * y + z * z + w * w, 1.0e-30f));
    x *= s; y *= s; z *= s; w *= s;
    R3 r;
    r.m[0] = 1 - 2 * (y * y + z * z); r.m[1] = 2 * (x * y - z * w); r.m[2] = 2 * (x * z + y * w);
    r.m[3] = 2 * (x * y + z * w); r.m[4] = 1 - 2 * (x * x + z * z); r.m[5] = 2 * (y * z - x * w);
    r.m[6] = 2 * (x * z - y * w); r.m[7] = 2 * (y * z + x * w); r.m[8] = 1 - 2 * (x * x + y * y);
    return r;
}
__device__ __forceinline__ void rot_t(const R3& R, float a, float b, float c, float* o) {
    o[0] = R.m[0] * a + R.m[3] * b + R.m[6] * c;
    o[1] = R.m[1] * a + R.m[4] * b + R.m[7] * c;
    o[2] = R.m[2] * a + R.m[5] * b + R.m[8] * c;
}
__device__ __forceinline__ void tf_point(const float* X, const float* p, float* o) {
    const R3 R = quat_rot(X[3], X[4], X[5], X[6]);
    o[0] = X[0] + R.m[0] * p[0] + R.m[1] * p[1] + R.m[2] * p[2];
    o[1] = X[1] + R.m[3] * p[0] + R.m[4] * p[1] + R.m[5] * p[2];
    o[2] = X[2] + R.m[6] * p[0] + R.m[7] * p[1] + R.m[8] * p[2];
}

__global__ __launch_bounds__(64) void sapx_features_kernel(FeatureParams P) {
    __shared__ float s_phi[kFeatContacts];
    __shared__ float s_n[kFeatContacts * 3];
    __shared__ int s_b0[kFeatContacts], s_b1[kFeatContacts];
    const DevModel& M = P.m;
    const int w = blockIdx.x, lane = threadIdx.x;
    const int J = M.J, NB = M.NB;
    float* out = P.out + (size_t)w * P.F;
    const float* q = P.jq + (size_t)w * M.NQ;
    const float* qd = P.jqd + (size_t)w * M.N;
    for (int j = lane; j < J; j += 64) {
        const int jt = M.t_jtype[j], qs = M.t_jqs[j], ds = M.t_jds[j], nd = M.t_jds[j + 1] - ds;
        float* o = out + P.feat_off[j];
        if (jt == kJointFree || jt == kJointDistance) {
            const R3 R = quat_rot(q[qs + 3], q[qs + 4], q[qs + 5], q[qs + 6]);
            o[0] = q[qs + 2];
            rot_t(R, 0.0f, 0.0f, 1.0f, o + 1);
            rot_t(R, qd[ds], qd[ds + 1], qd[ds + 2], o + 4);
            rot_t(R, qd[ds + 3], qd[ds + 4], qd[ds + 5], o + 7);
        } else {
            for (int k = 0; k < nd; ++k) {
                const size_t gd = (size_t)w * M.N + ds + k;
                const float qv = q[qs + k], qdv = qd[ds + k];
                const float tp = P.tpos ? P.tpos[gd] : 0.0f, tv = P.tvel ? P.tvel[gd] : 0.0f;
                o[4 * k] = qv;
                o[4 * k + 1] = qdv;
                o[4 * k + 2] = tp - qv;
                o[4 * k + 3] = tv - qdv;
            }
        }
    }
    // per-body contact summaries, accumulated in contact order by the body's lane
    int c_begin = 0, nc = 0;
    if (P.cworld_start) {
        c_begin = P.cworld_start[w];
        nc = P.cworld_start[w + 1] - c_begin;
        nc = nc < 0 ? 0 : (nc > P.C ? P.C : nc);
    }
    float cnt = 0.0f, pmin = kPhiClip, ns0 = 0.0f, ns1 = 0.0f, ns2 = 0.0f;
    const int b0w = w * NB;
    for (int cb = 0; cb < nc; cb += kFeatContacts) {
        const int cn = nc - cb < kFeatContacts ? nc - cb : kFeatContacts;
        __syncthreads();
        for (int t = lane; t < cn; t += 64) {
            const int idx = c_begin + cb + t;
            const int s0 = P.cshape0[idx], s1 = P.cshape1[idx];
            const int g0 = s0 >= 0 ? M.shape_body[s0] : -1, g1 = s1 >= 0 ? M.shape_body[s1] : -1;
            float x0[3], x1[3];
            const float* p0 = P.cpoint0 + (size_t)idx * 3;
            const float* p1 = P.cpoint1 + (size_t)idx * 3;
            if (g0 >= 0) tf_point(P.bq + (size_t)g0 * 7, p0, x0);
            else { x0[0] = p0[0]; x0[1] = p0[1]; x0[2] = p0[2]; }
            if (g1 >= 0) tf_point(P.bq + (size_t)g1 * 7, p1, x1);
            else { x1[0] = p1[0]; x1[1] = p1[1]; x1[2] = p1[2]; }
            const float* nn = P.cnormal + (size_t)idx * 3;
            float phi = (x1[0] - x0[0]) * nn[0] + (x1[1] - x0[1]) * nn[1] + (x1[2] - x0[2]) * nn[2] - P.cmargin0[idx] -
                        P.cmargin1[idx];
            phi = fminf(fmaxf(phi, -kPhiClip), kPhiClip);
            s_phi[t] = phi;
            s_n[3 * t] = nn[0];
            s_n[3 * t + 1] = nn[1];
            s_n[3 * t + 2] = nn[2];
            s_b0[t] = g0 >= 0 ? g0 - b0w : -1;
            s_b1[t] = g1 >= 0 ? g1 - b0w : -1;
        }
        __syncthreads();
        if (lane < NB) {
            const int b = lane;  // one body per lane (the model has at most 64 bodies per world)
            for (int t = 0; t < cn; ++t) {
                const float sg = s_b0[t] == b ? 1.0f : (s_b1[t] == b ? -1.0f : 0.0f);
                if (sg != 0.0f) {
                    cnt += 1.0f;
                    pmin = fminf(pmin, s_phi[t]);
                    ns0 += sg * s_n[3 * t];
                    ns1 += sg * s_n[3 * t + 1];
                    ns2 += sg * s_n[3 * t + 2];
                }
                // a body touching itself (both sides) counts twice, like the oracle
                if (s_b0[t] == b && s_b1[t] == b) {
                    cnt += 1.0f;
                    ns0 -= s_n[3 * t];
                    ns1 -= s_n[3 * t + 1];
                    ns2 -= s_n[3 * t + 2];
                }
            }
        }
    }
    // the previous step's correction (training target layout), zeros when absent
    for (int d = lane; d < M.N; d += 64)
        out[P.feat_off[J] + 5 * NB + d] = P.prev_target ? P.prev_target[(size_t)w * M.N + d] : 0.0f;
    if (lane < NB) {
        float* o = out + P.feat_off[J] + 5 * lane;
        float f[3] = {ns0, ns1, ns2};
        if (P.frame_qs >= 0) {
            const int qs = P.frame_qs;
            const R3 R = quat_rot(q[qs + 3], q[qs + 4], q[qs + 5], q[qs + 6]);
            rot_t(R, ns0, ns1, ns2, f);
        }
        o[0] = cnt;
        o[1] = pmin;
        o[2] = f[0];
        o[3] = f[1];
        o[4] = f[2];
    }
}

}  // namespace

hipError_t launch_features(const FeatureParams& P, hipStream_t stream) {
    hipLaunchKernelGGL(sapx_features_kernel, dim3(P.m.W), dim3(64), 0, stream, P);
    return hipGetLastError();
}

}  // namespace sapx

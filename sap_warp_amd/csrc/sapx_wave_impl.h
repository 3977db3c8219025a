// One-wave-per-world SAP timestep (gfx950).
//
// Each world is one 64-lane wavefront (one workgroup): lane i owns dof i (its
// row of the dynamics matrix A and of the Hessian / Cholesky factor live in
// VGPRs), lane c owns contact c.  Cross-lane data moves by v_readlane
// broadcasts and DPP reductions, so no phase needs a workgroup barrier and
// LDS only holds the contact Jacobian, the kinematics scratch, poses and the
// transposed factor (≈ 30 KB at 35 dofs / 16 contacts: five worlds per CU,
// every wave busy).  The algorithm and the reference citations are those of
// the fused step kernel (sapx_step.hip); this kernel re-maps the same phases
// onto a single wave:
//   0/1 loads + public -> SAP boundary      sap_helpers.py:1482-1522, 1811-1848, 2058-2072
//   2   forward kinematics (float64 poses)  free_motion.py:412-530
//   3   body dynamics                       free_motion.py:680-719
//   4   composites, tau, A (CRBA rows)      free_motion.py:93-147, 843-1179
//   5   v* = v0 + dt A^-1 tau               free_motion.py:1970-2178
//   6   contacts, J, weights, constants     contact_jacobian.py:153-562, contact_solve.py:791-822
//   7   PD / limits / participation         contact_solve.py:405-741
//   8   Newton loop with line search        contact_solve.py:743-3176, 6129-6285
//   9   integration + public FK             solver_sap.py:288-519, sap_kinematics.py:84-229
// NP = dofs rounded up to 16 (<= 64), CB = contact slots of this variant.
#ifndef SAPX_WAVE_IMPL_H
#define SAPX_WAVE_IMPL_H

#include "sapx_kcommon.h"

namespace sapx {
namespace wv {

// ---------------------------------------------------------------------------
// register-row dense kernels (lane i owns row i; rows / columns >= n are the
// identity in H and zero in A; lanes >= NP hold zero rows).
// ---------------------------------------------------------------------------

// sum_j A[row][j] x_j for the LDS row of this lane (A zero-padded to NP columns);
// x is lane j's value (zero on lanes >= n), staged in LDS (xs, 64 doubles) and read back two
// columns per uniform-address load
template <int NP>
__device__ __forceinline__ double lds_row_dot(const float* arow, double x, double* xs, int lane) {
    xs[lane] = x;
    wave_sync();
    const double2* x2 = reinterpret_cast<const double2*>(xs);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j += 4) {
        const double2 xa = x2[j / 2], xb = x2[j / 2 + 1];
        s0 += (double)arow[j] * xa.x;
        s1 += (double)arow[j + 1] * xa.y;
        s2 += (double)arow[j + 2] * xb.x;
        s3 += (double)arow[j + 3] * xb.y;
    }
    wave_sync();
    return (s0 + s1) + (s2 + s3);
}

// One pass over the dofs: lane i gets (A x)_i from its LDS row, lane c gets the three
// contact-Jacobian rows of contact c times x.  x_j (zero on lanes >= n) is staged in LDS (xs,
// 64 doubles) and read back two columns per uniform-address load; A and J rows are
// zero-padded to NP columns, so the loop is fully unrolled without guards (every lane's loads
// are independent of the chain).
template <int NP>
__device__ __forceinline__ void matvec_AJ(const float* arow, const float* J0, int js, double x, double& ax,
                                          double& jx0, double& jx1, double& jx2, double* xs, int lane) {
    xs[lane] = x;
    wave_sync();
    const double2* x2 = reinterpret_cast<const double2*>(xs);
    double s0 = 0.0, s1 = 0.0, s2 = 0.0, s3 = 0.0, c0 = 0.0, c1 = 0.0, c2 = 0.0;
#pragma unroll
    for (int j = 0; j < NP; j += 4) {
        const double2 xa = x2[j / 2], xb = x2[j / 2 + 1];
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const double xj = u == 0 ? xa.x : (u == 1 ? xa.y : (u == 2 ? xb.x : xb.y));
            const double t = (double)arow[j + u] * xj;
            if (u == 0) s0 += t; else if (u == 1) s1 += t; else if (u == 2) s2 += t; else s3 += t;
            c0 += (double)J0[j + u] * xj;
            c1 += (double)J0[js + j + u] * xj;
            c2 += (double)J0[2 * js + j + u] * xj;
        }
    }
    wave_sync();
    ax = (s0 + s1) + (s2 + s3);
    jx0 = c0; jx1 = c1; jx2 = c2;
}

// (J^T y)_lane = sum_c sum_r J[3c+r][lane] y_{c,r}, y broadcast from lane c (zero for
// c >= nc; J rows of unused slots are zero).  Contacts in groups of four: the capacity CB is
// a multiple of four, so the last group ends at slot CB - 1 at most.
__device__ __forceinline__ double jt_times(const float* Jm, int js, int nc, int lane, double y0, double y1, double y2) {
    double acc0 = 0.0, acc1 = 0.0;
    const int col = lane < js ? lane : 0;
    for (int c0 = 0; c0 < nc; c0 += 4) {
#pragma unroll
        for (int u = 0; u < 4; ++u) {
            const int c = c0 + u;
            const float* Jc = Jm + (c * 3) * js + col;
            const double t = (double)Jc[0] * readlane_d(y0, c) + (double)Jc[js] * readlane_d(y1, c) +
                             (double)Jc[2 * js] * readlane_d(y2, c);
            if (u & 1) acc1 += t; else acc0 += t;
        }
    }
    return acc0 + acc1;
}

// Right-looking Cholesky of the register rows in panels of eight columns: h becomes L
// (row `lane`, lower, strictly upper part zero), dinv = 1 / L_ii.  Inside a panel the
// pivots and the panel's own columns use v_readlane; the panel is then broadcast
// through `pan` (two float4 per row, rows < NP: 8 NP floats) and every trailing column k
// takes its eight updates from two broadcast reads.  Panels and trailing columns beyond
// the dof count n are skipped (the padding is the identity and needs no work).
template <int NP>
__device__ __forceinline__ void chol_rows(float (&h)[NP], float& dinv, int lane, int n, float4* pan) {
    dinv = 1.0f;
#pragma unroll
    for (int j0 = 0; j0 < NP; j0 += 8) {
        if (j0 < n) {
#pragma unroll
            for (int t = 0; t < 8; ++t) {
                const int j = j0 + t;
                const float pj = fmaxf(rl(h[j], j), 1.0e-30f);
                const float dj = __builtin_amdgcn_rsqf(pj);
                float lij = h[j] * dj;
                if (lane < j) lij = 0.0f;
                h[j] = lij;
                if (lane == j) dinv = dj;
#pragma unroll
                for (int u = t + 1; u < 8; ++u) h[j0 + u] = fmaf(-lij, rl(lij, j0 + u), h[j0 + u]);
            }
            if (j0 + 8 < NP && j0 + 8 < n) {
                if (lane < NP) {
                    pan[2 * lane] = make_float4(h[j0], h[j0 + 1], h[j0 + 2], h[j0 + 3]);
                    pan[2 * lane + 1] = make_float4(h[j0 + 4], h[j0 + 5], h[j0 + 6], h[j0 + 7]);
                }
                wave_sync();
#pragma unroll
                for (int k8 = j0 + 8; k8 < NP; k8 += 8) {
                    if (k8 < n) {
#pragma unroll
                        for (int k = k8; k < k8 + 8; ++k) {
                            const float4 pa = pan[2 * k], pb = pan[2 * k + 1];
                            float a = fmaf(-h[j0], pa.x, h[k]);
                            a = fmaf(-h[j0 + 1], pa.y, a);
                            a = fmaf(-h[j0 + 2], pa.z, a);
                            a = fmaf(-h[j0 + 3], pa.w, a);
                            a = fmaf(-h[j0 + 4], pb.x, a);
                            a = fmaf(-h[j0 + 5], pb.y, a);
                            a = fmaf(-h[j0 + 6], pb.z, a);
                            h[k] = fmaf(-h[j0 + 7], pb.w, a);
                        }
                    }
                }
                wave_sync();
            }
        }
    }
}

// forward substitution L y = b with the register rows
template <int NP>
__device__ __forceinline__ float fwd_rows(const float (&h)[NP], float dinv, float b, int lane) {
    float y = b;
#pragma unroll
    for (int j = 0; j < NP; ++j) {
        const float yj = rl(y * dinv, j);
        y = (lane == j) ? yj : fmaf(-h[j], yj, y);
    }
    return y;
}

// back substitution L^T x = y with the factor's rows stored in LDS (row j at
// Lt + j * ls); lane i reads column i in 16-row blocks ahead of the chain
template <int NP>
__device__ __forceinline__ float bwd_lds(const float* Lt, int ls, float dinv, float y, int lane) {
#pragma unroll
    for (int jb = NP - 16; jb >= 0; jb -= 16) {
        float Lc[16];
#pragma unroll
        for (int t = 0; t < 16; ++t) Lc[t] = Lt[(jb + t) * ls + lane];
#pragma unroll
        for (int t = 15; t >= 0; --t) {
            const int j = jb + t;
            const float xj = rl(y * dinv, j);
            y = (lane == j) ? xj : fmaf(-Lc[t], xj, y);
        }
    }
    return y;
}

template <int NP>
__device__ __forceinline__ void store_rows(const float (&h)[NP], float* Lt, int ls, int lane) {
    if (lane < NP) {
#pragma unroll
        for (int k = 0; k < NP; ++k) Lt[lane * ls + k] = h[k];
    }
}

// L L^T x = b (float32 factor), returns x on each lane
template <int NP>
__device__ __forceinline__ float chol_solve(const float (&h)[NP], const float* Lt, int ls, float dinv, float b, int lane) {
    const float y = fwd_rows<NP>(h, dinv, b, lane);
    wave_sync();
    return bwd_lds<NP>(Lt, ls, dinv, y, lane);
}

// ---------------------------------------------------------------------------
// per-dof PD / limit terms held in registers (contact_solve.py:1569-1612)
// ---------------------------------------------------------------------------
struct DofTerms {
    int pd, lo, up;
    float pda, pdg, pdl, lov, upv, lor, upr;
    double pd_cw, lo_ri, up_ri;  // dt / gain (0 without gain), 1 / r_lower, 1 / r_upper
};

// component k (0..5) of the free-joint block that starts at dof ds of x (float64, LDS),
// rotated by R(q) (joint body frame -> world) or by R(q)^T (world -> body frame); both
// three-row halves of the block rotate alike
__device__ __forceinline__ double free_rot(const double* x, int ds, int k, const float* q, bool to_body) {
    const int g = k < 3 ? 0 : 3;
    const D3 a = d3(x[ds + g], x[ds + g + 1], x[ds + g + 2]);
    DQ r = dqnormalize(DQ{(double)q[0], (double)q[1], (double)q[2], (double)q[3]});
    if (to_body) r = DQ{-r.x, -r.y, -r.z, r.w};
    const D3 b = dqrot_unit(r, a);
    const int c = k - g;
    return c == 0 ? b.x : (c == 1 ? b.y : b.z);
}

__device__ __forceinline__ double dof_terms(const DofTerms& m, double vi, double dt, double* imp, double* hd) {
    double cost = 0.0;
    *imp = 0.0;
    *hd = 0.0;
    if (m.pd) {
        const double gain = m.pdg;
        const double lim = m.pdl;
        const double y = (double)m.pda - gain * vi;
        *imp += dt * clamp_s(y, lim);
        *hd += dt * gain * clamp_d(y, lim);
        cost += m.pd_cw * clamp_a(y, lim);
    }
    if (m.lo) {
        const double r = m.lor;
        const double g = m.lo_ri * ((double)m.lov - vi);
        if (g > 0.0) { *imp += g; *hd += m.lo_ri; cost += 0.5 * r * g * g; }
    }
    if (m.up) {
        const double r = m.upr;
        const double g = m.up_ri * ((double)m.upv + vi);
        if (g > 0.0) { *imp -= g; *hd += m.up_ri; cost += 0.5 * r * g * g; }
    }
    return cost;
}

// ---------------------------------------------------------------------------
// the kernel
// ---------------------------------------------------------------------------
// One world of the step on one wave: `blk` is the world's slot in the launch (mapped through
// P.world_order when set).  Shared by the per-model kernel and the batch kernel.
template <int NP, int CB>
__device__ __forceinline__ void wave_body(const StepParams& P, const int blk) {
    extern __shared__ __attribute__((aligned(16))) char smem_raw[];
    const DevModel& M = P.m;
    const WaveLayout& L = P.wl;
    const int w = P.world_order ? P.world_order[blk] : blk;
    const int lane = threadIdx.x;
    const int n = M.N, nb = M.NB, nj = M.J;
    const int js = L.js, ls = L.ls;
    const int C = P.C;

    // contact range of this world; the variant for fewer contacts owns the world
    int nc = 0, c_begin = 0;
    if (P.cworld_start) {
        c_begin = P.cworld_start[w];
        nc = P.cworld_start[w + 1] - c_begin;
    }
    if (nc > C) nc = C;
    if (nc < 0) nc = 0;
    if (nc > CB || nc < P.wave_min_nc) return;

    const double dt = P.dt_world ? P.dt_world[w] : P.dt;
    const float h = (float)dt;

    int* topo = reinterpret_cast<int*>(smem_raw + L.o_topo);
    float* mp = reinterpret_cast<float*>(smem_raw + L.o_mp);
    float* q_in = reinterpret_cast<float*>(smem_raw + L.o_q);
    float* q_out = q_in + M.NQ;
    float* bq = reinterpret_cast<float*>(smem_raw + L.o_bq);
    double* bqd = reinterpret_cast<double*>(smem_raw + L.o_bqd);
    float* Sg = reinterpret_cast<float*>(smem_raw + L.o_S);
    double* vec = reinterpret_cast<double*>(smem_raw + L.o_vec);  // 2 x 64 doubles of cross-lane data
    double* ck = reinterpret_cast<double*>(smem_raw + L.o_ck);    // CB x kCk
    float* cG = reinterpret_cast<float*>(smem_raw + L.o_cG);      // CB x 9
    float* Jm = reinterpret_cast<float*>(smem_raw + L.o_J);       // 3CB x js
    float* Lt = reinterpret_cast<float*>(smem_raw + L.o_Lt);      // NP x ls (+64)
    float* Am = reinterpret_cast<float*>(smem_raw + L.o_A);       // n x ls dynamics matrix rows
    const float* arow = Am + (lane < n ? lane : 0) * ls;
    // kinematics scratch overlays the J / Lt region (phases 2-4 and 9)
    float* kV = Jm;
    float* kA = kV + nb * 6;
    float* kI = kA + nb * 6;      // nb x 12 (16-byte rows): inertia about the COM, COM offset, mass
    float* kIc = kI + nb * 12;    // nb x 12: subtree composite inertia about the body origin
    float* kF = kIc + nb * 12;    // nb x 6: body force, torque about the body origin
    float* kFc = kF + nb * 6;     // nb x 6: subtree composite force, torque about the body origin
    float* kExt = kFc + nb * 6 + n * 6;
    float* kQdPub = kExt + nb * 6;
    double* kXl = reinterpret_cast<double*>(kQdPub + ((n + 1) & ~1));
    double* kXpn = kXl + nj * 8;  // nj x 8: X_p with its rotation normalised (float64)
    double* kPose = kXpn + nj * 8;  // nb x 8: pointer-jumping partner buffer of bqd (float64)
    float* mXp = mp;
    float* mXc = mXp + nj * 8;
    float* mAx = mXc + nj * 8;
    float* mCom = mAx + n * 4;
    float* mI = mCom + nb * 4;
    float* mM = mI + nb * 9;

    // ------------------------------------------------------------------ 0/1
    // every global input of the prologue is requested in one batch (16-byte copies of the
    // topology template and of this world's parameter block, the state and control rows),
    // then written to LDS: one memory latency instead of one per dependent step
    constexpr int kPre = 6;  // float4 prefetch slots per lane (parameter blocks up to 1536 floats)
    const int wp4 = M.wp_stride >> 2, tp4 = M.topo_ints >> 2;
    const float4* wsrc = reinterpret_cast<const float4*>(M.world_params + (size_t)w * M.wp_stride);
    const int4* tsrc = reinterpret_cast<const int4*>(M.topo);
    // (clamped indices: every slot is loaded unconditionally, so the buffers stay in VGPRs)
    float4 wbuf[kPre];
    int4 tbuf[2];
#pragma unroll
    for (int k = 0; k < kPre; ++k) {
        const int i = k * 64 + lane;
        wbuf[k] = wsrc[i < wp4 ? i : 0];
    }
#pragma unroll
    for (int k = 0; k < 2; ++k) {
        const int i = k * 64 + lane;
        tbuf[k] = tsrc[i < tp4 ? i : 0];
    }
    float qd_l = 0.0f, jf_l = 0.0f;
    if (lane < n) {
        qd_l = P.jqd_in[(size_t)w * n + lane];
        if (P.c_jf) jf_l = P.c_jf[(size_t)w * n + lane];
    }
    float bq_l[7], bf_l[6];
#pragma unroll
    for (int k = 0; k < 7; ++k) bq_l[k] = 0.0f;
#pragma unroll
    for (int k = 0; k < 6; ++k) bf_l[k] = 0.0f;
    if (lane < nb) {
        const float* src = P.bq_in + ((size_t)w * nb + lane) * 7;
#pragma unroll
        for (int k = 0; k < 7; ++k) bq_l[k] = src[k];
        if (P.bf_in) {
            const float* f = P.bf_in + ((size_t)w * nb + lane) * 6;
#pragma unroll
            for (int k = 0; k < 6; ++k) bf_l[k] = f[k];
        }
    }
#pragma unroll
    for (int k = 0; k < kPre; ++k)
        if (k * 64 + lane < wp4) reinterpret_cast<float4*>(mp)[k * 64 + lane] = wbuf[k];
    for (int i = kPre * 64 + lane; i < wp4; i += 64) reinterpret_cast<float4*>(mp)[i] = wsrc[i];
#pragma unroll
    for (int k = 0; k < 2; ++k)
        if (k * 64 + lane < tp4) reinterpret_cast<int4*>(topo)[k * 64 + lane] = tbuf[k];
    for (int i = 128 + lane; i < tp4; i += 64) reinterpret_cast<int4*>(topo)[i] = tsrc[i];
    for (int k = lane; k < M.NQ; k += 64) q_in[k] = P.jq_in[(size_t)w * M.NQ + k];
    // public body poses (bq is free until the kinematics fill it), public velocities and forces
    if (lane < nb) {
#pragma unroll
        for (int k = 0; k < 7; ++k) bq[lane * 8 + k] = bq_l[k];
    }
    vec[lane] = qd_l;
    vec[64 + lane] = jf_l;
    wave_sync();
    const Topo T = topo_view(topo, M);

    // boundary: SAP-order v0 and joint force of dof `lane` (free joints: [w, v_O], [tau_O, f])
    double v0 = 0.0;
    float jf = 0.0f;
    if (lane < n) {
        const int i = lane;
        const int j = T.dof_joint[i];
        const int jt = T.jtype[j];
        v0 = (double)qd_l;
        jf = jf_l;
        if (jt == kJointFree || jt == kJointDistance) {
            const int ds = T.jds[j];
            const int a = i - ds;
            const int child = T.jchild[j];
            const float* xb = bq + child * 8;
            const Q4 xq = Q4{xb[3], xb[4], xb[5], xb[6]};
            const float* com = mCom + child * 4;
            const V3 r = qrot(qnormalize(xq), v3(com[0], com[1], com[2]));
            const double w0 = vec[ds + 3], w1 = vec[ds + 4], w2 = vec[ds + 5];
            const double rx = r.x, ry = r.y, rz = r.z;
            if (a < 3) {
                v0 = a == 0 ? w0 : (a == 1 ? w1 : w2);
            } else {
                const double vc = vec[ds + a - 3];
                const double wxr = a == 3 ? (w1 * rz - w2 * ry) : (a == 4 ? (w2 * rx - w0 * rz) : (w0 * ry - w1 * rx));
                v0 = vc - wxr;
            }
            if (P.c_jf) {
                const double* fj = vec + 64 + ds;
                const V3 fc = v3((float)fj[0], (float)fj[1], (float)fj[2]);
                const V3 tc = v3((float)fj[3], (float)fj[4], (float)fj[5]);
                const V3 to = tc + cross(r, fc);
                jf = a == 0 ? to.x : a == 1 ? to.y : a == 2 ? to.z : a == 3 ? fc.x : a == 4 ? fc.y : fc.z;
            }
        }
    }
    if (lane < nb) {
        // external body force about the body origin (public [f, tau] at the COM)
        float* e = kExt + lane * 6;
        if (P.bf_in) {
            const V3 fc = v3(bf_l[0], bf_l[1], bf_l[2]);
            const V3 tc = v3(bf_l[3], bf_l[4], bf_l[5]);
            const float* com = mCom + lane * 4;
            const V3 c = qrot(qnormalize(Q4{bq_l[3], bq_l[4], bq_l[5], bq_l[6]}), v3(com[0], com[1], com[2]));
            const V3 to = tc + cross(c, fc);
            e[0] = -to.x; e[1] = -to.y; e[2] = -to.z;
            e[3] = -fc.x; e[4] = -fc.y; e[5] = -fc.z;
        } else {
#pragma unroll
            for (int k = 0; k < 6; ++k) e[k] = 0.0f;
        }
    }
    wave_sync();
    vec[lane] = v0;  // v0 by dof for the joint-parallel phases
    wave_sync();

    // ------------------------------------------------------------------ 2
    // bodies that are no joint's child keep the identity pose
    for (int b = lane; b < nb; b += 64) store_dtf(kPose + b * 8, DTF{d3(0.0, 0.0, 0.0), DQ{0.0, 0.0, 0.0, 1.0}});
    wave_sync();
    // forward kinematics: joint-local parts in parallel; world poses in float64 by pointer
    // jumping over the root paths (each body's pose is the product of its path's local
    // transforms L = X_p X_j X_c^-1, combined in ceil(log2 depth) rounds); then the motion
    // subspaces and relative velocities in the world frame in parallel, and the velocity /
    // bias-acceleration recursion of the reference (free_motion.py:412-530) level by level
    for (int j = lane; j < nj; j += 64) {
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
        double plain = 0.0;
        SV vl = sv(v3(0, 0, 0), v3(0, 0, 0));
        SV al = vl;
        D6Ang d6;
        if (jt == kJointRevolute && xcid) {
            Xl.q = dq_axis_angle(axd, (double)q_in[qs]);
            plain = 1.0;
            const SV S = sv(axis, v3(0, 0, 0));
            store_sv(Sg + ds * 6, S);
            vl = sv_scale(S, (float)vec[ds]);
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
            Xl = dtfsel(xcid, Xj, dtf_compose_n(Xj, dtf_inverse_n(dtf_from(Xc))));
            const TF Xmb = tf_inverse_n(Xc);
            const Q4 qj = Q4{(float)Xj.q.x, (float)Xj.q.y, (float)Xj.q.z, (float)Xj.q.w};
            const V3 pmb = qrot(qj, Xmb.p);
            if (jt == kJointPrismatic) {
                const SV S = sv(v3(0, 0, 0), axis);
                store_sv(Sg + ds * 6, S);
                vl = sv_scale(S, (float)vec[ds]);
            } else if (jt == kJointRevolute) {
                const SV S = sv(axis, cross(axis, pmb));
                store_sv(Sg + ds * 6, S);
                const float qd = (float)vec[ds];
                const V3 wf = axis * qd;
                al = sv(v3(0, 0, 0), cross(wf, cross(wf, pmb)));
                vl = sv_scale(S, qd);
            } else if (jt == kJointBall) {
                // three angular mobilities in the joint frame (sap_helpers.py:1345-1361)
                const V3 wv = v3((float)vec[ds], (float)vec[ds + 1], (float)vec[ds + 2]);
                vl = sv(wv, cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
                for (int k = 0; k < 3; ++k) {
                    const V3 e = v3(k == 0, k == 1, k == 2);
                    store_sv(Sg + (ds + k) * 6, sv(e, cross(e, pmb)));
                }
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
                    store_sv(Sg + (ds + k) * 6, sv(vsel(lin, z, a), vsel(lin, a, cross(a, pmb))));
                    const float qd = (float)vec[ds + k];
                    wv = wv + vsel(lin, z, a * qd);
                    vv = vv + vsel(lin, a * qd, z);
                }
                vl = sv(wv, vv + cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
            } else if (jt == kJointFree || jt == kJointDistance) {
                const V3 wv = v3((float)vec[ds], (float)vec[ds + 1], (float)vec[ds + 2]);
                const V3 vv = v3((float)vec[ds + 3], (float)vec[ds + 4], (float)vec[ds + 5]);
                vl = sv(wv, vv + cross(wv, pmb));
                al = sv(v3(0, 0, 0), cross(wv, cross(wv, pmb)));
                for (int k = 0; k < 3; ++k) {
                    const V3 e = v3(k == 0, k == 1, k == 2);
                    store_sv(Sg + (ds + k) * 6, sv(e, cross(e, pmb)));
                    store_sv(Sg + (ds + 3 + k) * 6, sv(v3(0, 0, 0), e));
                }
            }
        }
        store_dtf(kXl + j * 8, Xl);
        kXl[j * 8 + 7] = plain;
        DTF Xpn = dtf_from(load_tf(mXp + j * 8));
        Xpn.q = dqnormalize(Xpn.q);
        store_dtf(kXpn + j * 8, Xpn);
        // the child's pose relative to its parent body (for the root: relative to the world)
        store_dtf(kPose + child * 8, dtf_compose_u(Xpn, Xl));
        store_sv(kV + child * 6, vl);
        store_sv(kA + child * 6, al);
    }
    wave_sync();
    {
        double* src = kPose;
        double* dst = bqd;
        for (int r = 0; r < M.fk_rounds; ++r) {
            if (lane < nb) {
                const int u = T.body_up[r * nb + lane];
                DTF x = load_dtf(src + lane * 8);
                if (u >= 0) x = dtf_compose_u(load_dtf(src + u * 8), x);
                store_dtf(dst + lane * 8, x);
            }
            wave_sync();
            double* t = src;
            src = dst;
            dst = t;
        }
        if (src != bqd) {
            if (lane < nb) store_dtf(bqd + lane * 8, load_dtf(src + lane * 8));
            wave_sync();
        }
    }
    // per joint, in parallel: the parent-joint frame X_wpj (rotation of the local quantities)
    for (int j = lane; j < nj; j += 64) {
        const int parent = T.jparent[j];
        const int child = T.jchild[j];
        const int ds = T.jds[j], nd = T.jds[j + 1] - ds;
        // the root's X_wpj is X_p as given, like the reference
        const DTF Xwpj_d = parent >= 0 ? dtf_compose_u(load_dtf(bqd + parent * 8), load_dtf(kXpn + j * 8))
                                       : dtf_from(load_tf(mXp + j * 8));
        const Q4 R = tf_from(Xwpj_d).q;
        for (int k = 0; k < nd; ++k) {
            const SV Sl = load_sv(Sg + (ds + k) * 6);
            store_sv(Sg + (ds + k) * 6, sv(qrot(R, Sl.w), qrot(R, Sl.v)));
        }
        const SV vl = load_sv(kV + child * 6);
        const SV al = load_sv(kA + child * 6);
        store_sv(kV + child * 6, sv(qrot(R, vl.w), qrot(R, vl.v)));
        store_sv(kA + child * 6, sv(qrot(R, al.w), qrot(R, al.v)));
        const TF Xwc = tf_from(load_dtf(bqd + child * 8));
        float* bqc = bq + child * 8;
        bqc[0] = Xwc.p.x; bqc[1] = Xwc.p.y; bqc[2] = Xwc.p.z;
        bqc[3] = Xwc.q.x; bqc[4] = Xwc.q.y; bqc[5] = Xwc.q.z; bqc[6] = Xwc.q.w;
    }
    wave_sync();
    // velocities and bias accelerations, root to leaves (kV / kA hold the relative ones)
    for (int Lv = 1; Lv < M.num_levels; ++Lv) {
        const int l0 = T.level_start[Lv], l1 = T.level_start[Lv + 1];
        for (int t = l0 + lane; t < l1; t += 64) {
            const int j = T.level_joint[t];
            const int parent = T.jparent[j];
            const int child = T.jchild[j];
            if (parent < 0) continue;
            const SV vPB = load_sv(kV + child * 6);
            const SV aPB = load_sv(kA + child * 6);
            const V3 pPB = v3(bq[child * 8], bq[child * 8 + 1], bq[child * 8 + 2]) -
                           v3(bq[parent * 8], bq[parent * 8 + 1], bq[parent * 8 + 2]);
            const SV Vp = load_sv(kV + parent * 6);
            const SV Ap = load_sv(kA + parent * 6);
            store_sv(kV + child * 6, compose_velocity(Vp, pPB, vPB));
            store_sv(kA + child * 6, compose_acceleration(Ap, Vp, pPB, vPB, aPB));
        }
        wave_sync();
    }

    // ------------------------------------------------------------------ 3
    // body dynamics about each body origin. The composite inertias, the bias forces and the
    // dynamics matrix are taken about the origin of the child body of each dof's joint, with the
    // lever arms between bodies formed from the float64 poses: a light link far down a chain
    // keeps its small inertia (about one reference point for the whole tree, float32 loses it
    // to the parallel-axis terms of the bodies between). The motion subspace about O = the
    // origin of the body's tree root is what the contact Jacobian needs from phase 6 on.
    auto tree_offset = [&](int b) {
        const int r = T.body_root[b];
        return v3((float)(bqd[b * 8] - bqd[r * 8]), (float)(bqd[b * 8 + 1] - bqd[r * 8 + 1]),
                  (float)(bqd[b * 8 + 2] - bqd[r * 8 + 2]));
    };
    for (int b = lane; b < nb; b += 64) {
        const float m = mM[b];
        const float* Ip = mI + b * 9;
        const float* cp = mCom + b * 4;
        const float* g = M.gravity + (size_t)w * 3;
        const float* bqb = bq + b * 8;
        const Q4 q = Q4{bqb[3], bqb[4], bqb[5], bqb[6]};
        const V3 c0 = qrot(q, v3(1, 0, 0)), c1 = qrot(q, v3(0, 1, 0)), c2 = qrot(q, v3(0, 0, 1));
        const float Rm[3][3] = {{c0.x, c1.x, c2.x}, {c0.y, c1.y, c2.y}, {c0.z, c1.z, c2.z}};
        float Tm[3][3], IW[3][3];
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
                Tm[r][cc] = Rm[r][0] * Ip[0 * 3 + cc] + Rm[r][1] * Ip[1 * 3 + cc] + Rm[r][2] * Ip[2 * 3 + cc];
        for (int r = 0; r < 3; ++r)
            for (int cc = 0; cc < 3; ++cc)
                IW[r][cc] = Tm[r][0] * Rm[cc][0] + Tm[r][1] * Rm[cc][1] + Tm[r][2] * Rm[cc][2];
        const V3 c = qrot(q, v3(cp[0], cp[1], cp[2]));
        const float cc2 = dot(c, c);
        float IB[3][3];
        const float cv_[3] = {c.x, c.y, c.z};
        for (int r = 0; r < 3; ++r)
            for (int k = 0; k < 3; ++k) IB[r][k] = IW[r][k] + m * ((r == k ? cc2 : 0.0f) - cv_[r] * cv_[k]);
        const SV V = load_sv(kV + b * 6);
        const SV A = load_sv(kA + b * 6);
        const V3 Iw = v3(IB[0][0] * V.w.x + IB[0][1] * V.w.y + IB[0][2] * V.w.z,
                         IB[1][0] * V.w.x + IB[1][1] * V.w.y + IB[1][2] * V.w.z,
                         IB[2][0] * V.w.x + IB[2][1] * V.w.y + IB[2][2] * V.w.z);
        const V3 Ia = v3(IB[0][0] * A.w.x + IB[0][1] * A.w.y + IB[0][2] * A.w.z,
                         IB[1][0] * A.w.x + IB[1][1] * A.w.y + IB[1][2] * A.w.z,
                         IB[2][0] * A.w.x + IB[2][1] * A.w.y + IB[2][2] * A.w.z);
        V3 fw = Ia + cross(c, A.v) * m;
        V3 fv = (A.v - cross(c, A.w)) * m;
        fw = fw + cross(V.w, Iw);
        fv = fv + cross(V.w, cross(V.w, c)) * m;
        const V3 mg = v3(g[0], g[1], g[2]) * m;
        fw = fw - cross(c, mg);
        fv = fv - mg;
        const float* e = kExt + b * 6;
        float* F = kF + b * 6;
        F[0] = fw.x + e[0]; F[1] = fw.y + e[1]; F[2] = fw.z + e[2];
        F[3] = fv.x + e[3]; F[4] = fv.y + e[4]; F[5] = fv.z + e[5];
        float* I = kI + b * 12;  // inertia about the COM (world axes), COM offset, mass
        int e6 = 0;
        for (int r = 0; r < 3; ++r)
            for (int k = r; k < 3; ++k) I[e6++] = IW[r][k];
        I[6] = c.x; I[7] = c.y; I[8] = c.z; I[9] = m;
    }
    wave_sync();

    // ------------------------------------------------------------------ 4
    // composite inertia and force of each body's subtree about the body's origin: reverse sweep
    // over the levels (free_motion.py:843-1179), each body pulling its children's finished
    // composites shifted by d = p_ch - p_c (float64 poses):
    //   own: I = I_com + m (|c|^2 1 - c c^T), h = m c;
    //   child: I += I_ch + m (|d|^2 1 - d d^T) + 2 (d.h) 1 - d h^T - h d^T, h += h_ch + m d,
    //          torque += torque_ch + d x f_ch, f += f_ch
    for (int Lv = M.num_levels - 1; Lv >= 0; --Lv) {
        const int l0 = T.level_start[Lv], l1 = T.level_start[Lv + 1];
        for (int t = l0 + lane; t < l1; t += 64) {
            const int c = T.jchild[T.level_joint[t]];
            const float* Io = kI + c * 12;
            const float m0 = Io[9];
            const V3 c0 = v3(Io[6], Io[7], Io[8]);
            const V3 h0 = c0 * m0;
            const float hc = dot(h0, c0);
            float I00 = Io[0] + (hc - h0.x * c0.x), I01 = Io[1] - h0.x * c0.y, I02 = Io[2] - h0.x * c0.z;
            float I11 = Io[3] + (hc - h0.y * c0.y), I12 = Io[4] - h0.y * c0.z, I22 = Io[5] + (hc - h0.z * c0.z);
            V3 hs = h0;
            float ms = m0;
            const float* Fo = kF + c * 6;
            V3 ts = v3(Fo[0], Fo[1], Fo[2]), fs = v3(Fo[3], Fo[4], Fo[5]);
            const double pcx = bqd[c * 8], pcy = bqd[c * 8 + 1], pcz = bqd[c * 8 + 2];
            for (int e = T.child_start[c]; e < T.child_start[c + 1]; ++e) {
                const int ch = T.child[e];
                const float* Ich = kIc + ch * 12;
                const V3 dl = v3((float)(bqd[ch * 8] - pcx), (float)(bqd[ch * 8 + 1] - pcy), (float)(bqd[ch * 8 + 2] - pcz));
                const float md = Ich[9];
                const V3 hd = v3(Ich[6], Ich[7], Ich[8]);
                const float dd = dot(dl, dl), dh2 = 2.0f * dot(dl, hd);
                I00 += Ich[0] + md * (dd - dl.x * dl.x) + dh2 - 2.0f * dl.x * hd.x;
                I01 += Ich[1] - md * dl.x * dl.y - dl.x * hd.y - hd.x * dl.y;
                I02 += Ich[2] - md * dl.x * dl.z - dl.x * hd.z - hd.x * dl.z;
                I11 += Ich[3] + md * (dd - dl.y * dl.y) + dh2 - 2.0f * dl.y * hd.y;
                I12 += Ich[4] - md * dl.y * dl.z - dl.y * hd.z - hd.y * dl.z;
                I22 += Ich[5] + md * (dd - dl.z * dl.z) + dh2 - 2.0f * dl.z * hd.z;
                hs = hs + hd + dl * md;
                ms += md;
                const float* Fch = kFc + ch * 6;
                const V3 fd = v3(Fch[3], Fch[4], Fch[5]);
                ts = ts + v3(Fch[0], Fch[1], Fch[2]) + cross(dl, fd);
                fs = fs + fd;
            }
            float* Ic = kIc + c * 12;
            Ic[0] = I00; Ic[1] = I01; Ic[2] = I02; Ic[3] = I11; Ic[4] = I12; Ic[5] = I22;
            Ic[6] = hs.x; Ic[7] = hs.y; Ic[8] = hs.z; Ic[9] = ms;
            float* Fc = kFc + c * 6;
            Fc[0] = ts.x; Fc[1] = ts.y; Fc[2] = ts.z; Fc[3] = fs.x; Fc[4] = fs.y; Fc[5] = fs.z;
        }
        wave_sync();
    }
    // tau and the dof's composite momentum, both about the origin of the dof's child body c
    float tau = 0.0f;
    SV Si = sv(v3(0, 0, 0), v3(0, 0, 0)), SOi = Si, Fii = Si;
    V3 pOi = v3(0, 0, 0);
    if (lane < n) {
        const int ci = T.jchild[T.dof_joint[lane]];
        Si = load_sv(Sg + lane * 6);
        pOi = tree_offset(ci);
        SOi = shift_velocity(Si, -pOi);
        tau = jf - sv_dot(Si, load_sv(kFc + ci * 6));
        const float* I = kIc + ci * 12;
        const V3 hv = v3(I[6], I[7], I[8]);
        const V3 Iw = v3(I[0] * Si.w.x + I[1] * Si.w.y + I[2] * Si.w.z, I[1] * Si.w.x + I[3] * Si.w.y + I[4] * Si.w.z,
                         I[2] * Si.w.x + I[4] * Si.w.y + I[5] * Si.w.z);
        Fii = sv(Iw + cross(hv, Si.v), Si.v * I[9] - cross(hv, Si.w));
    }
    // A_ij = A_ji = (I_c S_i) . S_j about p_ci when joint(j) is on joint(i)'s root path: the lane
    // of dof i writes both entries (S_j moves to p_ci by v + w x (p_ci - p_cj)); entries with
    // neither dof on the other's path are 0 (dof-level ancestor masks from the topology
    // template). Two dofs of one joint are each on the other's path; the higher index writes.
    // A is exactly symmetric: the gradient (A rows) and the line search (A dv) read A from
    // opposite sides, and an asymmetric A stalls the line search near the optimum.
    const unsigned long long dmask = lane < n ? T.dof_anc[lane] : 0ull;
    const float arm_i = (lane < n) ? M.armature[(size_t)w * n + lane] : 0.0f;
    float adiag = 1.0f;
    // p_ci - p_cj as a difference of the two tree offsets (column j's from lane j): I_c S_i about
    // p_ci is small, so the lever's float32 rounding (1e-7 m at 1 m from the root) stays at
    // float32 precision of A_ij; same-tree pairs only (other pairs are 0)
    for (int j = 0; j < n; ++j) {
        const unsigned long long mj = T.dof_anc[j];
        const SV Sj = load_sv(Sg + j * 6);
        const V3 dl = v3(pOi.x - rl(pOi.x, j), pOi.y - rl(pOi.y, j), pOi.z - rl(pOi.z, j));
        float val = sv_dot(Fii, sv(Sj.w, Sj.v + cross(Sj.w, dl)));
        const bool up_i = (dmask >> j) & 1ull, up_j = (mj >> lane) & 1ull;
        const bool mine = up_i && (!up_j || j <= lane);
        if (lane == j) { val += arm_i; adiag = val; }
        if (lane < n) {
            if (mine) {
                Am[lane * ls + j] = val;
                Am[j * ls + lane] = val;
            } else if (!up_j) {
                Am[lane * ls + j] = 0.0f;
            }
        }
    }
    if (lane < n)
        for (int j = n; j < NP; ++j) Am[lane * ls + j] = 0.0f;
    wave_sync();
    // from here on the motion subspace is only needed about O
    if (lane < n) store_sv(Sg + lane * 6, SOi);
    wave_sync();

    // ------------------------------------------------------------------ 5
    // free motion v* = v0 + dt A^-1 tau with one float64-residual refinement
    float hrow[NP];
    float dinv;
#pragma unroll
    for (int j = 0; j < NP; ++j) hrow[j] = (lane < n) ? (j < n ? arow[j] : 0.0f) : ((lane == j) ? 1.0f : 0.0f);
    chol_rows<NP>(hrow, dinv, lane, n, reinterpret_cast<float4*>(Lt));  // Lt is free until store_rows
    store_rows<NP>(hrow, Lt, ls, lane);
    wave_sync();
    const float x1 = chol_solve<NP>(hrow, Lt, ls, dinv, lane < n ? tau : 0.0f, lane);
    const double x1d = lane < n ? (double)x1 : 0.0;
    const double ax1 = lds_row_dot<NP>(arow, x1d, vec, lane);
    const double r1 = lane < n ? (double)tau - ax1 : 0.0;
    const float x2 = chol_solve<NP>(hrow, Lt, ls, dinv, (float)r1, lane);
    const double vs = lane < n ? v0 + dt * (x1d + (double)x2) : 0.0;
    double ainv = 1.0, dsc = 1.0;
    {
        double ad = (lane < n) ? (double)adiag : 1.0;
        if (ad < 1.0e-12 || !isfinite(ad)) ad = 1.0e-12;
        ainv = (float)(1.0 / ad);
        dsc = (float)(1.0 / sqrt(ad));
    }

    // ------------------------------------------------------------------ 6
    // contacts: lane c builds contact c (witnesses / phi0 in float64)
    int cb0 = -1, cb1 = -1;
    V3 cx0 = v3(0, 0, 0), cx1 = cx0, cpc = cx0, tx = cx0, ty = cx0, nrm = cx0;
    double phi = 0.0, kcomb = 1.0, taus = 0.0, muc = 0.0;
    D3 xs0d = d3(0.0, 0.0, 0.0), xs1d = xs0d;  // the Jacobian's points (witnesses or midpoint), float64
    if (lane < nc) {
        const int idx = c_begin + lane;
        const int sh0 = P.cshape0[idx], sh1 = P.cshape1[idx];
        const int b0g = sh0 >= 0 ? M.shape_body[sh0] : -1;
        const int b1g = sh1 >= 0 ? M.shape_body[sh1] : -1;
        cb0 = b0g >= 0 ? b0g - w * nb : -1;
        cb1 = b1g >= 0 ? b1g - w * nb : -1;
        D3 x0d = d3(P.cpoint0[idx * 3], P.cpoint0[idx * 3 + 1], P.cpoint0[idx * 3 + 2]);
        D3 x1d = d3(P.cpoint1[idx * 3], P.cpoint1[idx * 3 + 1], P.cpoint1[idx * 3 + 2]);
        // body poses in bqd carry unit rotations (normalised in the level walk)
        if (cb0 >= 0) { const DTF X = load_dtf(bqd + cb0 * 8); x0d = dadd(X.p, dqrot_unit(X.q, x0d)); }
        if (cb1 >= 0) { const DTF X = load_dtf(bqd + cb1 * 8); x1d = dadd(X.p, dqrot_unit(X.q, x1d)); }
        D3 nd = d3(P.cnormal[idx * 3], P.cnormal[idx * 3 + 1], P.cnormal[idx * 3 + 2]);
        const double nd2 = nd.x * nd.x + nd.y * nd.y + nd.z * nd.z;
        nd = dsel(nd2 <= 1.0e-32, d3(0.0, 0.0, 0.0), dscale(nd, 1.0 / sqrt(nd2)));
        cx0 = v3((float)x0d.x, (float)x0d.y, (float)x0d.z);
        cx1 = v3((float)x1d.x, (float)x1d.y, (float)x1d.z);
        nrm = v3((float)nd.x, (float)nd.y, (float)nd.z);
        // frame: MakeFromOneUnitVector (sap_helpers.py:467-503)
        const float ax = fabsf(nrm.x), ay = fabsf(nrm.y), az = fabsf(nrm.z);
        int im = 0;
        float mn = ax;
        if (ay < mn) { im = 1; mn = ay; }
        if (az < mn) { im = 2; mn = az; }
        const float mag = sqrtf(1.0f - mn * mn);
        const float rr = 1.0f / mag;
        const float u = im == 0 ? nrm.x : (im == 1 ? nrm.y : nrm.z);
        const float sg = -rr * u;
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
        const double kden = k0 + k1;
        double w0 = 0.5, w1 = 0.5;
        if (kden != 0.0) { w0 = k0 / kden; w1 = k1 / kden; }
        const D3 pcd = dadd(dscale(x0d, w0), dscale(x1d, w1));
        cpc = v3((float)pcd.x, (float)pcd.y, (float)pcd.z);
        xs0d = dsel(P.point_mode == 0, x0d, pcd);
        xs1d = dsel(P.point_mode == 0, x1d, pcd);
        if (!isfinite(k0)) kcomb = k1;
        else if (!isfinite(k1)) kcomb = k0;
        else if (kden <= 1.0e-12) kcomb = fmax(k0, k1);
        else kcomb = k0 * k1 / kden;
        double mu0 = P.fallback_mu, mu1 = P.fallback_mu;
        if (sh0 >= 0) mu0 = M.shape_mu[sh0];
        if (sh1 >= 0) mu1 = M.shape_mu[sh1];
        mu0 = fmax(mu0, 0.0);
        mu1 = fmax(mu1, 0.0);
        const double mud = mu0 + mu1;
        muc = (mud <= 1.0e-12) ? 0.0 : 2.0 * mu0 * mu1 / mud;
        double t0 = P.fallback_tau, t1 = P.fallback_tau;
        if (sh0 >= 0) t0 = M.shape_tau[sh0];
        if (!isfinite(t0) || t0 < 0.0) t0 = P.fallback_tau;
        if (!isfinite(t0) || t0 < 0.0) t0 = 0.0;
        if (sh1 >= 0) t1 = M.shape_tau[sh1];
        if (!isfinite(t1) || t1 < 0.0) t1 = P.fallback_tau;
        if (!isfinite(t1) || t1 < 0.0) t1 = 0.0;
        taus = t0 + t1;
        const D3 gapv = dsub(x1d, x0d);
        phi = gapv.x * nd.x + gapv.y * nd.y + gapv.z * nd.z - (double)P.cmargin0[idx] - (double)P.cmargin1[idx];
    }
    // contact Jacobian rows of contact `lane`: J_c[:, d] = R^T (u1 - u0); every row and
    // padding column of the region starts at zero, and G of unused slots is zero
    for (int t = lane; t < 3 * CB * js; t += 64) Jm[t] = 0.0f;
    for (int t = lane; t < CB * 9; t += 64) cG[t] = 0.0f;
    wave_sync();
    vec[64 + lane] = (lane < n) ? ainv : 0.0;   // diag(A)^-1 by dof for the Delassus weights
    if (lane < nc) {
        // x - O of the body's tree root, from float64 points and poses
        auto lever = [&](D3 x, int b) {
            const int r = T.body_root[b];
            return v3((float)(x.x - bqd[r * 8]), (float)(x.y - bqd[r * 8 + 1]), (float)(x.z - bqd[r * 8 + 2]));
        };
        // dof d moves the contact point x of a body on its root path at S_O.v + S_O.w x (x - O)
        // (S_O: motion subspace about the origin O of the dof's tree, which is the body's tree;
        // equal to S.v + S.w x (x - p_child))
        const unsigned long long anc1 = cb1 >= 0 ? T.body_anc_dof[cb1] : 0ull;
        const unsigned long long anc0 = cb0 >= 0 ? T.body_anc_dof[cb0] : 0ull;
        const V3 r1 = vsel(cb1 >= 0, lever(xs1d, cb1 >= 0 ? cb1 : 0), v3(0, 0, 0));
        const V3 r0 = vsel(cb0 >= 0, lever(xs0d, cb0 >= 0 ? cb0 : 0), v3(0, 0, 0));
        float* J0 = Jm + (lane * 3) * js;
#pragma unroll 2
        for (int d = 0; d < n; ++d) {
            const SV S = load_sv(Sg + d * 6);
            const float s1 = ((anc1 >> d) & 1ull) ? 1.0f : 0.0f;
            const float s0 = ((anc0 >> d) & 1ull) ? 1.0f : 0.0f;
            const V3 u1 = S.v + cross(S.w, r1), u0 = S.v + cross(S.w, r0);
            const V3 rel = u1 * s1 - u0 * s0;
            J0[d] = dot(tx, rel);
            J0[js + d] = dot(ty, rel);
            J0[2 * js + d] = dot(nrm, rel);
        }
    }
    wave_sync();
    // effective weights + regularisation constants (contact_jacobian.py:153-251, contact_solve.py:791-822)
    if (lane < nc) {
        double wvv;
        if (P.weight_mode == 0) {
            const V3 dirs[3] = {tx, ty, nrm};
            double acc = 0.0;
            for (int side = 0; side < 2; ++side) {
                const int lb = side == 0 ? cb0 : cb1;
                if (lb < 0) continue;
                const double m = mM[lb];
                if (m <= 1.0e-12) continue;
                const float* bb = bq + lb * 8;
                const Q4 q = qnormalize(Q4{bb[3], bb[4], bb[5], bb[6]});
                const float* com = mCom + lb * 4;
                const V3 comw = v3(bb[0], bb[1], bb[2]) + qrot(q, v3(com[0], com[1], com[2]));
                const V3 r = vsel(side == 0, cx0, cx1) - comw;
                const float* I = mI + lb * 9;
                const double ix = fmax((double)I[0], 1.0e-12), iy = fmax((double)I[4], 1.0e-12), iz = fmax((double)I[8], 1.0e-12);
                for (int k = 0; k < 3; ++k) {
                    V3 dn = dirs[k];
                    const float dl = sqrtf(dot(dn, dn));
                    dn = vsel(dl * dl <= 1.0e-32f, v3(0, 0, 0), dn * (1.0f / dl));
                    const V3 cbv = qrot(qconj(q), cross(r, dn));
                    acc += 1.0 / m + (double)cbv.x * cbv.x / ix + (double)cbv.y * cbv.y / iy + (double)cbv.z * cbv.z / iz;
                }
            }
            wvv = acc / 3.0;
        } else {
            double W00 = 0, W01 = 0, W02 = 0, W11 = 0, W12 = 0, W22 = 0;
            const float* J0 = Jm + (lane * 3) * js;
            for (int d = 0; d < n; ++d) {
                const double ai = vec[64 + d];
                const double a0 = J0[d], a1 = J0[js + d], a2 = J0[2 * js + d];
                W00 += ai * a0 * a0; W01 += ai * a0 * a1; W02 += ai * a0 * a2;
                W11 += ai * a1 * a1; W12 += ai * a1 * a2; W22 += ai * a2 * a2;
            }
            wvv = sqrt(W00 * W00 + W11 * W11 + W22 * W22 + 2.0 * (W01 * W01 + W02 * W02 + W12 * W12)) / 3.0;
        }
        if (wvv < 1.0e-12 || !isfinite(wvv)) wvv = 1.0e-12;
        double kc = kcomb;
        if (kc <= 0.0 || !isfinite(kc)) kc = 1.0;
        const double tau_c = fmax(taus, 0.0);
        const double bfac = P.beta * P.beta / (4.0 * kPi * kPi);
        const double rn_soft = 1.0 / (dt * kc * (dt + tau_c));
        const double rn = fmax(fmax(bfac * wvv, rn_soft), 1.0e-30);
        const double rt = fmax(P.sigma * wvv, 1.0e-30);
        double mu = muc;
        if (mu < 0.0 || !isfinite(mu)) mu = 0.0;
        double* k8 = ck + lane * kCk;
        k8[0] = rt;
        k8[1] = rn;
        k8[2] = -phi / (dt + tau_c);
        k8[3] = mu;
        k8[4] = 1.0 / rt;
        k8[5] = 1.0 / rn;
        const double mu_tilde = mu * sqrt(rt / rn);
        k8[6] = mu * rt / rn;
        k8[7] = 1.0 / (1.0 + mu_tilde * mu_tilde);
        if (P.dbg_phi0) P.dbg_phi0[(size_t)w * C + lane] = phi;
        if (P.dbg_w) P.dbg_w[(size_t)w * C + lane] = wvv;
    }
    wave_sync();

    // ------------------------------------------------------------------ 7
    // PD + limit terms (contact_solve.py:490-683) and participating dofs (422-487)
    DofTerms mt = {0, 0, 0, 0.0f, 0.0f, 0.0f, 0.0f, 0.0f, 1.0f, 1.0f, 0.0, 1.0, 1.0};
    int part = 0;
    if (lane < n) {
        const int i = lane;
        const int t = T.dof_target[i];
        const size_t gt = (size_t)w * n + t;
        const size_t gi = (size_t)w * n + i;
        const int coord = T.dof_coord[i];
        if (M.target_mode[gt] != 0) {
            const double kp = M.target_ke[gt], kd = M.target_kd[gt];
            if (!(kp <= 0.0 && kd <= 0.0) && !(coord < 0 && kp > 0.0)) {
                const double pbf = kPdBeta * kPdBeta / (4.0 * kPi * kPi);
                const double r_nr = pbf * fmax(ainv, 1.0e-12);
                const double common = dt * kp + kd;
                double kpe = 0.0, kde = 0.0;
                if (common > 0.0) {
                    const double r = 1.0 / (dt * common);
                    if (r >= r_nr) { kpe = kp; kde = kd; }
                    else if (dt * kp > kd && kp > 0.0) {
                        const double tc = kd / kp;
                        kpe = 1.0 / (dt * (dt + tc) * r_nr);
                        kde = tc * kpe;
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
                    mt.pd = 1;
                    mt.pda = (float)(kpe * (qt - q0) + kde * vd + u0);
                    mt.pdg = (float)gain;
                    mt.pdl = M.effort_limit[gt];
                    mt.pd_cw = mt.pdg > 0.0f ? dt / (double)mt.pdg : 0.0;
                }
            }
        }
        if (coord >= 0 && T.dof_limit_ok[i]) {
            const double ke = M.limit_ke[gi];
            if (ke > 0.0) {
                const double q0 = q_in[coord];
                const double lower = M.limit_lower[gi], upper = M.limit_upper[gi];
                const double window = kLimitWindow * dt * fmax(fabs(v0), fabs(vs));
                const double kd = M.limit_kd[gi];
                const double tau_d = kd > 0.0 ? kd / ke : 0.0;
                const double lbf = kLimitBeta * kLimitBeta / (4.0 * kPi * kPi);
                const double r = fmax(lbf * fmax(ainv, 1.0e-12), 1.0 / (dt * ke * (dt + tau_d)));
                if (isfinite(lower) && q0 - lower <= window) {
                    mt.lo = 1; mt.lov = (float)(-(q0 - lower) / (dt + tau_d)); mt.lor = (float)r;
                    mt.lo_ri = 1.0 / (double)mt.lor;
                }
                if (isfinite(upper) && upper - q0 <= window) {
                    mt.up = 1; mt.upv = (float)(-(upper - q0) / (dt + tau_d)); mt.upr = (float)r;
                    mt.up_ri = 1.0 / (double)mt.upr;
                }
            }
        }
        part = mt.pd | mt.lo | mt.up;
    }
    for (int c = 0; c < nc; ++c) {
        const int b0 = __builtin_amdgcn_readlane(cb0, c), b1 = __builtin_amdgcn_readlane(cb1, c);
        if (lane < n && !part) {
            for (int side = 0; side < 2; ++side) {
                const int lb = side == 0 ? b0 : b1;
                if (lb >= 0) {
                    const int st = T.body_art_start[lb], ct = T.body_art_count[lb];
                    if (st >= 0 && ct > 0 && lane >= st && lane < st + ct) part = 1;
                }
            }
            const float* J0 = Jm + (c * 3) * js;
            if (J0[lane] != 0.0f || J0[js + lane] != 0.0f || J0[2 * js + lane] != 0.0f) part = 1;
        }
    }

    // ------------------------------------------------------------------ 8
    // SAP Newton solve (contact_solve.py:6129-6285, 5474-5655)
    if (P.dbg_A) {
        for (int t = lane; t < n * n; t += 64) P.dbg_A[(size_t)w * n * n + t] = Am[(t / n) * ls + (t % n)];
    }
    if (P.dbg_J) {
        const size_t jb = (size_t)w * C * 3 * n;
        for (int t = lane; t < C * 3 * n; t += 64) {
            const int r = t / n;
            P.dbg_J[jb + t] = (r < 3 * nc) ? (double)Jm[r * js + (t % n)] : 0.0;
        }
    }
    const bool unconstrained = (nc == 0) && !__any(mt.pd | mt.lo | mt.up);
    double v = 0.0;
    if (lane < n) {
        v = P.use_guess ? P.v_guess[(size_t)w * n + lane] : v0;
        if (unconstrained || !part) v = vs;
    }
    double avs = lds_row_dot<NP>(arow, vs, vec, lane);
    if (lane >= n) avs = 0.0;
    const double* k8 = ck + (lane < CB ? lane : 0) * kCk;
    // learned warm start: v_plain + R delta (free-joint rows rotated out of the joint's body
    // frame) replaces the plain start when its SAP cost is lower; otherwise the plain start
    // stays (the fallback)
    const double v_plain = v;
    int start_kind = 0;  // 0 plain, 1 learned start taken, 2 learned start rejected
    int fr_ds = -1, fr_k = 0;
    float fq[4] = {0.0f, 0.0f, 0.0f, 1.0f};
    if (lane < n) {
        const int jd = T.dof_joint[lane];
        const int jt = T.jtype[jd];
        if (jt == kJointFree || jt == kJointDistance) {
            fr_ds = T.jds[jd];
            fr_k = lane - fr_ds;
            const int qs = T.jqs[jd];
            fq[0] = q_in[qs + 3]; fq[1] = q_in[qs + 4]; fq[2] = q_in[qs + 5]; fq[3] = q_in[qs + 6];
        }
    }
    if (P.learn_delta && !unconstrained) {
        auto sap_cost = [&](double x) -> double {
            double ax, c0, c1, c2;
            const float* J0 = Jm + ((lane < CB ? lane : 0) * 3) * js;
            matvec_AJ<NP>(arow, J0, js, x, ax, c0, c1, c2, vec, lane);
            double cst = 0.0;
            if (lane < nc) {
                const double vcv[3] = {c0, c1, c2};
                double gg[3];
                float Gd[9];
                cst += project_contact(k8, vcv, gg, Gd, false);
            }
            if (lane < n) {
                double imp, hdd;
                cst += dof_terms(mt, x, dt, &imp, &hdd);
                cst += 0.5 * (x - vs) * (ax - avs);
            }
            return wave_sum(cst);
        };
        vec[lane] = lane < n ? (double)P.learn_delta[(size_t)w * n + lane] : 0.0;
        wave_sync();
        double dl = lane < n ? vec[lane] : 0.0;
        if (fr_ds >= 0) dl = free_rot(vec, fr_ds, fr_k, fq, false);
        wave_sync();
        const double vp = (lane < n && part) ? v + dl : v;
        const double c_plain = sap_cost(v);
        const double c_pred = sap_cost(vp);
        if (c_pred < c_plain) {
            v = vp;
            start_kind = 1;
        } else {
            start_kind = 2;
        }
    }
    int iteration = 0, ls_total = 0, flags = 0, ls_status = 0;
    // lane-c contact state: vc, dvc, gamma (float64), G in LDS (cG)
    double vc0 = 0, vc1 = 0, vc2 = 0, dvc0 = 0, dvc1 = 0, dvc2 = 0, g0 = 0, g1 = 0, g2 = 0;
    // lane-i dof state
    double grad = 0.0, ar = 0.0, dvv = 0.0;
    float hd = 0.0f;
    double alpha = 1.0, prev_cost = 0.0;
    const double cost_min_alpha = P.ls_mode == kLsMonotone ? 0.0 : 0.5;  // contact_solve.py:6182
    // A v and J v carried between iterations: after a step v + alpha dv they are updated with the
    // line search's A dv and J dv (float64), so only the first evaluation runs the mat-vec
    double av_run = 0.0, adv_ls = 0.0;
    bool have_av = false;
    if (unconstrained) {
        flags = 1;
    } else {
        while (true) {
            // ---- evaluate at v: vc = J v, A v; projection + model terms; impulses / gradient / norms
            double av = av_run;
            if (!have_av) {
                const float* J0 = Jm + ((lane < CB ? lane : 0) * 3) * js;
                double c0, c1, c2;
                matvec_AJ<NP>(arow, J0, js, v, av, c0, c1, c2, vec, lane);
                vc0 = c0; vc1 = c1; vc2 = c2;
            }
            if (lane >= n) av = 0.0;
            ar = av - avs;
            double part4[4] = {0.0, 0.0, 0.0, 0.0};  // cost, |D g|^2, |D A v|^2, |D j|^2
            if (lane < nc) {
                const double vcv[3] = {vc0, vc1, vc2};
                double gg[3];
                part4[0] += project_contact(k8, vcv, gg, cG + lane * 9, true);
                g0 = gg[0]; g1 = gg[1]; g2 = gg[2];
            }
            double imp = 0.0, hdd = 0.0;
            if (lane < n) part4[0] += dof_terms(mt, v, dt, &imp, &hdd);
            hd = (float)hdd;
            if (lane >= nc) { g0 = 0.0; g1 = 0.0; g2 = 0.0; }
            const double jtg = jt_times(Jm, js, nc, lane, g0, g1, g2);
            double jc = imp + (lane < n ? jtg : 0.0);
            grad = ar - jc;
            if (lane < n) {
                part4[0] += 0.5 * (v - vs) * ar;
                if (part) {
                    part4[1] += (dsc * grad) * (dsc * grad);
                    part4[2] += (dsc * av) * (dsc * av);
                    part4[3] += (dsc * jc) * (dsc * jc);
                }
            }
#pragma unroll
            for (int k = 0; k < 4; ++k) part4[k] = wave_sum(part4[k]);
            const double cost = part4[0];
            const double gnorm = sqrt(fmax(part4[1], 0.0));
            const double pnorm = sqrt(fmax(part4[2], 0.0));
            const double jnorm = sqrt(fmax(part4[3], 0.0));
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
                break;
            }
            if (iteration >= P.max_iter) {
                flags = 8;
                break;
            }
            wave_sync();
            // ---- Hessian rows H = A + J^T G J + diag(hd + shift); J^T G J by MFMA f32 16x16x4
            //      into LDS: k-steps outer, all TB x TB tiles accumulate from one set of operands
            {
                constexpr int TB = NP / 16;
                typedef float f4 __attribute__((ext_vector_type(4)));
                f4 acc[TB][TB];
#pragma unroll
                for (int ib = 0; ib < TB; ++ib)
#pragma unroll
                    for (int jb = 0; jb < TB; ++jb) acc[ib][jb] = f4{0.0f, 0.0f, 0.0f, 0.0f};
                const int K = 3 * nc;
                const int col = lane & 15;
                for (int kk0 = 0; kk0 < K; kk0 += 4) {
                    const int k = kk0 + (lane >> 4);  // k < 3 CB: padded rows are zero
                    const int c = k / 3, r = k - 3 * (k / 3);
                    const float* G = cG + c * 9 + r * 3;
                    const float g0f = G[0], g1f = G[1], g2f = G[2];
                    const float* Jk = Jm + k * js;
                    const float* Jc = Jm + (c * 3) * js;
                    float av_[TB], bv_[TB];
#pragma unroll
                    for (int t = 0; t < TB; ++t) {
                        av_[t] = Jk[t * 16 + col];
                        bv_[t] = g0f * Jc[t * 16 + col] + g1f * Jc[js + t * 16 + col] + g2f * Jc[2 * js + t * 16 + col];
                    }
#pragma unroll
                    for (int ib = 0; ib < TB; ++ib)
#pragma unroll
                        for (int jb = 0; jb < TB; ++jb)
                            acc[ib][jb] = __builtin_amdgcn_mfma_f32_16x16x4f32(av_[ib], bv_[jb], acc[ib][jb], 0, 0, 0);
                }
#pragma unroll
                for (int ib = 0; ib < TB; ++ib)
#pragma unroll
                    for (int jb = 0; jb < TB; ++jb)
#pragma unroll
                        for (int e = 0; e < 4; ++e)
                            Lt[(ib * 16 + 4 * (lane >> 4) + e) * ls + jb * 16 + col] = acc[ib][jb][e];
            }
            wave_sync();
            const float hdiag = hd + (float)P.diag_shift;
#pragma unroll
            for (int j = 0; j < NP; ++j) {
                const float hc = Lt[(lane < NP ? lane : 0) * ls + j];
                hrow[j] = (lane < n) ? (j < n ? arow[j] + hc + ((lane == j) ? hdiag : 0.0f) : 0.0f) : ((lane == j) ? 1.0f : 0.0f);
            }
            wave_sync();
            chol_rows<NP>(hrow, dinv, lane, n, reinterpret_cast<float4*>(Lt));  // Lt is free until store_rows
            store_rows<NP>(hrow, Lt, ls, lane);
            wave_sync();
            float dvf = chol_solve<NP>(hrow, Lt, ls, dinv, lane < n ? (float)(-grad) : 0.0f, lane);
            dvv = lane < n ? (double)dvf : 0.0;
            // ---- mixed-precision refinement: r = -grad - (A + J^T G J + diag) dv in float64
            for (int rep = 0; rep < kNewtonRefine; ++rep) {
                double t0, t1, t2, adv;
                {
                    const float* J0 = Jm + ((lane < CB ? lane : 0) * 3) * js;
                    matvec_AJ<NP>(arow, J0, js, dvv, adv, t0, t1, t2, vec, lane);
                }
                double u0 = 0.0, u1 = 0.0, u2 = 0.0;
                if (lane < nc) {
                    const float* G = cG + lane * 9;
                    u0 = (double)G[0] * t0 + (double)G[1] * t1 + (double)G[2] * t2;
                    u1 = (double)G[3] * t0 + (double)G[4] * t1 + (double)G[5] * t2;
                    u2 = (double)G[6] * t0 + (double)G[7] * t1 + (double)G[8] * t2;
                }
                const double jtu = jt_times(Jm, js, nc, lane, u0, u1, u2);
                const double acc = adv + ((double)hd + P.diag_shift) * dvv + jtu;
                const double res = lane < n ? -grad - acc : 0.0;
                const float corr = chol_solve<NP>(hrow, Lt, ls, dinv, (float)res, lane);
                if (lane < n) dvv += (double)corr;
            }
            // ---- line search data (contact_solve.py:2234-2263)
            double ld[4] = {0, 0, 0, 0};  // dell0, dell_a0, d2ell_a, m0
            {
                double adv, d0, d1, d2;
                {
                    const float* J0 = Jm + ((lane < CB ? lane : 0) * 3) * js;
                    matvec_AJ<NP>(arow, J0, js, dvv, adv, d0, d1, d2, vec, lane);
                }
                dvc0 = d0; dvc1 = d1; dvc2 = d2;
                adv_ls = adv;
                if (lane < n) {
                    const double ri = v - vs;
                    ld[0] = grad * dvv;
                    ld[1] = adv * ri;
                    ld[2] = dvv * adv;
                    ld[3] = 0.5 * ri * ar;
                }
#pragma unroll
                for (int k = 0; k < 4; ++k) ld[k] = wave_sum(ld[k]);
            }
            const double dell0 = ld[0], dell_a0 = ld[1], d2ell_a = ld[2], m0 = ld[3];
            // trial evaluation of phi(a), phi'(a), phi''(a)
            auto trial = [&](double al, bool second, double* dd1, double* dd2) -> double {
                double tv[3] = {0.0, 0.0, 0.0};
                if (lane < nc) {
                    const double dvc[3] = {dvc0, dvc1, dvc2};
                    const double vcv[3] = {vc0 + al * dvc0, vc1 + al * dvc1, vc2 + al * dvc2};
                    double gg[3], cc;
                    if (second) {
                        tv[2] += project_curvature(k8, vcv, dvc, gg, &cc);
                    } else {
                        float Gd[9];
                        cc = project_contact(k8, vcv, gg, Gd, false);
                    }
                    tv[0] += cc;
                    tv[1] -= dvc[0] * gg[0] + dvc[1] * gg[1] + dvc[2] * gg[2];
                }
                if (lane < n) {
                    double impt, hdt;
                    tv[0] += dof_terms(mt, v + al * dvv, dt, &impt, &hdt);
                    tv[1] -= dvv * impt;
                    tv[2] += hdt * dvv * dvv;
                }
#pragma unroll
                for (int k = 0; k < 3; ++k) tv[k] = wave_sum(tv[k]);
                *dd1 = dell_a0 + al * d2ell_a + tv[1];
                *dd2 = d2ell_a + tv[2];
                return m0 + al * dell_a0 + 0.5 * al * al * d2ell_a + tv[0];
            };
            bool accepted = false;
            double a_acc = 1.0;
            int ls_it = 0;
            ls_status = 0;
            if (P.ls_mode == kLsMonotone) {
                // fused halving search (contact_solve.py:2486-2665): the first alpha = 1, 1/2, ...
                // whose cost does not exceed the current one, relaxed by 1e-14 + 1e-12 |cost|
                const double tol = kMonoRelaxA + kMonoRelaxR * fabs(cost);
                double al = 1.0, dd1, dd2;
                for (int it = 0; it < P.ls_max_iter; ++it) {
                    const double ell = trial(al, false, &dd1, &dd2);
                    ++ls_it;
                    if (ell <= cost + tol) { accepted = true; a_acc = al; break; }
                    al *= kMonoDecay;
                    if (al < kMonoMinAlpha) { ls_status = -2; break; }
                }
                if (!accepted) a_acc = al;
            } else if (!(dell0 < 0.0) || !isfinite(dell0)) {
                ls_status = -1;
            } else if (P.ls_mode == 0) {
                // Armijo backtracking (contact_solve.py:2670-2905, 5108-5305)
                const double amax = 1.0 / P.rho;
                double al = amax, dd1, dd2;
                double ell = trial(al, false, &dd1, &dd2);
                const double slop = (P.rel_slop / 10.0) * fmax(1.0, 0.5 * (fabs(ell) + fabs(cost)));
                if (dd1 < 0.0 || dd1 < slop) {
                    accepted = true;
                    a_acc = al;
                } else if (P.ls_max_iter > 1) {
                    double a_prev = al, e_prev = ell;
                    int it = 1;
                    while (true) {
                        al = al * P.rho;
                        ell = trial(al, false, &dd1, &dd2);
                        const double den = al - a_prev;
                        const double dapprox = fabs(den) > 0.0 ? (ell - e_prev) / den : 0.0;
                        if (fabs(dapprox) < slop) { accepted = true; a_acc = al; ls_it = it; break; }
                        if (ell > e_prev && ell < cost + P.armijo_c * al * dell0) {
                            accepted = true;
                            ls_it = it;
                            a_acc = (e_prev < cost + P.armijo_c * a_prev * dell0) ? a_prev : al;
                            break;
                        }
                        if (it >= P.ls_max_iter - 1) {
                            if (ell < cost + P.armijo_c * al * dell0) { accepted = true; a_acc = al; ls_it = P.ls_max_iter; }
                            else ls_status = -2;
                            break;
                        }
                        a_prev = al;
                        e_prev = ell;
                        ++it;
                    }
                } else {
                    ls_status = -2;
                }
            } else {
                // exact root (contact_solve.py:2913-3140, 5700-5908), cubic Hermite seed
                const double amax = kExactAlphaMax;
                double dd1, dd2;
                const double ell = trial(amax, true, &dd1, &dd2);
                if (dd1 <= 0.0) {
                    accepted = true; a_acc = amax;
                } else if (-dell0 < P.cost_abs + P.cost_rel * ell) {
                    accepted = true; a_acc = 1.0;
                } else {
                    const double scale = -dell0;
                    const double f_upper0 = dd1 / scale;
                    if (!isfinite(f_upper0) || f_upper0 <= 0.0) {
                        ls_status = -3;
                    } else {
                        double guess = -dell0 / dd2;
                        const double hh = amax;
                        const double theta = dell0 + dd1 - 3.0 * (ell - cost) / hh;
                        const double disc = theta * theta - dell0 * dd1;
                        if (isfinite(disc) && disc >= 0.0) {
                            const double gm = sqrt(disc);
                            const double den = dd1 - dell0 + 2.0 * gm;
                            if (den != 0.0) {
                                const double cubic = hh - hh * (dd1 + gm - theta) / den;
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
                                trial(root, true, &dd1, &dd2);
                                const double f = dd1 / scale, df = dd2 / scale;
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
                if (lane < n) v += alpha * dvv;
                av_run = av + alpha * adv_ls;
                vc0 += alpha * dvc0; vc1 += alpha * dvc1; vc2 += alpha * dvc2;
            } else {
                av_run = av;  // v unchanged
                if (P.ls_mode == kLsMonotone) {
                    // a rejected decay search keeps v and records the current cost as the previous
                    // one (contact_solve.py:2647-2662): the cost criterion ends the solve next
                    alpha = a_acc;
                    prev_cost = cost;
                    ls_total += ls_it;
                }
            }
            have_av = true;
            ++iteration;
        }
    }

    // training target of the learned warm start: converged v minus the plain start, with
    // free-joint rows in the joint's body frame
    if (P.learn_target) {
        vec[lane] = lane < n ? v - v_plain : 0.0;
        wave_sync();
        double t = lane < n ? vec[lane] : 0.0;
        if (fr_ds >= 0) t = free_rot(vec, fr_ds, fr_k, fq, true);
        if (lane < n) P.learn_target[(size_t)w * n + lane] = (float)t;
        wave_sync();
    }

    // ------------------------------------------------------------------ 9
    // integration (solver_sap.py:288-519) + SAP -> public velocity
    vec[lane] = v0;
    vec[64 + lane] = v;
    for (int k = lane; k < M.NQ; k += 64) q_out[k] = q_in[k];
    wave_sync();
    const double* v0v = vec;
    const double* vv = vec + 64;
    float* qd_pub = kQdPub;
    for (int j = lane; j < nj; j += 64) {
        const int jt = T.jtype[j];
        const int qs = T.jqs[j], ds = T.jds[j], nd = T.jds[j + 1] - T.jds[j];
        if (jt == kJointFree || jt == kJointDistance || jt == kJointBall) {
            // free: (p, quaternion) with [w, v_O]; ball: quaternion with w (solver_sap.py:288-519)
            const bool ball = jt == kJointBall;
            const int qr = ball ? qs : qs + 3;
            const Q4 q = Q4{q_in[qr], q_in[qr + 1], q_in[qr + 2], q_in[qr + 3]};
            const V3 wn = v3((float)vv[ds], (float)vv[ds + 1], (float)vv[ds + 2]);
            const V3 von = ball ? v3(0, 0, 0) : v3((float)vv[ds + 3], (float)vv[ds + 4], (float)vv[ds + 5]);
            const V3 wo = v3((float)v0v[ds], (float)v0v[ds + 1], (float)v0v[ds + 2]);
            const V3 voo = ball ? v3(0, 0, 0) : v3((float)v0v[ds + 3], (float)v0v[ds + 4], (float)v0v[ds + 5]);
            Q4 qn;
            V3 vmid;
            if (P.integ_mode == 0) {
                const V3 wm = (wo + wn) * 0.5f;
                const float ang = length(wm) * h;
                qn = q;
                if (ang > 1.0e-12f) {
                    const V3 ax = wm * (1.0f / length(wm));
                    const float sa = sinf(0.5f * ang), ca = cosf(0.5f * ang);
                    const Q4 r = Q4{ax.x * sa, ax.y * sa, ax.z * sa, ca};
                    Q4 o = qmul(r, q);
                    const float nn = sqrtf(o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w);
                    if (nn > 0.0f) { const float inv = 1.0f / nn; o = Q4{o.x * inv, o.y * inv, o.z * inv, o.w * inv}; }
                    qn = o;
                }
                vmid = (voo + von) * 0.5f;
            } else {
                const float hdl = 0.5f * h;
                float x = q.x + hdl * (q.w * wn.x + wn.y * q.z - wn.z * q.y);
                float y = q.y + hdl * (q.w * wn.y + wn.z * q.x - wn.x * q.z);
                float z = q.z + hdl * (q.w * wn.z + wn.x * q.y - wn.y * q.x);
                float sw = q.w - hdl * (wn.x * q.x + wn.y * q.y + wn.z * q.z);
                const float nn = sqrtf(x * x + y * y + z * z + sw * sw);
                if (nn > 0.0f) { const float inv = 1.0f / nn; x *= inv; y *= inv; z *= inv; sw *= inv; }
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
            const float* com = mCom + T.jchild[j] * 4;
            const V3 rc = qrot(qnormalize(qn), v3(com[0], com[1], com[2]));
            const V3 vcom = von + cross(wn, rc);
            qd_pub[ds] = vcom.x; qd_pub[ds + 1] = vcom.y; qd_pub[ds + 2] = vcom.z;
            qd_pub[ds + 3] = wn.x; qd_pub[ds + 4] = wn.y; qd_pub[ds + 5] = wn.z;
        } else if (jt == kJointFixed) {
            for (int a2 = 0; a2 < nd; ++a2) qd_pub[ds + a2] = (float)vv[ds + a2];
        } else {
            for (int a2 = 0; a2 < nd; ++a2) {
                const float qd = (float)vv[ds + a2];
                const float qm = P.integ_mode == 0 ? 0.5f * ((float)v0v[ds + a2] + qd) : qd;
                q_out[qs + a2] = q_in[qs + a2] + h * qm;
                qd_pub[ds + a2] = qd;
            }
        }
    }
    wave_sync();
    // public forward kinematics (sap_kinematics.py:84-229): world poses in float32 by pointer
    // jumping (each body's local part L = X_p X_j X_c^-1), then the parent-joint frames and
    // the joint velocity terms in parallel, and the body-velocity recursion per level
    float* pbq = bq;
    float* pbqd = kV;
    float* kPl = reinterpret_cast<float*>(kXl);       // nj x 16: X_jc, then (rp, R jl, R ja)
    float* kPB = reinterpret_cast<float*>(kPose);     // nb x 8: partner buffer of pbq
    for (int b = lane; b < nb; b += 64) {
        float* o = kPB + b * 8;
        o[0] = 0.f; o[1] = 0.f; o[2] = 0.f; o[3] = 0.f; o[4] = 0.f; o[5] = 0.f; o[6] = 1.f;
    }
    wave_sync();
    for (int j = lane; j < nj; j += 64) {
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
        store_tf(kPB + T.jchild[j] * 8, tf_mul(load_tf(mXp + j * 8), Xjc));
    }
    wave_sync();
    {
        float* src = kPB;
        float* dst = pbq;
        for (int r = 0; r < M.fk_rounds; ++r) {
            if (lane < nb) {
                const int u = T.body_up[r * nb + lane];
                TF x = load_tf(src + lane * 8);
                if (u >= 0) x = tf_mul(load_tf(src + u * 8), x);
                store_tf(dst + lane * 8, x);
            }
            wave_sync();
            float* t = src;
            src = dst;
            dst = t;
        }
        if (src != pbq) {
            if (lane < nb) store_tf(pbq + lane * 8, load_tf(src + lane * 8));
            wave_sync();
        }
    }
    for (int j = lane; j < nj; j += 64) {
        const int parent = T.jparent[j];
        const TF Xp = load_tf(mXp + j * 8);
        TF Xwpj = Xp;
        V3 rp = v3(0, 0, 0);
        if (parent >= 0) {
            const TF Xwp = load_tf(pbq + parent * 8);
            Xwpj = tf_mul(Xwp, Xp);
            const float* com = mCom + parent * 4;
            rp = Xwpj.p - (Xwp.p + qrot(Xwp.q, v3(com[0], com[1], com[2])));
        }
        float* o = kPl + j * 16;
        const V3 Rjl = qrot(Xwpj.q, v3(o[7], o[8], o[9]));
        const V3 Rja = qrot(Xwpj.q, v3(o[10], o[11], o[12]));
        o[7] = rp.x; o[8] = rp.y; o[9] = rp.z;
        o[10] = Rjl.x; o[11] = Rjl.y; o[12] = Rjl.z;
        o[13] = Rja.x; o[14] = Rja.y; o[15] = Rja.z;
    }
    wave_sync();
    for (int Lv = 0; Lv < M.num_levels; ++Lv) {
        const int l0 = T.level_start[Lv], l1 = T.level_start[Lv + 1];
        for (int t = l0 + lane; t < l1; t += 64) {
            const int j = T.level_joint[t];
            const int parent = T.jparent[j], child = T.jchild[j];
            const float* pl = kPl + j * 16;
            V3 vlin = v3(0, 0, 0), vang = v3(0, 0, 0);
            if (parent >= 0) {
                const float* vp = pbqd + parent * 6;
                vang = v3(vp[3], vp[4], vp[5]);
                vlin = v3(vp[0], vp[1], vp[2]) + cross(vang, v3(pl[7], pl[8], pl[9]));
            }
            const V3 lv = vlin + v3(pl[10], pl[11], pl[12]);
            const V3 av3 = vang + v3(pl[13], pl[14], pl[15]);
            float* ov = pbqd + child * 6;
            ov[0] = lv.x; ov[1] = lv.y; ov[2] = lv.z; ov[3] = av3.x; ov[4] = av3.y; ov[5] = av3.z;
        }
        wave_sync();
    }
    // ------------------------------------------------------------------ outputs
    for (int k = lane; k < M.NQ; k += 64) P.jq_out[(size_t)w * M.NQ + k] = q_out[k];
    if (lane < n) {
        const size_t g = (size_t)w * n + lane;
        P.jqd_out[g] = qd_pub[lane];
        P.v_guess[g] = v;
        if (P.dbg_v) P.dbg_v[g] = v;
        if (P.dbg_vstar) P.dbg_vstar[g] = vs;
        if (P.dbg_v0) P.dbg_v0[g] = v0;
        if (P.dbg_adiag) P.dbg_adiag[g] = adiag;
    }
    for (int t = lane; t < nb * 7; t += 64) {
        const int b = t / 7, k = t % 7;
        P.bq_out[((size_t)w * nb + b) * 7 + k] = pbq[b * 8 + k];
    }
    if (P.bqd_out)
        for (int t = lane; t < nb * 6; t += 64) P.bqd_out[(size_t)w * nb * 6 + t] = pbqd[t];
    if (P.bf_out && P.bf_in && P.bf_out != P.bf_in)
        for (int t = lane; t < nb * 6; t += 64) P.bf_out[(size_t)w * nb * 6 + t] = P.bf_in[(size_t)w * nb * 6 + t];
    if (P.dbg_gamma) {
        for (int t = nc * 3 + lane; t < C * 3; t += 64) P.dbg_gamma[(size_t)w * C * 3 + t] = 0.0;
        if (lane < nc) {
            P.dbg_gamma[(size_t)w * C * 3 + lane * 3] = g0;
            P.dbg_gamma[(size_t)w * C * 3 + lane * 3 + 1] = g1;
            P.dbg_gamma[(size_t)w * C * 3 + lane * 3 + 2] = g2;
        }
    }
    if (lane == 0) {
        P.st_iters[w] = iteration;
        P.st_ls[w] = ls_total;
        P.st_nc[w] = nc;
        P.st_flags[w] = flags | ((ls_status & 0xff) << 8);
        if (P.st_hist) {
            atomicAdd(&P.st_hist[iteration < 255 ? iteration : 255], 1ull);
            atomicAdd(&P.st_hist[256], (unsigned long long)ls_total);
            atomicAdd(&P.st_hist[257], 1ull);
            if (start_kind) atomicAdd(&P.st_hist[start_kind == 1 ? 258 : 259], 1ull);
        }
    }
}

template <int NP, int CB>
__global__ __launch_bounds__(64, 1) void sapx_wave_kernel(StepParams P) {
    wave_body<NP, CB>(P, (int)blockIdx.x);
}

template <int NP, int CB>
hipError_t launch_variant(const StepParams& P, hipStream_t stream) {
    static bool attr_set = false;
    if (!attr_set) {
        (void)hipFuncSetAttribute((const void*)sapx_wave_kernel<NP, CB>, hipFuncAttributeMaxDynamicSharedMemorySize,
                                  160 * 1024);
        attr_set = true;
    }
    hipLaunchKernelGGL((sapx_wave_kernel<NP, CB>), dim3(P.m.W), dim3(64), P.wl.bytes, stream, P);
    return hipGetLastError();
}

}  // namespace wv
}  // namespace sapx

#endif  // SAPX_WAVE_IMPL_H

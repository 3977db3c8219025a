// One world of free rigid bodies on the whole GPU (gfx950): collision, free motion, contact
// rows, the large-world Newton solve and integration.
//
// The step kernels give each world one wave; a single world holding a thousand bodies (a bin
// of boxes, spheres and capsules) spreads every stage over the GPU instead:
//
//   collide       one thread per shape (world pose, AABB), one thread per candidate pair (the
//                 narrowphase and contact writer of the per-world kernel, staged in the pair's
//                 own slots), an exclusive scan of the per-pair counts and one thread per pair
//                 placing its contacts under the per-world capacity rule (sapx_collide.hip)
//   free motion   one thread per free joint: boundary conversions (sap_helpers.py:1482-1522,
//                 1811-1848, 2058-2072), the body's spatial inertia, bias and gravity wrenches
//                 at its origin (sap_helpers.py:1027-1105), tau and the 6 x 6 solve of
//                 v* = v0 + dt A^-1 tau (free_motion.py:2094-2178) in float64; the body's
//                 block of the dense float32 dynamics matrix A
//   contacts      one thread per contact: witnesses, frame, signed distance, materials, the
//                 twelve Jacobian columns of the two bodies, weights and the participating dofs
//                 (contact_jacobian.py:404-562, sap_helpers.py:148-338, contact_solve.py:422-487)
//   solve         sapx_large_solve (sapx_large.hip), started from the previous step's solution
//                 after the first step (warm start; sapx_bigworld_reset clears it)
//   integration   one thread per free joint (solver_sap.py:288-519); the public body pose and
//                 velocity of a root free body are its coordinates (sap_kinematics.py:84-229)
//
// The model must hold one world whose joints are all free joints rooted at the world with
// identity joint frames, one per body, and no drives or limits.
#include <hip/hip_runtime.h>
#include <hipcub/hipcub.hpp>

#include <cmath>
#include <cstdint>
#include <string>
#include <vector>

#include "../../include/sapx.h"
#include "sapx_device.h"
#include "sapx_params.h"

namespace sapx {
void set_last_error(const char* msg);
}

namespace {

using namespace sapx;

constexpr int NT = 256;
constexpr int KB = 12;  // Jacobian columns per contact: the six dofs of each body

int32_t bfail(int32_t code, const std::string& msg) {
    sapx::set_last_error(msg.c_str());
    return code;
}

struct BwModel {
    int J, NB, n;
    const int* jinfo;  // J x 3: child body, first coordinate, first dof
    const int* bdof;   // NB: first dof of the body's free joint
    const int* brange; // NB x 2: the body's articulation dofs (start, count; start < 0: none)
    const float* body_mass;
    const float* body_inertia;  // x 9
    const float* body_com;      // x 3
    const float* armature;
    const float* gravity;
    const int* shape_body;
    const float* shape_mu;
    const float* shape_ke;
    const float* shape_tau;
};

__device__ __forceinline__ D3 ld3(const float* a) { return d3(a[0], a[1], a[2]); }
__device__ __forceinline__ double ddot(D3 a, D3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ D3 dqrot_n(DQ q, D3 v) { return dqrot_unit(dqnormalize(q), v); }

// ---------------------------------------------------------------------------
// free motion of one free joint (thread per joint), float64
// ---------------------------------------------------------------------------
struct FreeArgs {
    BwModel m;
    const float* jq;
    const float* jqd;
    const float* bq;   // state body poses (the lever arms of the boundary conversions)
    const float* bf;   // public body wrenches [f, tau] at the COM, may be null
    const float* jf;   // public joint forces, may be null
    double dt;
    float* A;          // n x n, off-diagonal blocks stay zero
    double* adiag;     // n
    double* v0;
    double* vstar;
    double* pose;      // NB x 8: body pose of the free motion (the joint coordinates)
    int* part;
};

__global__ __launch_bounds__(NT) void bw_free(FreeArgs a) {
    const int j = blockIdx.x * NT + threadIdx.x;
    if (j >= a.m.J) return;
    const int c = a.m.jinfo[3 * j], qs = a.m.jinfo[3 * j + 1], ds = a.m.jinfo[3 * j + 2];
    const float* q = a.jq + qs;
    const D3 p = ld3(q);
    const DQ R{q[3], q[4], q[5], q[6]};
    double* po = a.pose + (size_t)c * 8;
    po[0] = p.x; po[1] = p.y; po[2] = p.z; po[3] = R.x; po[4] = R.y; po[5] = R.z; po[6] = R.w; po[7] = 0.0;
    const D3 com = ld3(a.m.body_com + (size_t)c * 3);
    const float* bqs = a.bq + (size_t)c * 7;
    const D3 rs = dqrot_n(DQ{bqs[3], bqs[4], bqs[5], bqs[6]}, com);  // COM offset at the state pose
    // SAP velocity [w, v_O] from the public [v_com, w]
    const D3 vcom = ld3(a.jqd + ds), w = ld3(a.jqd + ds + 3);
    const D3 vo = dsub(vcom, dcross(w, rs));
    double v0[6] = {w.x, w.y, w.z, vo.x, vo.y, vo.z};
    // generalized force: joint force [f, t] at the COM -> [t + r x f, f]; body wrench negated
    double fs[6] = {0, 0, 0, 0, 0, 0};
    if (a.jf) {
        const D3 f = ld3(a.jf + ds), t = ld3(a.jf + ds + 3);
        const D3 tt = dadd(t, dcross(rs, f));
        fs[0] = tt.x; fs[1] = tt.y; fs[2] = tt.z; fs[3] = f.x; fs[4] = f.y; fs[5] = f.z;
    }
    double fb[6] = {0, 0, 0, 0, 0, 0};
    if (a.bf) {
        const D3 f = ld3(a.bf + (size_t)c * 6), t = ld3(a.bf + (size_t)c * 6 + 3);
        const D3 tt = dadd(t, dcross(rs, f));
        fb[0] = -tt.x; fb[1] = -tt.y; fb[2] = -tt.z; fb[3] = -f.x; fb[4] = -f.y; fb[5] = -f.z;
    }
    // spatial inertia at the body origin, world frame, from the (unnormalised) coordinate rotation
    const double m = a.m.body_mass[c];
    const float* Ic = a.m.body_inertia + (size_t)c * 9;
    const D3 e0 = dqrot_unit(R, d3(1, 0, 0)), e1 = dqrot_unit(R, d3(0, 1, 0)), e2 = dqrot_unit(R, d3(0, 0, 1));
    const double Rm[3][3] = {{e0.x, e1.x, e2.x}, {e0.y, e1.y, e2.y}, {e0.z, e1.z, e2.z}};
    double RI[3][3], IO[3][3];
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int l = 0; l < 3; ++l) s += Rm[r][l] * (double)Ic[3 * l + k];
            RI[r][k] = s;
        }
    const D3 cw = dqrot_unit(R, com);
    const double cc[3] = {cw.x, cw.y, cw.z};
    const double c2 = ddot(cw, cw);
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            double s = 0.0;
            for (int l = 0; l < 3; ++l) s += RI[r][l] * Rm[k][l];
            IO[r][k] = s + m * ((r == k ? c2 : 0.0) - cc[r] * cc[k]);
        }
    double M[6][6];
    const double sk[3][3] = {{0.0, -cw.z, cw.y}, {cw.z, 0.0, -cw.x}, {-cw.y, cw.x, 0.0}};
    for (int r = 0; r < 3; ++r)
        for (int k = 0; k < 3; ++k) {
            M[r][k] = IO[r][k];
            M[r][3 + k] = m * sk[r][k];
            M[3 + r][k] = -m * sk[r][k];
            M[3 + r][3 + k] = r == k ? m : 0.0;
        }
    // bias [w x I_O w, m w x (w x c)] minus gravity [c x m g, m g] (the velocity of the body is v0)
    const D3 Iw = d3(IO[0][0] * w.x + IO[0][1] * w.y + IO[0][2] * w.z, IO[1][0] * w.x + IO[1][1] * w.y + IO[1][2] * w.z,
                     IO[2][0] * w.x + IO[2][1] * w.y + IO[2][2] * w.z);
    const D3 b0 = dcross(w, Iw);
    const D3 b1 = dscale(dcross(w, dcross(w, cw)), m);
    const D3 mg = dscale(ld3(a.m.gravity), m);
    const D3 g0 = dcross(cw, mg);
    const double fbias[6] = {b0.x - g0.x, b0.y - g0.y, b0.z - g0.z, b1.x - mg.x, b1.y - mg.y, b1.z - mg.z};
    double tau[6];
    for (int k = 0; k < 6; ++k) tau[k] = -(fbias[k] + fb[k]) + fs[k];
    // A = M + diag(armature) (M symmetric by construction); Cholesky solve of A x = tau
    double L[6][6];
    for (int r = 0; r < 6; ++r)
        for (int k = 0; k < 6; ++k) {
            const double v = M[r][k] + (r == k ? (double)a.m.armature[ds + r] : 0.0);
            L[r][k] = v;
            a.A[(size_t)(ds + r) * a.m.n + ds + k] = (float)v;
        }
    for (int r = 0; r < 6; ++r) a.adiag[ds + r] = L[r][r];
    for (int k = 0; k < 6; ++k) {
        double d = L[k][k];
        for (int l = 0; l < k; ++l) d -= L[k][l] * L[k][l];
        d = sqrt(fmax(d, 1.0e-300));
        L[k][k] = d;
        for (int r = k + 1; r < 6; ++r) {
            double s = L[r][k];
            for (int l = 0; l < k; ++l) s -= L[r][l] * L[k][l];
            L[r][k] = s / d;
        }
    }
    double y[6];
    for (int r = 0; r < 6; ++r) {
        double s = tau[r];
        for (int l = 0; l < r; ++l) s -= L[r][l] * y[l];
        y[r] = s / L[r][r];
    }
    for (int r = 5; r >= 0; --r) {
        double s = y[r];
        for (int l = r + 1; l < 6; ++l) s -= L[l][r] * y[l];
        y[r] = s / L[r][r];
    }
    for (int k = 0; k < 6; ++k) {
        a.v0[ds + k] = v0[k];
        a.vstar[ds + k] = v0[k] + a.dt * y[k];
        a.part[ds + k] = 0;
    }
}

// ---------------------------------------------------------------------------
// contact rows (thread per contact), float64
// ---------------------------------------------------------------------------
struct RowArgs {
    BwModel m;
    int nc;
    const int* shape0;
    const int* shape1;
    const float* point0;
    const float* point1;
    const float* normal;
    const float* margin0;
    const float* margin1;
    const double* pose;
    const double* adiag;
    double fallback_mu, fallback_k, fallback_tau;
    int weight_mode, point_mode;
    int* part;
    int* cols;      // nc x KB
    float* vals;    // nc x 3 x KB
    double* phi0;
    double* weight;
    double* stiff;
    double* tau;
    double* mu;
};

__device__ __forceinline__ double material_k(const RowArgs& a, int sh) {
    double k = a.fallback_k;
    if (sh >= 0) k = a.m.shape_ke[sh];
    if (!isfinite(k) || k <= 0.0) k = a.fallback_k;
    if (!isfinite(k) || k <= 0.0) k = 1.0;
    return k;
}
__device__ __forceinline__ double material_tau(const RowArgs& a, int sh) {
    double t = a.fallback_tau;
    if (sh >= 0) t = a.m.shape_tau[sh];
    if (!isfinite(t) || t < 0.0) t = a.fallback_tau;
    if (!isfinite(t) || t < 0.0) t = 0.0;
    return t;
}
__device__ __forceinline__ D3 safe_unit(D3 v) {
    const double n2 = ddot(v, v);
    return n2 <= 1.0e-32 ? d3(0.0, 0.0, 0.0) : dscale(v, 1.0 / sqrt(n2));
}
// 1/m + |I^-1/2 (R^T (r x d))|^2 of one body (sap_helpers.py:307-338)
__device__ __forceinline__ double direction_weight(const BwModel& m, int b, DQ q, D3 r, D3 d) {
    const double mass = m.body_mass[b];
    if (mass <= 1.0e-12) return 0.0;
    const D3 cwd = dcross(r, d);
    const double n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    const DQ qi{-q.x / n2, -q.y / n2, -q.z / n2, q.w / n2};
    const D3 cb = dqrot_n(qi, cwd);
    const float* I = m.body_inertia + (size_t)b * 9;
    const double ix = fmax((double)I[0], 1.0e-12), iy = fmax((double)I[4], 1.0e-12), iz = fmax((double)I[8], 1.0e-12);
    return 1.0 / mass + cb.x * cb.x / ix + cb.y * cb.y / iy + cb.z * cb.z / iz;
}

__global__ __launch_bounds__(NT) void bw_rows(RowArgs a) {
    const int i = blockIdx.x * NT + threadIdx.x;
    if (i >= a.nc) return;
    const int sh0 = a.shape0[i], sh1 = a.shape1[i];
    const int b0 = sh0 >= 0 ? a.m.shape_body[sh0] : -1;
    const int b1 = sh1 >= 0 ? a.m.shape_body[sh1] : -1;
    D3 x0 = ld3(a.point0 + 3 * (size_t)i), x1 = ld3(a.point1 + 3 * (size_t)i);
    D3 p0 = d3(0, 0, 0), p1 = p0;
    DQ q0{0, 0, 0, 1}, q1{0, 0, 0, 1};
    if (b0 >= 0) {
        const double* P = a.pose + (size_t)b0 * 8;
        p0 = d3(P[0], P[1], P[2]);
        q0 = DQ{P[3], P[4], P[5], P[6]};
        x0 = dadd(p0, dqrot_n(q0, x0));
    }
    if (b1 >= 0) {
        const double* P = a.pose + (size_t)b1 * 8;
        p1 = d3(P[0], P[1], P[2]);
        q1 = DQ{P[3], P[4], P[5], P[6]};
        x1 = dadd(p1, dqrot_n(q1, x1));
    }
    const D3 nz = safe_unit(ld3(a.normal + 3 * (size_t)i));
    // MakeFromOneUnitVector with z = n (sap_helpers.py:467-503)
    const double ax = fabs(nz.x), ay = fabs(nz.y), az = fabs(nz.z);
    int im = 0;
    double mn = ax;
    if (ay < mn) { im = 1; mn = ay; }
    if (az < mn) { im = 2; mn = az; }
    const double mag = sqrt(1.0 - mn * mn);
    const double rr = 1.0 / mag;
    const double u = im == 0 ? nz.x : (im == 1 ? nz.y : nz.z);
    const double sg = -rr * u;
    D3 tx, ty;
    if (im == 0) { tx = d3(0.0, -rr * nz.z, rr * nz.y); ty = d3(mag, sg * nz.y, sg * nz.z); }
    else if (im == 1) { tx = d3(rr * nz.z, 0.0, -rr * nz.x); ty = d3(sg * nz.x, mag, sg * nz.z); }
    else { tx = d3(-rr * nz.y, rr * nz.x, 0.0); ty = d3(sg * nz.x, sg * nz.y, mag); }
    // materials (sap_helpers.py:148-237)
    const double k0 = material_k(a, sh0), k1 = material_k(a, sh1);
    const double kden = k0 + k1;
    double w0 = 0.5, w1 = 0.5;
    if (kden != 0.0) { w0 = k0 / kden; w1 = k1 / kden; }
    const D3 pc = dadd(dscale(x0, w0), dscale(x1, w1));
    double kc;
    if (!isfinite(k0)) kc = k1;
    else if (!isfinite(k1)) kc = k0;
    else if (kden <= 1.0e-12) kc = fmax(k0, k1);
    else kc = k0 * k1 / kden;
    const double mu0 = fmax(sh0 >= 0 ? (double)a.m.shape_mu[sh0] : a.fallback_mu, 0.0);
    const double mu1 = fmax(sh1 >= 0 ? (double)a.m.shape_mu[sh1] : a.fallback_mu, 0.0);
    const double mud = mu0 + mu1;
    a.mu[i] = mud <= 1.0e-12 ? 0.0 : 2.0 * mu0 * mu1 / mud;
    a.stiff[i] = kc;
    a.tau[i] = material_tau(a, sh0) + material_tau(a, sh1);
    a.phi0[i] = ddot(dsub(x1, x0), nz) - (double)a.margin0[i] - (double)a.margin1[i];
    // rows t1, t2, n over each body's [w, v_O] columns: sign * [r x d, d] with r = x - O
    const D3 xj0 = a.point_mode == 0 ? x0 : pc;
    const D3 xj1 = a.point_mode == 0 ? x1 : pc;
    const D3 dirs[3] = {tx, ty, nz};
    double V[3][KB];
    int C[KB];
    for (int side = 0; side < 2; ++side) {
        const int b = side == 0 ? b0 : b1;
        const double sgn = side == 0 ? -1.0 : 1.0;
        const D3 r = side == 0 ? dsub(xj0, p0) : dsub(xj1, p1);
        for (int k = 0; k < 6; ++k) C[6 * side + k] = b >= 0 ? a.m.bdof[b] + k : -1;
        for (int rw = 0; rw < 3; ++rw) {
            const D3 rxd = dcross(r, dirs[rw]);
            const double e[6] = {rxd.x, rxd.y, rxd.z, dirs[rw].x, dirs[rw].y, dirs[rw].z};
            for (int k = 0; k < 6; ++k) V[rw][6 * side + k] = b >= 0 ? sgn * e[k] : 0.0;
        }
    }
    if (b0 >= 0 && b0 == b1) {  // both sides on one body: one set of columns
        for (int rw = 0; rw < 3; ++rw)
            for (int k = 0; k < 6; ++k) {
                V[rw][6 + k] += V[rw][k];
                V[rw][k] = 0.0;
            }
        for (int k = 0; k < 6; ++k) C[k] = -1;
    }
    for (int k = 0; k < KB; ++k) {
        a.cols[(size_t)i * KB + k] = C[k];
        for (int rw = 0; rw < 3; ++rw) a.vals[((size_t)i * 3 + rw) * KB + k] = (float)V[rw][k];
        if (C[k] >= 0 && (V[0][k] != 0.0 || V[1][k] != 0.0 || V[2][k] != 0.0)) a.part[C[k]] = 1;
    }
    // the whole articulation of each touched body takes part (contact_solve.py:422-487)
    for (int side = 0; side < 2; ++side) {
        const int b = side == 0 ? b0 : b1;
        if (b < 0 || a.m.brange[2 * b] < 0) continue;
        for (int k = 0; k < a.m.brange[2 * b + 1]; ++k) a.part[a.m.brange[2 * b] + k] = 1;
    }
    // effective weight (contact_jacobian.py:153-251)
    double wv;
    if (a.weight_mode == SAPX_WEIGHT_BODY_INERTIA) {
        double acc = 0.0;
        for (int side = 0; side < 2; ++side) {
            const int b = side == 0 ? b0 : b1;
            if (b < 0) continue;
            const D3 P = side == 0 ? p0 : p1;
            const DQ Q = side == 0 ? q0 : q1;
            const D3 comw = dadd(P, dqrot_n(Q, ld3(a.m.body_com + (size_t)b * 3)));
            const D3 r = dsub(side == 0 ? x0 : x1, comw);
            for (int k = 0; k < 3; ++k) acc += direction_weight(a.m, b, Q, r, safe_unit(dirs[k]));
        }
        wv = acc / 3.0;
    } else {
        double W2 = 0.0;
        for (int r = 0; r < 3; ++r)
            for (int s = 0; s < 3; ++s) {
                double x = 0.0;
                for (int k = 0; k < KB; ++k) {
                    if (C[k] < 0) continue;
                    const double ad = a.adiag[C[k]];
                    x += V[r][k] * V[s][k] / fmax(ad, 1.0e-12);
                }
                W2 += x * x;
            }
        wv = sqrt(W2) / 3.0;
    }
    a.weight[i] = fmax(wv, 1.0e-12);
}

// ---------------------------------------------------------------------------
// integration of one free joint (thread per joint), float64; public outputs float32
// ---------------------------------------------------------------------------
struct IntArgs {
    BwModel m;
    const float* jq;
    const double* v0;
    const double* v;
    double dt;
    int mode;
    float* jq_out;
    float* jqd_out;
    float* bq_out;
    float* bqd_out;
};

__global__ __launch_bounds__(NT) void bw_integrate(IntArgs a) {
    const int j = blockIdx.x * NT + threadIdx.x;
    if (j >= a.m.J) return;
    const int c = a.m.jinfo[3 * j], qs = a.m.jinfo[3 * j + 1], ds = a.m.jinfo[3 * j + 2];
    const float* qi = a.jq + qs;
    const D3 p = ld3(qi);
    const DQ q{qi[3], qi[4], qi[5], qi[6]};
    const double* vp = a.v0 + ds;
    const double* vn = a.v + ds;
    const D3 wn = d3(vn[0], vn[1], vn[2]), von = d3(vn[3], vn[4], vn[5]);
    const double h = a.dt;
    D3 vom;
    DQ qn;
    if (a.mode == SAPX_INTEGRATE_MIDPOINT) {
        // solver_sap.py:227-244: rotate by the midpoint angular velocity, renormalise
        const D3 wm = d3(0.5 * (vp[0] + vn[0]), 0.5 * (vp[1] + vn[1]), 0.5 * (vp[2] + vn[2]));
        vom = d3(0.5 * (vp[3] + vn[3]), 0.5 * (vp[4] + vn[4]), 0.5 * (vp[5] + vn[5]));
        const double wl = sqrt(ddot(wm, wm));
        const double angle = wl * h;
        qn = q;
        if (angle > 1.0e-12) {
            const D3 ax = dscale(wm, 1.0 / wl);
            double s, cs;
            sincos(0.5 * angle, &s, &cs);
            const DQ r{ax.x * s, ax.y * s, ax.z * s, cs};
            DQ o = dqmul(r, q);
            const double nrm = sqrt(o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w);
            if (nrm > 0.0) o = DQ{o.x / nrm, o.y / nrm, o.z / nrm, o.w / nrm};
            qn = o;
        }
    } else {
        // solver_sap.py:161-181
        vom = von;
        const double hd = 0.5 * h;
        double x = q.x + hd * (q.w * wn.x + wn.y * q.z - wn.z * q.y);
        double y = q.y + hd * (q.w * wn.y + wn.z * q.x - wn.x * q.z);
        double z = q.z + hd * (q.w * wn.z + wn.x * q.y - wn.y * q.x);
        double s = q.w - hd * (wn.x * q.x + wn.y * q.y + wn.z * q.z);
        const double nrm = sqrt(x * x + y * y + z * z + s * s);
        if (nrm > 0.0) { x /= nrm; y /= nrm; z /= nrm; s /= nrm; }
        qn = DQ{x, y, z, s};
    }
    const D3 po = dadd(p, dscale(vom, h));
    const D3 r = dqrot_n(qn, ld3(a.m.body_com + (size_t)c * 3));
    const D3 vc = dadd(von, dcross(wn, r));
    const float qo[7] = {(float)po.x, (float)po.y, (float)po.z, (float)qn.x, (float)qn.y, (float)qn.z, (float)qn.w};
    const float qdo[6] = {(float)vc.x, (float)vc.y, (float)vc.z, (float)wn.x, (float)wn.y, (float)wn.z};
    for (int k = 0; k < 7; ++k) a.jq_out[qs + k] = qo[k];
    for (int k = 0; k < 6; ++k) a.jqd_out[ds + k] = qdo[k];
    if (a.bq_out)
        for (int k = 0; k < 7; ++k) a.bq_out[(size_t)c * 7 + k] = qo[k];
    if (a.bqd_out)
        for (int k = 0; k < 6; ++k) a.bqd_out[(size_t)c * 6 + k] = qdo[k];
}

}  // namespace

struct SapxBigWorld {
    const SapxModel* model = nullptr;
    SapxSolverOpts opts{};
    int J = 0, NB = 0, n = 0, S = 0, P = 0, C = 0;
    BwModel bm{};
    int *jinfo = nullptr, *bdof = nullptr, *brange = nullptr, *pairs = nullptr;
    // collide scratch
    float *X = nullptr, *AB = nullptr, *st_f = nullptr;
    int *cnt = nullptr, *cvx = nullptr, *off = nullptr, *st_sh = nullptr, *unsupported = nullptr;
    void* scan_tmp = nullptr;
    size_t scan_bytes = 0;
    // step
    float* A = nullptr;
    double *adiag = nullptr, *v0 = nullptr, *vstar = nullptr, *pose = nullptr, *v = nullptr;
    int* part = nullptr;
    int* cols = nullptr;
    float* vals = nullptr;
    double *phi0 = nullptr, *weight = nullptr, *stiff = nullptr, *tau = nullptr, *mu = nullptr;
    int* hcount = nullptr;  // pinned
    SapxLarge* large = nullptr;
    int has_guess = 0;      // the previous step's v (in v) starts the next solve
};

extern "C" {

void sapx_bigworld_destroy(SapxBigWorld* b) {
    if (!b) return;
    sapx_large_destroy(b->large);
    for (void* p : {(void*)b->jinfo, (void*)b->bdof, (void*)b->brange, (void*)b->pairs, (void*)b->X, (void*)b->AB,
                    (void*)b->st_f, (void*)b->cnt, (void*)b->cvx, (void*)b->off, (void*)b->st_sh, (void*)b->unsupported,
                    b->scan_tmp, (void*)b->A, (void*)b->adiag, (void*)b->v0, (void*)b->vstar, (void*)b->pose,
                    (void*)b->v, (void*)b->part, (void*)b->cols, (void*)b->vals, (void*)b->phi0, (void*)b->weight,
                    (void*)b->stiff, (void*)b->tau, (void*)b->mu})
        if (p) (void)hipFree(p);
    if (b->hcount) (void)hipHostFree(b->hcount);
    delete b;
}

int32_t sapx_bigworld_create(const SapxModel* model, const SapxSolverOpts* opts, SapxBigWorld** out) {
    if (!model || !opts || !out) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_create: null argument");
    *out = nullptr;
    const DevModel& dm = sapx::model_dev(model);
    if (dm.W != 1) return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: the model must hold exactly one world");
    if (opts->max_rigid_contact <= 0) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_create: max_rigid_contact must be positive");
    if (opts->line_search != SAPX_LS_EXACT_ROOT && opts->line_search != SAPX_LS_ARMIJO_DECAY)
        return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: line searches exact_root and armijo_decay only");
    const int J = dm.J, NB = dm.NB, n = dm.N;
    if (J <= 0 || J != NB || n != 6 * J || dm.NQ != 7 * J)
        return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: one free joint per body required");
    // host copies of the joint template and the drive / limit tables
    std::vector<int> jt(J), jp(J), jc(J), jqs(J + 1), jds(J + 1), xci(J), tmode(n), lim_ok(n), art_s(NB), art_c(NB);
    std::vector<float> xp(7 * (size_t)J), tke(n), tkd(n), lke(n);
    std::vector<int> tp(2 * (size_t)dm.P);
    auto get = [](void* dst, const void* src, size_t bytes) {
        return bytes == 0 || hipMemcpy(dst, src, bytes, hipMemcpyDeviceToHost) == hipSuccess;
    };
    bool ok = get(jt.data(), dm.t_jtype, J * 4) && get(jp.data(), dm.t_jparent, J * 4) &&
              get(jc.data(), dm.t_jchild, J * 4) && get(jqs.data(), dm.t_jqs, (J + 1) * 4) &&
              get(jds.data(), dm.t_jds, (J + 1) * 4) && get(xci.data(), dm.j_xc_identity, J * 4) &&
              get(xp.data(), dm.joint_X_p, 7 * (size_t)J * 4) && get(tmode.data(), dm.target_mode, n * 4) &&
              get(tke.data(), dm.target_ke, n * 4) && get(tkd.data(), dm.target_kd, n * 4) &&
              get(lim_ok.data(), dm.t_dof_limit_ok, n * 4) && get(lke.data(), dm.limit_ke, n * 4) &&
              get(art_s.data(), dm.t_body_art_start, NB * 4) && get(art_c.data(), dm.t_body_art_count, NB * 4) &&
              get(tp.data(), dm.t_pairs, tp.size() * 4);
    if (!ok) return bfail(SAPX_ERR_HIP, "sapx_bigworld_create: reading the model failed");
    std::vector<int> jinfo(3 * (size_t)J), bdof(NB, -1), brange(2 * (size_t)NB);
    for (int j = 0; j < J; ++j) {
        const float* x = xp.data() + 7 * (size_t)j;
        const bool xp_ident = x[0] == 0.f && x[1] == 0.f && x[2] == 0.f && x[3] == 0.f && x[4] == 0.f && x[5] == 0.f &&
                              x[6] == 1.f;
        if (jt[j] != kJointFree || jp[j] != -1 || !xp_ident || !xci[j] || jds[j + 1] - jds[j] != 6 ||
            jqs[j + 1] - jqs[j] != 7)
            return bfail(SAPX_ERR_UNSUPPORTED,
                         "sapx_bigworld_create: every joint must be a free joint rooted at the world with identity frames");
        if (jc[j] < 0 || jc[j] >= NB || bdof[jc[j]] >= 0)
            return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: one free joint per body required");
        bdof[jc[j]] = jds[j];
        jinfo[3 * j] = jc[j];
        jinfo[3 * j + 1] = jqs[j];
        jinfo[3 * j + 2] = jds[j];
    }
    for (int d = 0; d < n; ++d) {
        if (tmode[d] != 0 && (tke[d] > 0.f || tkd[d] > 0.f))
            return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: joint drives are not supported in a large world");
        if (lim_ok[d] && lke[d] > 0.f)
            return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: joint limits are not supported in a large world");
    }
    for (int b = 0; b < NB; ++b) {
        brange[2 * b] = art_s[b];
        brange[2 * b + 1] = art_c[b];
    }
    // the explicit candidate pairs, as shape indices (the template holds a global shape s as -s-1);
    // a pair of two static shapes would give contacts outside every body's world
    const int P = dm.P;
    for (int k = 0; k < 2 * P; ++k) tp[k] = tp[k] >= 0 ? tp[k] : -tp[k] - 1;
    std::vector<int> sbody(dm.total_shapes);
    if (!get(sbody.data(), dm.shape_body, sbody.size() * 4))
        return bfail(SAPX_ERR_HIP, "sapx_bigworld_create: reading the model failed");
    for (int k = 0; k < P; ++k) {
        const int a = tp[2 * k], c = tp[2 * k + 1];
        if (a < 0 || c < 0 || a >= dm.total_shapes || c >= dm.total_shapes)
            return bfail(SAPX_ERR_INVALID, "sapx_bigworld_create: pair references a bad shape");
        if (sbody[a] < 0 && sbody[c] < 0)
            return bfail(SAPX_ERR_UNSUPPORTED, "sapx_bigworld_create: a candidate pair of two static shapes");
    }

    SapxBigWorld* b = new SapxBigWorld();
    b->model = model;
    b->opts = *opts;
    b->J = J;
    b->NB = NB;
    b->n = n;
    b->S = dm.total_shapes;
    b->P = P;
    b->C = opts->max_rigid_contact;
    const size_t Pc = P > 0 ? P : 1, C = b->C, S = b->S;
    ok = hipMalloc(&b->jinfo, jinfo.size() * 4) == hipSuccess && hipMalloc(&b->bdof, NB * 4) == hipSuccess &&
         hipMalloc(&b->brange, 2 * (size_t)NB * 4) == hipSuccess && hipMalloc(&b->pairs, 2 * Pc * 4) == hipSuccess &&
         hipMalloc(&b->X, S * 8 * 4) == hipSuccess && hipMalloc(&b->AB, S * 6 * 4) == hipSuccess &&
         hipMalloc(&b->cnt, Pc * 4) == hipSuccess && hipMalloc(&b->cvx, Pc * 4) == hipSuccess &&
         hipMalloc(&b->off, Pc * 4) == hipSuccess && hipMalloc(&b->st_sh, Pc * kBigSlots * 2 * 4) == hipSuccess &&
         hipMalloc(&b->st_f, Pc * kBigSlots * 11 * 4) == hipSuccess && hipMalloc(&b->unsupported, 4) == hipSuccess &&
         hipMalloc(&b->A, (size_t)n * n * 4) == hipSuccess && hipMalloc(&b->adiag, n * 8) == hipSuccess &&
         hipMalloc(&b->v0, n * 8) == hipSuccess && hipMalloc(&b->vstar, n * 8) == hipSuccess &&
         hipMalloc(&b->pose, (size_t)NB * 8 * 8) == hipSuccess && hipMalloc(&b->v, n * 8) == hipSuccess &&
         hipMalloc(&b->part, n * 4) == hipSuccess && hipMalloc(&b->cols, C * KB * 4) == hipSuccess &&
         hipMalloc(&b->vals, C * 3 * KB * 4) == hipSuccess && hipMalloc(&b->phi0, C * 8) == hipSuccess &&
         hipMalloc(&b->weight, C * 8) == hipSuccess && hipMalloc(&b->stiff, C * 8) == hipSuccess &&
         hipMalloc(&b->tau, C * 8) == hipSuccess && hipMalloc(&b->mu, C * 8) == hipSuccess &&
         hipHostMalloc(&b->hcount, 4) == hipSuccess;
    ok = ok && hipMemcpy(b->jinfo, jinfo.data(), jinfo.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->bdof, bdof.data(), NB * 4, hipMemcpyHostToDevice) == hipSuccess &&
         hipMemcpy(b->brange, brange.data(), brange.size() * 4, hipMemcpyHostToDevice) == hipSuccess &&
         (P == 0 || hipMemcpy(b->pairs, tp.data(), 2 * (size_t)P * 4, hipMemcpyHostToDevice) == hipSuccess) &&
         hipMemset(b->A, 0, (size_t)n * n * 4) == hipSuccess && hipMemset(b->unsupported, 0, 4) == hipSuccess;
    if (ok && P > 0) {
        ok = hipcub::DeviceScan::ExclusiveSum(nullptr, b->scan_bytes, b->cnt, b->off, P) == hipSuccess &&
             hipMalloc(&b->scan_tmp, b->scan_bytes > 0 ? b->scan_bytes : 1) == hipSuccess;
    }
    ok = ok && hipDeviceSynchronize() == hipSuccess;
    if (!ok) {
        sapx_bigworld_destroy(b);
        return bfail(SAPX_ERR_HIP, "sapx_bigworld_create: device allocation failed");
    }
    b->bm = BwModel{J, NB, n, b->jinfo, b->bdof, b->brange, dm.body_mass, dm.body_inertia, dm.body_com, dm.armature,
                    dm.gravity, dm.shape_body, dm.shape_mu, dm.shape_ke, dm.shape_tau};
    const int32_t rc = sapx_large_create(n, b->C, KB, &b->large);
    if (rc != SAPX_OK) {
        sapx_bigworld_destroy(b);
        return rc;
    }
    *out = b;
    return SAPX_OK;
}

void sapx_bigworld_reset(SapxBigWorld* b) {
    if (b) b->has_guess = 0;
}

int32_t sapx_bigworld_collide(SapxBigWorld* b, const float* body_q, const SapxContactsView* c, void* stream) {
    if (!b || !body_q || !c) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_collide: null argument");
    if (!c->count || !c->world_start || !c->shape0 || !c->shape1 || !c->point0 || !c->point1 || !c->normal ||
        !c->margin0 || !c->margin1)
        return bfail(SAPX_ERR_INVALID, "sapx_bigworld_collide: contact view needs every array");
    if (c->capacity < b->C)
        return bfail(SAPX_ERR_CAPACITY, "sapx_bigworld_collide: contact capacity below max_rigid_contact");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const DevModel& dm = sapx::model_dev(b->model);
    BigCollideParams B{};
    B.nshapes = b->S;
    B.P = b->P;
    B.C = b->C;
    B.shape_type = dm.shape_type;
    B.shape_body = dm.shape_body;
    B.shape_transform = dm.shape_transform;
    B.shape_scale = dm.shape_scale;
    B.shape_radius = dm.shape_radius;
    B.shape_margin = dm.shape_margin;
    B.shape_gap = dm.shape_gap;
    B.pairs = b->pairs;
    B.body_q = body_q;
    B.X = b->X;
    B.AB = b->AB;
    B.cnt = b->cnt;
    B.cvx = b->cvx;
    B.off = b->off;
    B.st_sh = b->st_sh;
    B.st_f = b->st_f;
    B.unsupported = b->unsupported;
    B.out_shape0 = c->shape0;
    B.out_shape1 = c->shape1;
    B.out_p0 = c->point0;
    B.out_p1 = c->point1;
    B.out_n = c->normal;
    B.out_m0 = c->margin0;
    B.out_m1 = c->margin1;
    B.out_pair = nullptr;
    B.world_start = c->world_start;
    B.count = c->count;
    if (launch_big_pairs(B, st) != hipSuccess) return bfail(SAPX_ERR_HIP, "sapx_bigworld_collide: launch failed");
    if (b->P > 0 &&
        hipcub::DeviceScan::ExclusiveSum(b->scan_tmp, b->scan_bytes, b->cnt, b->off, b->P, st) != hipSuccess)
        return bfail(SAPX_ERR_HIP, "sapx_bigworld_collide: scan failed");
    if (launch_big_write(B, st) != hipSuccess) return bfail(SAPX_ERR_HIP, "sapx_bigworld_collide: launch failed");
    return SAPX_OK;
}

int32_t sapx_bigworld_step(SapxBigWorld* b, const SapxStateView* in, const SapxStateView* out,
                           const SapxControlView* control, const SapxContactsView* c, float dt, double* v_out,
                           double* gamma_out, SapxLargeStats* stats, void* stream) {
    if (!b || !in || !out || !c || !stats) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_step: null argument");
    if (!in->joint_q || !in->joint_qd || !in->body_q || !out->joint_q || !out->joint_qd)
        return bfail(SAPX_ERR_INVALID, "sapx_bigworld_step: state views need joint_q, joint_qd (and body_q in)");
    if (!c->count || !c->shape0 || !c->shape1 || !c->point0 || !c->point1 || !c->normal || !c->margin0 || !c->margin1)
        return bfail(SAPX_ERR_INVALID, "sapx_bigworld_step: contact view needs every array");
    if (!(dt > 0.0f)) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_step: dt must be positive");
    hipStream_t st = reinterpret_cast<hipStream_t>(stream);
    const SapxSolverOpts& o = b->opts;
    FreeArgs fa{b->bm, in->joint_q, in->joint_qd, in->body_q, in->body_f, control ? control->joint_f : nullptr,
                (double)dt, b->A, b->adiag, b->v0, b->vstar, b->pose, b->part};
    hipLaunchKernelGGL(bw_free, dim3((b->J + NT - 1) / NT), dim3(NT), 0, st, fa);
    if (hipMemcpyAsync(b->hcount, c->count, 4, hipMemcpyDeviceToHost, st) != hipSuccess ||
        hipStreamSynchronize(st) != hipSuccess)
        return bfail(SAPX_ERR_HIP, "sapx_bigworld_step: free motion failed");
    int nc = *b->hcount;
    if (nc < 0) nc = 0;
    if (nc > b->C) nc = b->C;  // contacts beyond the capacity are dropped (contact_jacobian.py:470-472)
    if (nc > c->capacity) return bfail(SAPX_ERR_INVALID, "sapx_bigworld_step: contact count above the view's capacity");
    if (nc > 0) {
        RowArgs ra{b->bm,       nc,          c->shape0,  c->shape1, c->point0, c->point1, c->normal,
                   c->margin0,  c->margin1,  b->pose,    b->adiag,  o.fallback_mu, o.fallback_stiffness,
                   o.fallback_tau_d, o.contact_weight_mode, o.contact_point_mode, b->part, b->cols, b->vals,
                   b->phi0,     b->weight,   b->stiff,   b->tau,    b->mu};
        hipLaunchKernelGGL(bw_rows, dim3((nc + NT - 1) / NT), dim3(NT), 0, st, ra);
    }
    if (hipGetLastError() != hipSuccess) return bfail(SAPX_ERR_HIP, "sapx_bigworld_step: launch failed");
    SapxLargeProblem p{};
    p.n = b->n;
    p.contact_count = nc;
    p.kmax = KB;
    p.A = b->A;
    p.v_star = b->vstar;
    // warm start: the previous step's solution once there is one (solver_sap.py:1004-1017,
    // 1388-1494); l_init reads it before the solve writes the new v into the same buffer
    p.v_guess = b->has_guess ? b->v : b->v0;
    p.participating = b->part;
    p.J_cols = b->cols;
    p.J_vals = b->vals;
    p.phi0 = b->phi0;
    p.weight = b->weight;
    p.stiffness = b->stiff;
    p.tau = b->tau;
    p.mu = b->mu;
    p.dt = dt;
    SapxLargeOpts lo{};
    lo.line_search = o.line_search;
    lo.max_iterations = o.max_iterations;
    lo.ls_max_iterations = o.line_search_max_iterations;
    lo.refine = 1;
    lo.optimality_abs_tol = o.optimality_abs_tol;
    lo.optimality_rel_tol = o.optimality_rel_tol;
    lo.cost_abs_tol = o.cost_abs_tol;
    lo.cost_rel_tol = o.cost_rel_tol;
    lo.armijo_c = o.armijo_c;
    lo.rho = o.rho;
    lo.relative_slop = o.line_search_relative_slop;
    lo.beta = o.contact_beta;
    lo.sigma = o.contact_sigma;
    lo.diag_shift = o.diag_shift;
    const int32_t rc = sapx_large_solve(b->large, &p, &lo, b->v, gamma_out, stats, stream);
    if (rc != SAPX_OK) {
        b->has_guess = 0;
        return rc;
    }
    b->has_guess = 1;
    IntArgs ia{b->bm, in->joint_q, b->v0, b->v, (double)dt, o.position_integration, out->joint_q, out->joint_qd,
               out->body_q, out->body_qd};
    hipLaunchKernelGGL(bw_integrate, dim3((b->J + NT - 1) / NT), dim3(NT), 0, st, ia);
    if (v_out && hipMemcpyAsync(v_out, b->v, (size_t)b->n * 8, hipMemcpyDeviceToDevice, st) != hipSuccess)
        return bfail(SAPX_ERR_HIP, "sapx_bigworld_step: copy of v failed");
    if (hipGetLastError() != hipSuccess) return bfail(SAPX_ERR_HIP, "sapx_bigworld_step: launch failed");
    return SAPX_OK;
}

}  // extern "C"

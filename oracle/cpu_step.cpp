/*
 * CPU restatement of the SAP step in C++17 — TEST INFRASTRUCTURE AND CPU BASELINE ONLY.
 *
 * The same algorithm as the float64 numpy oracle (oracle/sap_oracle.py), stage for stage,
 * in double precision with one OpenMP thread per world at a time.  Reference lines:
 *   boundary conversions     sim/sap_helpers.py:1482-1522, 1811-1848, 2058-2072
 *   free motion              sim/free_motion.py:412-530 (kinematics), 680-719 (body dynamics),
 *                            93-147 (torques), 722-1179 (Jacobian, H = J^T M J), 2094-2178 (v*)
 *   contact Jacobian         sim/contact_jacobian.py:404-562 (scatter, frames, phi0, J rows),
 *                            153-251 (weights); sim/sap_helpers.py:148-237 (materials),
 *                            307-338 (body direction weight), 467-503 (contact frame)
 *   projection               sim/contact_solve.py:743-921
 *   PD / limits              sim/contact_solve.py:490-683; sim/sap_helpers.py:2251-2340
 *   Newton, line searches    sim/contact_solve.py:1493-1654, 1970-2087, 2234-3176, 6129-6285
 *   integration              sim/solver_sap.py:288-519
 *   public FK                sim/sap_kinematics.py:84-229
 * Collision is not here: contacts come in world-grouped (as oracle/narrowphase.c writes them).
 *
 * Only tests/ and the CPU-baseline leg of bench.py load this library.
 */
#include <omp.h>

#include <algorithm>
#include <array>
#include <cmath>
#include <cstdint>
#include <cstring>
#include <vector>

namespace {

constexpr int kPrismatic = 0, kRevolute = 1, kBall = 2, kFixed = 3, kFree = 4, kDistance = 5, kD6 = 6;
constexpr double kPi = 3.14159265358979323846;
constexpr double kSoftNormTol = 1.0e-7;     // contact_solve.py SOFT_NORM_TOL
constexpr double kPdBeta = 0.1;             // sap_helpers.py near-rigid PD beta
constexpr double kLimitBeta = 0.1;
constexpr double kLimitWindow = 2.0;
constexpr double kExactAlphaMax = 1.5;
constexpr double kExactFTol = 1.0e-8;
constexpr double kMonoDecay = 0.5, kMonoMinAlpha = 1.0e-8, kMonoRelaxR = 1.0e-12, kMonoRelaxA = 1.0e-14;

struct V3 {
    double x = 0, y = 0, z = 0;
};
inline V3 v3(double x, double y, double z) { return V3{x, y, z}; }
inline V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
inline V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
inline V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
inline V3 operator*(V3 a, double s) { return v3(a.x * s, a.y * s, a.z * s); }
inline V3 operator*(double s, V3 a) { return a * s; }
inline double dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
inline V3 cross(V3 a, V3 b) { return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x); }
inline double at(V3 a, int i) { return i == 0 ? a.x : (i == 1 ? a.y : a.z); }
inline V3 ld3(const float* p) { return v3(p[0], p[1], p[2]); }

struct Q {
    double x = 0, y = 0, z = 0, w = 1;
};
inline Q qmul(Q a, Q b) {
    return Q{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
             a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
inline Q qnormalize(Q q) {
    const double n = std::sqrt(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n <= 0.0) return Q{};
    return Q{q.x / n, q.y / n, q.z / n, q.w / n};
}
inline Q qconj(Q q) { return Q{-q.x, -q.y, -q.z, q.w}; }
inline V3 qrot_unit(Q q, V3 v) {
    const V3 qv = v3(q.x, q.y, q.z);
    const V3 t = 2.0 * cross(qv, v);
    return v + q.w * t + cross(qv, t);
}
inline V3 qrot(Q q, V3 v) { return qrot_unit(qnormalize(q), v); }
inline Q q_axis_angle(V3 a, double ang) {
    const double s = std::sin(0.5 * ang), c = std::cos(0.5 * ang);
    return Q{a.x * s, a.y * s, a.z * s, c};
}

struct TF {
    V3 p;
    Q q;
};
inline TF ld_tf(const float* a) { return TF{v3(a[0], a[1], a[2]), Q{a[3], a[4], a[5], a[6]}}; }
inline TF tcompose(TF a, TF b) {
    const Q aq = qnormalize(a.q), bq = qnormalize(b.q);
    return TF{a.p + qrot_unit(aq, b.p), qnormalize(qmul(aq, bq))};
}
inline TF tinverse(TF x) {
    const Q qi = qconj(qnormalize(x.q));
    return TF{-qrot_unit(qi, x.p), qi};
}
// Quaternion of the matrix with columns c0, c1, c2, normalised (``_quatd_from_matrix_cols``,
// sap_helpers.py:607-652).
inline Q q_from_cols(V3 c0, V3 c1, V3 c2) {
    const double tr = c0.x + c1.y + c2.z;
    Q q;
    if (tr > 0.0) {
        const double s = std::sqrt(tr + 1.0) * 2.0;
        q = Q{(c1.z - c2.y) / s, (c2.x - c0.z) / s, (c0.y - c1.x) / s, 0.25 * s};
    } else if (c0.x > c1.y && c0.x > c2.z) {
        const double s = std::sqrt(1.0 + c0.x - c1.y - c2.z) * 2.0;
        q = Q{0.25 * s, (c1.x + c0.y) / s, (c2.x + c0.z) / s, (c1.z - c2.y) / s};
    } else if (c1.y > c2.z) {
        const double s = std::sqrt(1.0 + c1.y - c0.x - c2.z) * 2.0;
        q = Q{(c1.x + c0.y) / s, 0.25 * s, (c2.y + c1.z) / s, (c2.x - c0.z) / s};
    } else {
        const double s = std::sqrt(1.0 + c2.z - c0.x - c1.y) * 2.0;
        q = Q{(c2.x + c0.z) / s, (c2.y + c1.z) / s, 0.25 * s, (c0.y - c1.x) / s};
    }
    return qnormalize(q);
}
// Rotation of a D6 joint's na angular coordinates q and each angular dof's current axis (joint
// frame): one axis turns by its angle; two or three are first replaced by the frame their
// matrix defines (third column a0 x a1 for two), then turned in sequence, each later axis
// carried by the rotations before it (sap_helpers.py:1109-1161, 1247-1286).
template <class T>
inline Q d6_rotation(int na, const V3* a, const T* q, V3* axis) {
    if (na <= 0) return Q{};
    if (na == 1) {
        axis[0] = a[0];
        return q_axis_angle(a[0], (double)q[0]);
    }
    const Q off = q_from_cols(a[0], a[1], na == 2 ? cross(a[0], a[1]) : a[2]);
    const V3 l0 = qrot(off, v3(1, 0, 0)), l1 = qrot(off, v3(0, 1, 0)), l2 = qrot(off, v3(0, 0, 1));
    const Q q0 = q_axis_angle(l0, (double)q[0]);
    axis[0] = l0;
    axis[1] = qrot(q0, l1);
    const Q q10 = qmul(q_axis_angle(axis[1], (double)q[1]), q0);
    if (na == 2) return q10;
    axis[2] = qrot(q10, l2);
    return qmul(q_axis_angle(axis[2], (double)q[2]), q10);
}
inline V3 tpoint(TF x, V3 p) { return x.p + qrot(x.q, p); }
inline bool is_identity(const float* x) {
    return x[0] == 0 && x[1] == 0 && x[2] == 0 && x[3] == 0 && x[4] == 0 && x[5] == 0 && x[6] == 1;
}

// spatial vectors, angular first
struct SV {
    V3 w, v;
};
inline SV operator+(SV a, SV b) { return SV{a.w + b.w, a.v + b.v}; }
inline SV operator-(SV a, SV b) { return SV{a.w - b.w, a.v - b.v}; }
inline SV operator*(SV a, double s) { return SV{a.w * s, a.v * s}; }
inline double sdot(SV a, SV b) { return dot(a.w, b.w) + dot(a.v, b.v); }
inline SV shift_velocity(SV V, V3 p) { return SV{V.w, V.v + cross(V.w, p)}; }
inline SV shift_acceleration(SV A, V3 p, V3 w) { return SV{A.w, A.v + cross(A.w, p) + cross(w, cross(w, p))}; }
inline SV shift_force(SV F, V3 p) { return SV{F.w - cross(p, F.v), F.v}; }
inline SV compose_velocity(SV Vp, V3 p, SV Vb) { return SV{Vp.w + Vb.w, Vp.v + cross(Vp.w, p) + Vb.v}; }
inline SV compose_acceleration(SV Ap, SV Vp, V3 p, SV Vb, SV Ab) {
    const V3 w = Vp.w;
    const SV As = shift_acceleration(Ap, p, w);
    return SV{As.w + Ab.w + cross(w, Vb.w), As.v + Ab.v + 2.0 * cross(w, Vb.v)};
}
inline SV mob_to_body(TF X_WF, V3 p, SV V) {
    return SV{qrot_unit(X_WF.q, V.w), qrot_unit(X_WF.q, V.v + cross(V.w, p))};
}
inline SV mob_bias_to_body(TF X_WF, V3 p, SV V, SV A) {
    const SV s = shift_acceleration(A, p, V.w);
    return SV{qrot_unit(X_WF.q, s.w), qrot_unit(X_WF.q, s.v)};
}

using M3 = std::array<double, 9>;  // row-major
inline V3 m3v(const M3& m, V3 v) {
    return v3(m[0] * v.x + m[1] * v.y + m[2] * v.z, m[3] * v.x + m[4] * v.y + m[5] * v.z,
              m[6] * v.x + m[7] * v.y + m[8] * v.z);
}

// spatial inertia about the body origin, world frame: [I_O, m[c]x; -m[c]x, m 1]
struct SI {
    M3 Io;
    V3 c;
    double m;
};
SI spatial_inertia(TF X, const float* I_com, double m, V3 com) {
    M3 R;
    const V3 c0 = qrot_unit(X.q, v3(1, 0, 0)), c1 = qrot_unit(X.q, v3(0, 1, 0)), c2 = qrot_unit(X.q, v3(0, 0, 1));
    R = {c0.x, c1.x, c2.x, c0.y, c1.y, c2.y, c0.z, c1.z, c2.z};
    M3 IW{};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) {
            double s = 0;
            for (int k = 0; k < 3; ++k)
                for (int l = 0; l < 3; ++l) s += R[i * 3 + k] * (double)I_com[k * 3 + l] * R[j * 3 + l];
            IW[i * 3 + j] = s;
        }
    const V3 c = qrot_unit(X.q, com);
    const double cc = dot(c, c);
    const double cv[3] = {c.x, c.y, c.z};
    for (int i = 0; i < 3; ++i)
        for (int j = 0; j < 3; ++j) IW[i * 3 + j] += m * ((i == j ? cc : 0.0) - cv[i] * cv[j]);
    return SI{IW, c, m};
}
inline SV si_mul(const SI& I, SV a) {
    // [I_O w + m c x v ; -m c x w + m v]
    return SV{m3v(I.Io, a.w) + I.m * cross(I.c, a.v), -(I.m * cross(I.c, a.w)) + I.m * a.v};
}

// ------------------------------------------------------------------ model and options
struct Model {
    int world_count, joint_count, dof_count, coord_count, body_count, articulation_count;
    const int *joint_type, *joint_parent, *joint_child, *joint_q_start, *joint_qd_start, *joint_dof_dim,
        *joint_ancestor, *articulation_start;
    const float *joint_axis, *joint_X_p, *joint_X_c, *joint_armature;
    const float *body_mass, *body_inertia, *body_com, *gravity;
    const int* joint_target_mode;
    const float *joint_target_ke, *joint_target_kd, *joint_effort_limit, *joint_limit_lower, *joint_limit_upper,
        *joint_limit_ke, *joint_limit_kd;
    const int* shape_body;
    const float *shape_mu, *shape_ke, *shape_tau;
};

struct Opts {
    int max_rigid_contact;
    double fallback_mu, fallback_k, fallback_tau, beta, sigma, diag_shift;
    int max_iterations;
    double opt_abs, opt_rel, cost_abs, cost_rel;
    int line_search;  // 0 armijo_decay, 1 exact_root, 2 monotone_decay
    int ls_max_iterations;
    double armijo_c, rho, relative_slop;
    int weight_mode;  // 0 body_inertia, 1 diag_delassus
    int point_mode;   // 0 witness, 1 midpoint
    int integration;  // 0 midpoint, 1 sap_euler
};

struct State {
    const float *joint_q, *joint_qd, *body_q, *body_f;  // body_f may be null
};
struct Control {
    const float *joint_f, *joint_target_pos, *joint_target_vel, *joint_act;
};
struct Contacts {
    const int *world_start, *shape0, *shape1;
    const float *point0, *point1, *normal, *margin0, *margin1;
};
struct Out {
    float *joint_q, *joint_qd, *body_q, *body_qd;
    double* v_sap;
    int *iterations, *ls_iterations, *contact_count, *converged;
};

// dense n x n row-major helpers
struct Mat {
    int n = 0;
    std::vector<double> a;
    void resize(int k) { n = k; a.assign((size_t)k * k, 0.0); }
    double& operator()(int i, int j) { return a[(size_t)i * n + j]; }
    double operator()(int i, int j) const { return a[(size_t)i * n + j]; }
};

// Cholesky solve of an SPD matrix (any correct factorisation is the reference's blocked one)
bool chol_solve(Mat& A, std::vector<double>& b) {
    const int n = A.n;
    for (int j = 0; j < n; ++j) {
        double d = A(j, j);
        for (int k = 0; k < j; ++k) d -= A(j, k) * A(j, k);
        if (!(d > 0.0)) return false;
        d = std::sqrt(d);
        A(j, j) = d;
        for (int i = j + 1; i < n; ++i) {
            double s = A(i, j);
            for (int k = 0; k < j; ++k) s -= A(i, k) * A(j, k);
            A(i, j) = s / d;
        }
    }
    for (int i = 0; i < n; ++i) {
        double s = b[i];
        for (int k = 0; k < i; ++k) s -= A(i, k) * b[k];
        b[i] = s / A(i, i);
    }
    for (int i = n - 1; i >= 0; --i) {
        double s = b[i];
        for (int k = i + 1; k < n; ++k) s -= A(k, i) * b[k];
        b[i] = s / A(i, i);
    }
    return true;
}

// ------------------------------------------------------------------ per-world problem
struct Contact {
    int b0, b1;  // local bodies or -1
    V3 x0, x1, pc;
    V3 cx, cy, cz;
    double phi0, mu, k, tau, w;
    double rt, rn, vhat;
    std::vector<double> J;  // 3 x n rows t1, t2, n
};

struct Terms {
    std::vector<int> pd_active, lo_active, up_active;
    std::vector<double> pd_a, pd_gain, pd_limit, lo_vhat, up_vhat, lo_r, up_r;
};

struct Eval {
    double cost;
    std::vector<double> grad, jimp, Av, hdiag;
    std::vector<std::array<double, 3>> gamma;
    std::vector<std::array<double, 9>> G;
};

inline double clamp_scalar(double x, double lim) {
    if (lim <= 0.0 || !std::isfinite(lim)) return x;
    return std::max(-lim, std::min(lim, x));
}
inline double clamp_derivative(double x, double lim) {
    if (lim <= 0.0 || !std::isfinite(lim)) return 1.0;
    return (-lim <= x && x <= lim) ? 1.0 : 0.0;
}
inline double clamp_antiderivative(double x, double lim) {
    if (lim <= 0.0 || !std::isfinite(lim)) return 0.5 * x * x;
    if (x < -lim) return -lim * (x + 0.5 * lim);
    if (x <= lim) return 0.5 * x * x;
    return lim * (x - 0.5 * lim);
}

void effective_pd_gains(double kp, double kd, double dt, double delassus, double* kp_e, double* kd_e) {
    *kp_e = 0.0;
    *kd_e = 0.0;
    if (kp <= 0.0 && kd <= 0.0) return;
    const double bf = kPdBeta * kPdBeta / (4.0 * kPi * kPi);
    const double r_nr = bf * std::max(delassus, 1.0e-12);
    const double common = dt * kp + kd;
    if (common <= 0.0) return;
    const double r = 1.0 / (dt * common);
    if (r >= r_nr) {
        *kp_e = kp;
        *kd_e = kd;
        return;
    }
    if (dt * kp > kd && kp > 0.0) {
        const double tau = kd / kp;
        *kp_e = 1.0 / (dt * (dt + tau) * r_nr);
        *kd_e = tau * *kp_e;
    } else if (kd > 0.0) {
        const double tau_inv = kp / kd;
        *kd_e = 1.0 / (dt * (1.0 + dt * tau_inv) * r_nr);
        *kp_e = tau_inv * *kd_e;
    } else {
        *kp_e = 1.0 / (dt * dt * r_nr);
    }
}

// projection onto the regularised friction cone: gamma, G = dgamma/dy R^-1, cost
double project(const double vc[3], double rt, double rn, double vhat, double mu, double g[3], double G[9]) {
    const double rti = 1.0 / rt, rni = 1.0 / rn;
    const double y[3] = {-rti * vc[0], -rti * vc[1], rni * (vhat - vc[2])};
    const double yr = std::sqrt(y[0] * y[0] + y[1] * y[1] + kSoftNormTol * kSoftNormTol);
    const double t0 = yr > 0.0 ? y[0] / yr : 0.0, t1 = yr > 0.0 ? y[1] / yr : 0.0;
    g[0] = g[1] = g[2] = 0.0;
    for (int i = 0; i < 9; ++i) G[i] = 0.0;
    if (mu <= 1.0e-12) {
        if (y[2] > 0.0) {
            g[2] = y[2];
            G[8] = rni;
        }
    } else {
        const double mu_tilde = mu * std::sqrt(rt / rn), mu_hat = mu * rt / rn;
        const double f = 1.0 / (1.0 + mu_tilde * mu_tilde);
        if (yr <= mu * y[2]) {
            g[0] = y[0];
            g[1] = y[1];
            g[2] = y[2];
            G[0] = rti;
            G[4] = rti;
            G[8] = rni;
        } else if (-mu_hat * yr < y[2] && y[2] < yr / mu) {
            const double gn = (y[2] + mu_hat * yr) * f;
            g[0] = mu * gn * t0;
            g[1] = mu * gn * t1;
            g[2] = gn;
            const double gy = yr > 0.0 ? gn / yr : 0.0;
            const double P00 = t0 * t0, P01 = t0 * t1, P11 = t1 * t1;
            G[0] = mu * (gy * (1.0 - P00) + mu_hat * f * P00) * rti;
            G[1] = mu * (gy * (-P01) + mu_hat * f * P01) * rti;
            G[3] = mu * (gy * (-P01) + mu_hat * f * P01) * rti;
            G[4] = mu * (gy * (1.0 - P11) + mu_hat * f * P11) * rti;
            G[2] = mu * f * t0 * rni;
            G[5] = mu * f * t1 * rni;
            G[6] = mu_hat * f * t0 * rti;
            G[7] = mu_hat * f * t1 * rti;
            G[8] = f * rni;
        }
    }
    return 0.5 * (rt * (g[0] * g[0] + g[1] * g[1]) + rn * g[2] * g[2]);
}

struct World {
    const Model& M;
    const Opts& O;
    int w, j0, j1, d0, d1, c0, c1, b0, b1, n, nb;
    double dt;
    // free motion
    std::vector<TF> bq;
    std::vector<SV> S, V, Aacc;
    std::vector<SI> I;
    std::vector<double> body_jac;  // nb x 6 x n (angular rows 0-2, linear rows 3-5)
    Mat A;
    std::vector<double> v0, vstar, tau;
    // contacts and terms
    std::vector<Contact> cs;
    Terms mt;
    std::vector<int> part;
    std::vector<double> Dsc;

    World(const Model& m, const Opts& o, int w_, double dt_) : M(m), O(o), w(w_), dt(dt_) {
        const int jpw = M.joint_count / M.world_count, bpw = M.body_count / M.world_count;
        j0 = w * jpw;
        j1 = j0 + jpw;
        d0 = M.joint_qd_start[j0];
        d1 = M.joint_qd_start[j1];
        c0 = M.joint_q_start[j0];
        c1 = M.joint_q_start[j1];
        b0 = w * bpw;
        b1 = b0 + bpw;
        n = d1 - d0;
        nb = b1 - b0;
    }
    int lb(int b) const { return b < 0 ? -1 : b - b0; }
    double& jac(int b, int r, int c) { return body_jac[((size_t)b * 6 + r) * n + c]; }

    // public -> SAP velocity and forces (free joints: COM velocity / wrench to the body origin)
    void boundary(const State& st, const Control& ctl, std::vector<double>& fj, std::vector<SV>& fb) {
        v0.assign(n, 0.0);
        fj.assign(n, 0.0);
        for (int i = 0; i < n; ++i) {
            v0[i] = st.joint_qd[d0 + i];
            fj[i] = ctl.joint_f ? ctl.joint_f[d0 + i] : 0.0;
        }
        for (int j = j0; j < j1; ++j) {
            const int jt = M.joint_type[j];
            if (jt != kFree && jt != kDistance) continue;
            const int child = M.joint_child[j];
            const int ds = M.joint_qd_start[j] - d0;
            const TF x = ld_tf(st.body_q + (size_t)child * 7);
            const V3 r = qrot(x.q, ld3(M.body_com + (size_t)child * 3));
            const V3 vcom = ld3(st.joint_qd + d0 + ds), wv = ld3(st.joint_qd + d0 + ds + 3);
            const V3 vo = vcom - cross(wv, r);
            v0[ds] = wv.x; v0[ds + 1] = wv.y; v0[ds + 2] = wv.z;
            v0[ds + 3] = vo.x; v0[ds + 4] = vo.y; v0[ds + 5] = vo.z;
            if (ctl.joint_f) {
                const V3 f = ld3(ctl.joint_f + d0 + ds), t = ld3(ctl.joint_f + d0 + ds + 3);
                const V3 to = t + cross(r, f);
                fj[ds] = to.x; fj[ds + 1] = to.y; fj[ds + 2] = to.z;
                fj[ds + 3] = f.x; fj[ds + 4] = f.y; fj[ds + 5] = f.z;
            }
        }
        fb.assign(nb, SV{});
        if (st.body_f) {
            for (int b = 0; b < nb; ++b) {
                const int g = b0 + b;
                const V3 f = ld3(st.body_f + (size_t)g * 6), t = ld3(st.body_f + (size_t)g * 6 + 3);
                const TF x = ld_tf(st.body_q + (size_t)g * 7);
                const V3 c = qrot(x.q, ld3(M.body_com + (size_t)g * 3));
                fb[b] = SV{-(t + cross(c, f)), -f};
            }
        }
    }

    void free_motion(const float* joint_q, const std::vector<double>& fj, const std::vector<SV>& fb) {
        bq.assign(nb, TF{});
        V.assign(nb, SV{});
        Aacc.assign(nb, SV{});
        S.assign(n, SV{});
        const V3 g = ld3(M.gravity + (size_t)w * 3);
        for (int j = j0; j < j1; ++j) {
            const int jt = M.joint_type[j];
            const int parent = lb(M.joint_parent[j]), child = lb(M.joint_child[j]);
            const int qs = M.joint_q_start[j], ds = M.joint_qd_start[j] - d0;
            const int nd = M.joint_qd_start[j + 1] - M.joint_qd_start[j];
            const TF Xp = ld_tf(M.joint_X_p + (size_t)j * 7), Xc = ld_tf(M.joint_X_c + (size_t)j * 7);
            const bool xc_id = is_identity(M.joint_X_c + (size_t)j * 7);
            const TF Xwpj = parent < 0 ? Xp : tcompose(bq[parent], Xp);
            const double* qd = v0.data() + ds;
            V3 ax0 = nd > 0 ? ld3(M.joint_axis + (size_t)(d0 + ds) * 3) : V3{};
            V3 d6ax[6], d6cur[3];
            int nlin = 0;
            TF Xwc;
            SV vPB, aPB;
            if (jt == kRevolute && xc_id) {
                const Q qFM = q_axis_angle(ax0, joint_q[qs]);
                Xwc = TF{Xwpj.p, qmul(Xwpj.q, qFM)};
                S[ds] = SV{qrot_unit(Xwpj.q, ax0), V3{}};
                vPB = S[ds] * qd[0];
                aPB = SV{};
            } else {
                TF Xj;
                if (jt == kPrismatic) Xj = TF{ax0 * joint_q[qs], Q{}};
                else if (jt == kRevolute) Xj = TF{V3{}, q_axis_angle(ax0, joint_q[qs])};
                else if (jt == kBall) Xj = TF{V3{}, Q{joint_q[qs], joint_q[qs + 1], joint_q[qs + 2], joint_q[qs + 3]}};
                else if (jt == kFree || jt == kDistance)
                    Xj = TF{v3(joint_q[qs], joint_q[qs + 1], joint_q[qs + 2]),
                            Q{joint_q[qs + 3], joint_q[qs + 4], joint_q[qs + 5], joint_q[qs + 6]}};
                else if (jt == kD6) {
                    nlin = M.joint_dof_dim[2 * j];
                    for (int k = 0; k < nd; ++k) d6ax[k] = ld3(M.joint_axis + (size_t)(d0 + ds + k) * 3);
                    V3 pos{};
                    for (int k = 0; k < nlin; ++k) pos = pos + d6ax[k] * joint_q[qs + k];
                    Xj = TF{pos, d6_rotation(nd - nlin, d6ax + nlin, joint_q + qs + nlin, d6cur)};
                } else Xj = TF{};
                const TF XWM = tcompose(Xwpj, Xj);
                Xwc = xc_id ? XWM : tcompose(XWM, tinverse(Xc));
                const TF XMB = tinverse(Xc);
                const V3 p = qrot_unit(Xj.q, XMB.p);
                if (jt == kPrismatic) {
                    S[ds] = mob_to_body(Xwpj, p, SV{V3{}, ax0});
                    const SV VFM{V3{}, ax0 * qd[0]};
                    aPB = mob_bias_to_body(Xwpj, p, VFM, SV{});
                    vPB = S[ds] * qd[0];
                } else if (jt == kRevolute) {
                    S[ds] = SV{qrot_unit(Xwpj.q, ax0), qrot_unit(Xwpj.q, cross(ax0, p))};
                    const V3 wFM = ax0 * qd[0];
                    aPB = SV{V3{}, qrot_unit(Xwpj.q, cross(wFM, cross(wFM, p)))};
                    vPB = S[ds] * qd[0];
                } else if (jt == kBall) {
                    const SV VFM{v3(qd[0], qd[1], qd[2]), V3{}};
                    vPB = mob_to_body(Xwpj, p, VFM);
                    for (int k = 0; k < 3; ++k) {
                        V3 e{};
                        (k == 0 ? e.x : (k == 1 ? e.y : e.z)) = 1.0;
                        S[ds + k] = mob_to_body(Xwpj, p, SV{e, V3{}});
                    }
                    aPB = mob_bias_to_body(Xwpj, p, VFM, SV{});
                } else if (jt == kFree || jt == kDistance) {
                    const SV VFM{v3(qd[0], qd[1], qd[2]), v3(qd[3], qd[4], qd[5])};
                    vPB = mob_to_body(Xwpj, p, VFM);
                    for (int k = 0; k < 3; ++k) {
                        V3 e{};
                        (k == 0 ? e.x : (k == 1 ? e.y : e.z)) = 1.0;
                        S[ds + k] = mob_to_body(Xwpj, p, SV{e, V3{}});
                        S[ds + 3 + k] = mob_to_body(Xwpj, p, SV{V3{}, e});
                    }
                    aPB = mob_bias_to_body(Xwpj, p, VFM, SV{});
                } else if (jt == kD6) {
                    // linear axes fixed in the joint frame, angular axes at the current angles
                    // (sap_helpers.py:1387-1432)
                    SV VFM{};
                    for (int k = 0; k < nd; ++k) {
                        const SV col = k < nlin ? SV{V3{}, d6ax[k]} : SV{d6cur[k - nlin], V3{}};
                        S[ds + k] = mob_to_body(Xwpj, p, col);
                        VFM = VFM + col * qd[k];
                    }
                    vPB = mob_to_body(Xwpj, p, VFM);
                    aPB = mob_bias_to_body(Xwpj, p, VFM, SV{});
                } else {
                    vPB = SV{};
                    aPB = SV{};
                }
            }
            bq[child] = Xwc;
            if (parent >= 0) {
                const V3 pPB = Xwc.p - bq[parent].p;
                V[child] = compose_velocity(V[parent], pPB, vPB);
                Aacc[child] = compose_acceleration(Aacc[parent], V[parent], pPB, vPB, aPB);
            } else {
                V[child] = vPB;
                Aacc[child] = aPB;
            }
        }
        // body dynamics
        I.resize(nb);
        std::vector<SV> fbody(nb);
        for (int b = 0; b < nb; ++b) {
            const int gb = b0 + b;
            const double m = M.body_mass[gb];
            const V3 com = ld3(M.body_com + (size_t)gb * 3);
            I[b] = spatial_inertia(bq[b], M.body_inertia + (size_t)gb * 9, m, com);
            const SV IA = si_mul(I[b], Aacc[b]);
            const V3 wv = V[b].w;
            const SV bias{cross(wv, m3v(I[b].Io, wv)), m * cross(wv, cross(wv, I[b].c))};
            const V3 f = m * g;
            const SV grav{cross(I[b].c, f), f};
            fbody[b] = IA + bias - grav;
        }
        // torques, reverse recursion
        tau.assign(n, 0.0);
        std::vector<SV> ft(nb);
        for (int j = j1 - 1; j >= j0; --j) {
            const int parent = lb(M.joint_parent[j]), child = lb(M.joint_child[j]);
            const int ds = M.joint_qd_start[j] - d0;
            const int nd = M.joint_qd_start[j + 1] - M.joint_qd_start[j];
            const SV fs = fbody[child] + ft[child] + fb[child];
            for (int k = 0; k < nd; ++k) tau[ds + k] = -sdot(S[ds + k], fs) + fj[ds + k];
            if (parent >= 0) ft[parent] = ft[parent] + shift_force(fs, bq[parent].p - bq[child].p);
        }
        // body-origin Jacobians, H = sum_b J_b^T I_b J_b
        std::vector<int> joint_of_body(nb, -1);
        for (int j = j0; j < j1; ++j) joint_of_body[lb(M.joint_child[j])] = j;
        body_jac.assign((size_t)nb * 6 * n, 0.0);
        for (int b = 0; b < nb; ++b) {
            int j = joint_of_body[b];
            if (j < 0) continue;
            const V3 prow = bq[b].p;
            while (j != -1) {
                const int cb = lb(M.joint_child[j]);
                const int ds = M.joint_qd_start[j] - d0;
                const int nd = M.joint_qd_start[j + 1] - M.joint_qd_start[j];
                for (int k = 0; k < nd; ++k) {
                    const SV s = shift_velocity(S[ds + k], prow - bq[cb].p);
                    jac(b, 0, ds + k) = s.w.x; jac(b, 1, ds + k) = s.w.y; jac(b, 2, ds + k) = s.w.z;
                    jac(b, 3, ds + k) = s.v.x; jac(b, 4, ds + k) = s.v.y; jac(b, 5, ds + k) = s.v.z;
                }
                j = M.joint_ancestor[j];
            }
        }
        A.resize(n);
        std::vector<SV> IJ(n);
        for (int b = 0; b < nb; ++b) {
            for (int c = 0; c < n; ++c) {
                const SV col{v3(jac(b, 0, c), jac(b, 1, c), jac(b, 2, c)), v3(jac(b, 3, c), jac(b, 4, c), jac(b, 5, c))};
                IJ[c] = si_mul(I[b], col);
            }
            for (int r = 0; r < n; ++r) {
                const SV row{v3(jac(b, 0, r), jac(b, 1, r), jac(b, 2, r)), v3(jac(b, 3, r), jac(b, 4, r), jac(b, 5, r))};
                if (row.w.x == 0 && row.w.y == 0 && row.w.z == 0 && row.v.x == 0 && row.v.y == 0 && row.v.z == 0)
                    continue;
                for (int c = 0; c < n; ++c) A(r, c) += sdot(row, IJ[c]);
            }
        }
        for (int r = 0; r < n; ++r)
            for (int c = r + 1; c < n; ++c) {
                const double s = 0.5 * (A(r, c) + A(c, r));
                A(r, c) = s;
                A(c, r) = s;
            }
        for (int i = 0; i < n; ++i) A(i, i) += M.joint_armature[d0 + i];
        Mat L = A;
        std::vector<double> qdd = tau;
        chol_solve(L, qdd);
        vstar.resize(n);
        for (int i = 0; i < n; ++i) vstar[i] = v0[i] + dt * qdd[i];
    }

    static V3 safe_normalize(V3 v) {
        const double n2 = dot(v, v);
        if (n2 <= 1.0e-32) return V3{};
        return v * (1.0 / std::sqrt(n2));
    }
    static void make_frame(V3 nv, V3* cx, V3* cy) {
        const double ax = std::fabs(nv.x), ay = std::fabs(nv.y), az = std::fabs(nv.z);
        int i = 0;
        double mn = ax;
        if (ay < mn) { i = 1; mn = ay; }
        if (az < mn) { i = 2; mn = az; }
        const double mag = std::sqrt(1.0 - mn * mn), r = 1.0 / mag;
        const double s = -r * at(nv, i);
        if (i == 0) { *cx = v3(0.0, -r * nv.z, r * nv.y); *cy = v3(mag, s * nv.y, s * nv.z); }
        else if (i == 1) { *cx = v3(r * nv.z, 0.0, -r * nv.x); *cy = v3(s * nv.x, mag, s * nv.z); }
        else { *cx = v3(-r * nv.y, r * nv.x, 0.0); *cy = v3(s * nv.x, s * nv.y, mag); }
    }
    double shape_k(int s) const {
        double v = O.fallback_k;
        if (s >= 0) v = M.shape_ke[s];
        if (!std::isfinite(v) || v <= 0.0) v = O.fallback_k;
        if (!std::isfinite(v) || v <= 0.0) v = 1.0;
        return v;
    }
    double shape_mu(int s) const { return std::max(s >= 0 ? (double)M.shape_mu[s] : O.fallback_mu, 0.0); }
    double shape_tau(int s) const {
        double v = O.fallback_tau;
        if (s >= 0) v = M.shape_tau[s];
        if (!std::isfinite(v) || v < 0.0) v = O.fallback_tau;
        if (!std::isfinite(v) || v < 0.0) v = 0.0;
        return v;
    }
    double direction_weight(int gb, const TF& pose, V3 r, V3 d) const {
        const double m = M.body_mass[gb];
        if (m <= 1.0e-12) return 0.0;
        const V3 cw = cross(r, d);
        const Q q = pose.q;
        const double n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
        const Q qi{-q.x / n2, -q.y / n2, -q.z / n2, q.w / n2};
        const V3 cb = qrot(qi, cw);
        const float* Ib = M.body_inertia + (size_t)gb * 9;
        const double ix = std::max((double)Ib[0], 1e-12), iy = std::max((double)Ib[4], 1e-12),
                     iz = std::max((double)Ib[8], 1e-12);
        return 1.0 / m + cb.x * cb.x / ix + cb.y * cb.y / iy + cb.z * cb.z / iz;
    }

    void contacts(const Contacts& c) {
        cs.clear();
        const int begin = c.world_start[w], end = c.world_start[w + 1];
        for (int i = begin; i < end && (int)cs.size() < O.max_rigid_contact; ++i) {
            Contact k;
            const int s0 = c.shape0[i], s1 = c.shape1[i];
            const int gb0 = s0 >= 0 ? M.shape_body[s0] : -1, gb1 = s1 >= 0 ? M.shape_body[s1] : -1;
            k.b0 = lb(gb0);
            k.b1 = lb(gb1);
            V3 x0 = ld3(c.point0 + (size_t)i * 3), x1 = ld3(c.point1 + (size_t)i * 3), p0{}, p1{};
            if (k.b0 >= 0) { p0 = bq[k.b0].p; x0 = tpoint(bq[k.b0], x0); }
            if (k.b1 >= 0) { p1 = bq[k.b1].p; x1 = tpoint(bq[k.b1], x1); }
            k.cz = safe_normalize(ld3(c.normal + (size_t)i * 3));
            make_frame(k.cz, &k.cx, &k.cy);
            const double k0 = shape_k(s0), k1 = shape_k(s1), den = k0 + k1;
            const double w0 = den != 0.0 ? k0 / den : 0.5, w1 = den != 0.0 ? k1 / den : 0.5;
            k.pc = w0 * x0 + w1 * x1;
            k.x0 = x0;
            k.x1 = x1;
            k.phi0 = dot(x1 - x0, k.cz) - (double)c.margin0[i] - (double)c.margin1[i];
            const double m0 = shape_mu(s0), m1 = shape_mu(s1), dm = m0 + m1;
            k.mu = dm <= 1e-12 ? 0.0 : 2.0 * m0 * m1 / dm;
            if (!std::isfinite(k0)) k.k = k1;
            else if (!std::isfinite(k1)) k.k = k0;
            else k.k = den <= 1e-12 ? std::max(k0, k1) : k0 * k1 / den;
            k.tau = shape_tau(s0) + shape_tau(s1);
            const bool wit = O.point_mode == 0;
            const V3 xj0 = wit ? x0 : k.pc, xj1 = wit ? x1 : k.pc;
            k.J.assign((size_t)3 * n, 0.0);
            for (int col = 0; col < n; ++col) {
                V3 rel{};
                if (k.b0 >= 0) {
                    const V3 wj = v3(jac(k.b0, 0, col), jac(k.b0, 1, col), jac(k.b0, 2, col));
                    const V3 vj = v3(jac(k.b0, 3, col), jac(k.b0, 4, col), jac(k.b0, 5, col));
                    rel = rel - (vj + cross(wj, xj0 - p0));
                }
                if (k.b1 >= 0) {
                    const V3 wj = v3(jac(k.b1, 0, col), jac(k.b1, 1, col), jac(k.b1, 2, col));
                    const V3 vj = v3(jac(k.b1, 3, col), jac(k.b1, 4, col), jac(k.b1, 5, col));
                    rel = rel + (vj + cross(wj, xj1 - p1));
                }
                k.J[col] = dot(k.cx, rel);
                k.J[n + col] = dot(k.cy, rel);
                k.J[2 * n + col] = dot(k.cz, rel);
            }
            // weight
            if (O.weight_mode == 0) {
                const V3 ex = safe_normalize(k.cx), ey = safe_normalize(k.cy), ez = safe_normalize(k.cz);
                double acc = 0.0;
                const int bs[2] = {k.b0, k.b1};
                const V3 xs[2] = {x0, x1};
                for (int side = 0; side < 2; ++side) {
                    if (bs[side] < 0) continue;
                    const int gb = b0 + bs[side];
                    const TF& pose = bq[bs[side]];
                    const V3 comw = pose.p + qrot(pose.q, ld3(M.body_com + (size_t)gb * 3));
                    const V3 r = xs[side] - comw;
                    acc += direction_weight(gb, pose, r, ex) + direction_weight(gb, pose, r, ey) +
                           direction_weight(gb, pose, r, ez);
                }
                k.w = std::max(acc / 3.0, 1e-12);
            } else {
                double Wm[9] = {0};
                for (int col = 0; col < n; ++col) {
                    const double ai = 1.0 / std::max(A(col, col), 1e-12);
                    for (int r = 0; r < 3; ++r)
                        for (int s = 0; s < 3; ++s) Wm[r * 3 + s] += k.J[r * n + col] * ai * k.J[s * n + col];
                }
                double f = 0.0;
                for (double x : Wm) f += x * x;
                k.w = std::max(std::sqrt(f) / 3.0, 1e-12);
            }
            // regularisation
            double wi = k.w;
            if (wi < 1e-12 || !std::isfinite(wi)) wi = 1e-12;
            double kc = k.k;
            if (kc <= 0.0 || !std::isfinite(kc)) kc = 1.0;
            const double tc = std::max(k.tau, 0.0);
            const double rn = std::max(O.beta * O.beta / (4.0 * kPi * kPi) * wi, 1.0 / (dt * kc * (dt + tc)));
            k.rt = std::max(O.sigma * wi, 1e-30);
            k.rn = std::max(rn, 1e-30);
            k.vhat = -k.phi0 / (dt + tc);
            if (k.mu < 0.0 || !std::isfinite(k.mu)) k.mu = 0.0;
            cs.push_back(std::move(k));
        }
    }

    void model_terms(const float* joint_q_all, const Control& ctl, const int* coord, const int* target,
                     const int* limit_ok) {
        mt.pd_active.assign(n, 0); mt.lo_active.assign(n, 0); mt.up_active.assign(n, 0);
        mt.pd_a.assign(n, 0); mt.pd_gain.assign(n, 0); mt.pd_limit.assign(n, 0);
        mt.lo_vhat.assign(n, 0); mt.up_vhat.assign(n, 0); mt.lo_r.assign(n, 1.0); mt.up_r.assign(n, 1.0);
        for (int i = 0; i < n; ++i) {
            const int dof = d0 + i, c = coord[dof], t = target[dof];
            double aii = A(i, i);
            if (!(aii >= 1e-12 && std::isfinite(aii))) aii = 1e-12;
            const double a_inv = 1.0 / aii;
            if (M.joint_target_mode[t] != 0) {
                const double kp = M.joint_target_ke[t], kd = M.joint_target_kd[t];
                if (!(kp <= 0.0 && kd <= 0.0) && !(c < 0 && kp > 0.0)) {
                    double kpe, kde;
                    effective_pd_gains(kp, kd, dt, a_inv, &kpe, &kde);
                    const double gain = dt * kpe + kde;
                    if (gain > 0.0) {
                        const double q0 = c >= 0 ? (double)joint_q_all[c] : 0.0;
                        const double qd = (c >= 0 && ctl.joint_target_pos) ? (double)ctl.joint_target_pos[t] : 0.0;
                        const double vd = ctl.joint_target_vel ? (double)ctl.joint_target_vel[t] : 0.0;
                        const double u0 = ctl.joint_act ? (double)ctl.joint_act[t] : 0.0;
                        mt.pd_active[i] = 1;
                        mt.pd_a[i] = kpe * (qd - q0) + kde * vd + u0;
                        mt.pd_gain[i] = gain;
                        mt.pd_limit[i] = M.joint_effort_limit[t];
                    }
                }
            }
            if (c < 0 || limit_ok[dof] == 0) continue;
            const double ke = M.joint_limit_ke[dof];
            if (ke <= 0.0) continue;
            const double q0 = joint_q_all[c];
            const double lower = M.joint_limit_lower[dof], upper = M.joint_limit_upper[dof];
            const double window = kLimitWindow * dt * std::max(std::fabs(v0[i]), std::fabs(vstar[i]));
            const double kd = M.joint_limit_kd[dof];
            const double taud = kd > 0.0 ? kd / ke : 0.0;
            const double bf = kLimitBeta * kLimitBeta / (4.0 * kPi * kPi);
            const double r = std::max(bf * std::max(a_inv, 1e-12), 1.0 / (dt * ke * (dt + taud)));
            if (std::isfinite(lower) && q0 - lower <= window) {
                mt.lo_active[i] = 1;
                mt.lo_vhat[i] = -(q0 - lower) / (dt + taud);
                mt.lo_r[i] = r;
            }
            if (std::isfinite(upper) && upper - q0 <= window) {
                mt.up_active[i] = 1;
                mt.up_vhat[i] = -(upper - q0) / (dt + taud);
                mt.up_r[i] = r;
            }
        }
    }

    double eval_terms(const std::vector<double>& v, std::vector<double>& imp, std::vector<double>& hd) const {
        double cost = 0.0;
        for (int i = 0; i < n; ++i) {
            imp[i] = 0.0;
            hd[i] = 0.0;
            if (mt.pd_active[i]) {
                const double gain = mt.pd_gain[i], y = mt.pd_a[i] - gain * v[i];
                imp[i] += dt * clamp_scalar(y, mt.pd_limit[i]);
                hd[i] += dt * gain * clamp_derivative(y, mt.pd_limit[i]);
                if (gain > 0.0) cost += (dt / gain) * clamp_antiderivative(y, mt.pd_limit[i]);
            }
            if (mt.lo_active[i]) {
                const double gg = (mt.lo_vhat[i] - v[i]) / mt.lo_r[i];
                if (gg > 0.0) {
                    imp[i] += gg;
                    hd[i] += 1.0 / mt.lo_r[i];
                    cost += 0.5 * mt.lo_r[i] * gg * gg;
                }
            }
            if (mt.up_active[i]) {
                const double gg = (mt.up_vhat[i] + v[i]) / mt.up_r[i];
                if (gg > 0.0) {
                    imp[i] -= gg;
                    hd[i] += 1.0 / mt.up_r[i];
                    cost += 0.5 * mt.up_r[i] * gg * gg;
                }
            }
        }
        return cost;
    }

    void matvec(const std::vector<double>& x, std::vector<double>& y) const {
        for (int i = 0; i < n; ++i) {
            double s = 0.0;
            for (int j = 0; j < n; ++j) s += A(i, j) * x[j];
            y[i] = s;
        }
    }
    void jvec(const std::vector<double>& x, std::vector<double>& vc) const {
        const int nc = (int)cs.size();
        vc.assign((size_t)3 * nc, 0.0);
        for (int c = 0; c < nc; ++c)
            for (int r = 0; r < 3; ++r) {
                double s = 0.0;
                const double* row = cs[c].J.data() + (size_t)r * n;
                for (int j = 0; j < n; ++j) s += row[j] * x[j];
                vc[(size_t)c * 3 + r] = s;
            }
    }

    void evaluate(const std::vector<double>& v, Eval& ev) const {
        const int nc = (int)cs.size();
        std::vector<double> r(n), vc;
        for (int i = 0; i < n; ++i) r[i] = v[i] - vstar[i];
        std::vector<double> Ar(n);
        matvec(r, Ar);
        jvec(v, vc);
        ev.gamma.resize(nc);
        ev.G.resize(nc);
        double ccost = 0.0;
        for (int c = 0; c < nc; ++c) {
            ccost += project(&vc[(size_t)c * 3], cs[c].rt, cs[c].rn, cs[c].vhat, cs[c].mu, ev.gamma[c].data(),
                             ev.G[c].data());
        }
        ev.jimp.assign(n, 0.0);
        ev.hdiag.assign(n, 0.0);
        std::vector<double> imp(n);
        const double mcost = eval_terms(v, imp, ev.hdiag);
        for (int j = 0; j < n; ++j) {
            double s = imp[j];
            for (int c = 0; c < nc; ++c)
                for (int k = 0; k < 3; ++k) s += cs[c].J[(size_t)k * n + j] * ev.gamma[c][k];
            ev.jimp[j] = s;
        }
        ev.grad.resize(n);
        double rAr = 0.0;
        for (int i = 0; i < n; ++i) {
            ev.grad[i] = Ar[i] - ev.jimp[i];
            rAr += r[i] * Ar[i];
        }
        ev.cost = 0.5 * rAr + ccost + mcost;
        ev.Av.resize(n);
        matvec(v, ev.Av);
    }

    void participating(const int* art_first_dof, const int* art_dof_count, const int* body_art) {
        part.assign(n, 0);
        for (const Contact& k : cs) {
            const int bs[2] = {k.b0, k.b1};
            for (int b : bs) {
                if (b < 0) continue;
                const int a = body_art[b0 + b];
                if (a < 0) continue;
                const int s = art_first_dof[a] - d0, cnt = art_dof_count[a];
                for (int i = s; i < s + cnt; ++i) part[i] = 1;
            }
            for (int col = 0; col < n; ++col)
                if (k.J[col] != 0.0 || k.J[n + col] != 0.0 || k.J[2 * n + col] != 0.0) part[col] = 1;
        }
        for (int i = 0; i < n; ++i)
            if (mt.pd_active[i] || mt.lo_active[i] || mt.up_active[i]) part[i] = 1;
        Dsc.resize(n);
        for (int i = 0; i < n; ++i) {
            double d = A(i, i);
            if (!(d >= 1e-12 && std::isfinite(d))) d = 1e-12;
            Dsc[i] = 1.0 / std::sqrt(d);
        }
    }
    void opt_norms(const Eval& ev, double* gn, double* tol) const {
        double g = 0, p = 0, j = 0;
        for (int i = 0; i < n; ++i) {
            if (!part[i]) continue;
            g += (Dsc[i] * ev.grad[i]) * (Dsc[i] * ev.grad[i]);
            p += (Dsc[i] * ev.Av[i]) * (Dsc[i] * ev.Av[i]);
            j += (Dsc[i] * ev.jimp[i]) * (Dsc[i] * ev.jimp[i]);
        }
        *gn = std::sqrt(g);
        *tol = O.opt_abs + O.opt_rel * std::max(std::sqrt(p), std::sqrt(j));
    }

    // ---- line searches on phi(alpha) = ell(v + alpha dv)
    struct LineData {
        std::vector<double> dp, vc, dvc;
        double dell0, dell_a0, d2ell_a, m0;
    };
    void line_data(const std::vector<double>& v, const std::vector<double>& dv, const Eval& ev, LineData& L) const {
        L.dp.resize(n);
        matvec(dv, L.dp);
        double dell0 = 0, da0 = 0, d2 = 0, m0 = 0;
        std::vector<double> r(n), Ar(n);
        for (int i = 0; i < n; ++i) r[i] = v[i] - vstar[i];
        matvec(r, Ar);
        for (int i = 0; i < n; ++i) {
            dell0 += ev.grad[i] * dv[i];
            da0 += L.dp[i] * r[i];
            d2 += dv[i] * L.dp[i];
            m0 += r[i] * Ar[i];
        }
        L.dell0 = dell0;
        L.dell_a0 = da0;
        L.d2ell_a = d2;
        L.m0 = 0.5 * m0;
        jvec(v, L.vc);
        jvec(dv, L.dvc);
    }
    void trial(const std::vector<double>& v, const std::vector<double>& dv, const LineData& L, double alpha,
               bool second, double* cost, double* d1, double* d2) const {
        const int nc = (int)cs.size();
        std::vector<double> vt(n), imp(n), hd(n);
        for (int i = 0; i < n; ++i) vt[i] = v[i] + alpha * dv[i];
        double ccost = 0.0, d1c = 0.0, d2c = 0.0;
        for (int c = 0; c < nc; ++c) {
            double vc[3], g[3], G[9];
            for (int k = 0; k < 3; ++k) vc[k] = L.vc[(size_t)c * 3 + k] + alpha * L.dvc[(size_t)c * 3 + k];
            // J (v + alpha dv) formed as J v + alpha J dv
            ccost += project(vc, cs[c].rt, cs[c].rn, cs[c].vhat, cs[c].mu, g, G);
            const double* dvc = &L.dvc[(size_t)c * 3];
            d1c += dvc[0] * g[0] + dvc[1] * g[1] + dvc[2] * g[2];
            if (second)
                for (int r = 0; r < 3; ++r)
                    for (int s = 0; s < 3; ++s) d2c += dvc[r] * G[r * 3 + s] * dvc[s];
        }
        const double mcost = eval_terms(vt, imp, hd);
        *cost = L.m0 + alpha * L.dell_a0 + 0.5 * alpha * alpha * L.d2ell_a + ccost + mcost;
        double a1 = 0.0, dimp = 0.0, dh = 0.0;
        for (int i = 0; i < n; ++i) {
            a1 += L.dp[i] * (vt[i] - vstar[i]);
            dimp += dv[i] * imp[i];
            dh += hd[i] * dv[i] * dv[i];
        }
        *d1 = a1 - d1c - dimp;
        *d2 = second ? L.d2ell_a + d2c + dh : 0.0;
    }

    int exact_search(const std::vector<double>& v, const std::vector<double>& dv, const Eval& ev, double* a_out,
                     int* it_out) const {
        LineData L;
        line_data(v, dv, ev, L);
        const double ell0 = ev.cost, dell0 = L.dell0, amax = kExactAlphaMax, tol = kExactFTol;
        *it_out = 0;
        *a_out = amax;
        if (dell0 >= 0.0 || !std::isfinite(dell0)) return -1;
        double ell, d1, d2;
        trial(v, dv, L, amax, true, &ell, &d1, &d2);
        if (d1 <= 0.0) return 0;
        if (-dell0 < O.cost_abs + O.cost_rel * ell) {
            *a_out = 1.0;
            return 0;
        }
        const double scale = -dell0;
        double f_hi = d1 / scale;
        if (!std::isfinite(f_hi) || f_hi <= 0.0) return -3;
        double guess = -dell0 / d2;
        const double h = amax;
        const double theta = dell0 + d1 - 3.0 * (ell - ell0) / h;
        const double disc = theta * theta - dell0 * d1;
        if (std::isfinite(disc) && disc >= 0.0) {
            const double gm = std::sqrt(disc), den = d1 - dell0 + 2.0 * gm;
            if (den != 0.0) {
                const double cubic = h - h * (d1 + gm - theta) / den;
                if (std::isfinite(cubic) && 0.0 < cubic && cubic <= h) guess = cubic;
            }
        }
        if (guess > amax) guess = amax;
        if (!std::isfinite(guess) || guess < 0.0 || guess > amax || tol * guess <= 0.0) return -4;
        double xlo = 0.0, xhi = amax, root = guess, mdx = -amax, mdx_prev = -amax;
        const double xtol = tol * guess;
        for (int it = 1;; ++it) {
            double c, dd1, dd2;
            trial(v, dv, L, root, true, &c, &dd1, &dd2);
            const double f = dd1 / scale, df = dd2 / scale;
            if ((f < 0.0 && f_hi >= 0.0) || (f >= 0.0 && f_hi < 0.0)) xlo = root;
            else { xhi = root; f_hi = f; }
            *it_out = it;
            if (std::fabs(f) < tol) { *a_out = root; return 0; }
            if (it >= O.ls_max_iterations) { *a_out = root; return -5; }
            const bool slow = 2.0 * std::fabs(f) > std::fabs(mdx_prev * df);
            mdx_prev = mdx;
            const double bdx = 0.5 * (xlo - xhi), broot = xlo - bdx;
            double ndx = 0.0, nroot = root;
            if (std::fabs(df) > 0.0) { ndx = f / df; nroot = root - ndx; }
            const bool bis = slow || nroot < xlo || nroot > xhi || !std::isfinite(nroot) || !std::isfinite(ndx);
            const double nxt = bis ? broot : nroot, dx = bis ? bdx : ndx;
            if (std::fabs(dx) < xtol) { *a_out = nxt; return 0; }
            root = nxt;
            mdx = dx;
        }
    }

    int armijo_search(const std::vector<double>& v, const std::vector<double>& dv, const Eval& ev, double* a_out,
                      int* it_out, bool* accepted) const {
        LineData L;
        line_data(v, dv, ev, L);
        const double ell0 = ev.cost, dell0 = L.dell0, amax = 1.0 / O.rho;
        *it_out = 0;
        *accepted = false;
        *a_out = 1.0;
        if (dell0 >= 0.0 || !std::isfinite(dell0)) return -1;
        auto ok = [&](double a, double e) { return e < ell0 + O.armijo_c * a * dell0; };
        double alpha = amax, ell, d1, d2;
        trial(v, dv, L, alpha, false, &ell, &d1, &d2);
        const double scale = std::max(1.0, 0.5 * (std::fabs(ell) + std::fabs(ell0)));
        const double slop = (O.relative_slop / 10.0) * scale;
        *a_out = alpha;
        if (d1 < 0.0 || d1 < slop) { *accepted = true; return 0; }
        double ap = alpha, ep = ell;
        if (O.ls_max_iterations <= 1) return -2;
        for (int it = 1;; ++it) {
            alpha *= O.rho;
            trial(v, dv, L, alpha, false, &ell, &d1, &d2);
            const double den = alpha - ap;
            const double dapprox = std::fabs(den) > 0.0 ? (ell - ep) / den : 0.0;
            *it_out = it;
            if (std::fabs(dapprox) < slop) { *a_out = alpha; *accepted = true; return 0; }
            if (ell > ep && ok(alpha, ell)) {
                *a_out = ok(ap, ep) ? ap : alpha;
                *accepted = true;
                return 0;
            }
            if (it >= O.ls_max_iterations - 1) {
                *a_out = alpha;
                if (ok(alpha, ell)) { *it_out = O.ls_max_iterations; *accepted = true; return 0; }
                return -2;
            }
            ap = alpha;
            ep = ell;
        }
    }

    int monotone_search(const std::vector<double>& v, const std::vector<double>& dv, const Eval& ev, double* a_out,
                        int* it_out, bool* accepted) const {
        LineData L;
        line_data(v, dv, ev, L);
        const double cur = ev.cost, tol = kMonoRelaxA + kMonoRelaxR * std::fabs(cur);
        double alpha = 1.0;
        *accepted = false;
        *it_out = 0;
        for (int k = 0; k < O.ls_max_iterations; ++k) {
            double ell, d1, d2;
            trial(v, dv, L, alpha, false, &ell, &d1, &d2);
            *it_out = k + 1;
            if (ell <= cur + tol) { *a_out = alpha; *accepted = true; return 0; }
            alpha *= kMonoDecay;
            if (alpha < kMonoMinAlpha) { *a_out = alpha; return -2; }
        }
        *a_out = alpha;
        return 0;
    }

    // SAP Newton solve; returns the iteration count, fills v
    int solve(std::vector<double>& v, const double* v_guess, int* ls_total, int* converged) {
        const bool unconstrained = cs.empty() && std::none_of(mt.pd_active.begin(), mt.pd_active.end(), [](int x) { return x; }) &&
                                   std::none_of(mt.lo_active.begin(), mt.lo_active.end(), [](int x) { return x; }) &&
                                   std::none_of(mt.up_active.begin(), mt.up_active.end(), [](int x) { return x; });
        *ls_total = 0;
        *converged = 1;
        if (unconstrained) {
            v = vstar;
            return 0;
        }
        v.resize(n);
        for (int i = 0; i < n; ++i) v[i] = v_guess ? v_guess[i] : v0[i];
        for (int i = 0; i < n; ++i)
            if (!part[i]) v[i] = vstar[i];
        const double cost_min_alpha = O.line_search == 2 ? 0.0 : 0.5;
        double alpha = 1.0, prev = 0.0;
        Eval ev;
        std::vector<double> dv(n);
        Mat H;
        for (int it = 0;; ++it) {
            evaluate(v, ev);
            double gn, tol;
            opt_norms(ev, &gn, &tol);
            bool cost_ok = false;
            if (it > 0 && alpha > cost_min_alpha) {
                const double sc = 0.5 * (std::fabs(ev.cost) + std::fabs(prev));
                cost_ok = std::fabs(ev.cost - prev) < O.cost_abs + O.cost_rel * sc;
            }
            if (gn <= tol || cost_ok) return it;
            if (it >= O.max_iterations) {
                *converged = 0;
                return it;
            }
            H = A;
            for (int c = 0; c < (int)cs.size(); ++c) {
                const double* J = cs[c].J.data();
                const double* G = ev.G[c].data();
                // J^T G J, only the nonzero columns of this contact
                for (int a = 0; a < n; ++a) {
                    const double ja[3] = {J[a], J[n + a], J[2 * n + a]};
                    if (ja[0] == 0.0 && ja[1] == 0.0 && ja[2] == 0.0) continue;
                    double gj[3];
                    for (int s = 0; s < 3; ++s) gj[s] = ja[0] * G[s] + ja[1] * G[3 + s] + ja[2] * G[6 + s];
                    for (int b = 0; b < n; ++b) H(a, b) += gj[0] * J[b] + gj[1] * J[n + b] + gj[2] * J[2 * n + b];
                }
            }
            for (int i = 0; i < n; ++i) H(i, i) += ev.hdiag[i] + O.diag_shift;
            for (int i = 0; i < n; ++i) dv[i] = -ev.grad[i];
            chol_solve(H, dv);
            double a_new = 1.0;
            int lsit = 0;
            bool acc = false;
            if (O.line_search == 1) {
                const int st = exact_search(v, dv, ev, &a_new, &lsit);
                acc = st == 0;
            } else if (O.line_search == 0) {
                armijo_search(v, dv, ev, &a_new, &lsit, &acc);
            } else {
                monotone_search(v, dv, ev, &a_new, &lsit, &acc);
            }
            if (acc) {
                alpha = a_new;
                prev = ev.cost;
                for (int i = 0; i < n; ++i) v[i] += alpha * dv[i];
                *ls_total += lsit;
            } else if (O.line_search == 2) {
                alpha = a_new;
                prev = ev.cost;
                *ls_total += lsit;
            }
        }
    }

    static Q quat_sap_euler(Q q, V3 wv, double h) {
        const double hd = 0.5 * h;
        double x = q.x + hd * (q.w * wv.x + wv.y * q.z - wv.z * q.y);
        double y = q.y + hd * (q.w * wv.y + wv.z * q.x - wv.x * q.z);
        double z = q.z + hd * (q.w * wv.z + wv.x * q.y - wv.y * q.x);
        double s = q.w - hd * (wv.x * q.x + wv.y * q.y + wv.z * q.z);
        const double nr = std::sqrt(x * x + y * y + z * z + s * s);
        if (nr > 0.0) { x /= nr; y /= nr; z /= nr; s /= nr; }
        return Q{x, y, z, s};
    }
    static Q quat_midpoint(Q q, V3 wm, double h) {
        const double nw = std::sqrt(dot(wm, wm));
        const double ang = nw * h;
        if (ang > 1e-12) {
            const V3 ax = wm * (1.0 / nw);
            const double s = std::sin(0.5 * ang), c = std::cos(0.5 * ang);
            Q o = qmul(Q{ax.x * s, ax.y * s, ax.z * s, c}, q);
            const double nr = std::sqrt(o.x * o.x + o.y * o.y + o.z * o.z + o.w * o.w);
            if (nr > 0.0) o = Q{o.x / nr, o.y / nr, o.z / nr, o.w / nr};
            return o;
        }
        return q;
    }

    void integrate_and_fk(const float* joint_q_all, const std::vector<double>& v, Out& out) {
        std::vector<double> q(c1 - c0), qd(n, 0.0);
        for (int i = 0; i < c1 - c0; ++i) q[i] = joint_q_all[c0 + i];
        const bool mid = O.integration == 0;
        for (int j = j0; j < j1; ++j) {
            const int jt = M.joint_type[j];
            const int qs = M.joint_q_start[j] - c0, ds = M.joint_qd_start[j] - d0;
            const int nd = M.joint_qd_start[j + 1] - M.joint_qd_start[j];
            if (jt == kBall) {
                const Q q0{q[qs], q[qs + 1], q[qs + 2], q[qs + 3]};
                const V3 wn = v3(v[ds], v[ds + 1], v[ds + 2]);
                const Q qn = mid ? quat_midpoint(q0, 0.5 * (v3(v0[ds], v0[ds + 1], v0[ds + 2]) + wn), dt)
                                 : quat_sap_euler(q0, wn, dt);
                q[qs] = qn.x; q[qs + 1] = qn.y; q[qs + 2] = qn.z; q[qs + 3] = qn.w;
                for (int k = 0; k < 3; ++k) qd[ds + k] = v[ds + k];
            } else if (jt == kFree || jt == kDistance) {
                const Q q0{q[qs + 3], q[qs + 4], q[qs + 5], q[qs + 6]};
                const V3 wn = v3(v[ds], v[ds + 1], v[ds + 2]), von = v3(v[ds + 3], v[ds + 4], v[ds + 5]);
                V3 vom;
                Q qn;
                if (mid) {
                    const V3 wm = 0.5 * (v3(v0[ds], v0[ds + 1], v0[ds + 2]) + wn);
                    vom = 0.5 * (v3(v0[ds + 3], v0[ds + 4], v0[ds + 5]) + von);
                    qn = quat_midpoint(q0, wm, dt);
                } else {
                    vom = von;
                    qn = quat_sap_euler(q0, wn, dt);
                }
                q[qs] += dt * vom.x; q[qs + 1] += dt * vom.y; q[qs + 2] += dt * vom.z;
                q[qs + 3] = qn.x; q[qs + 4] = qn.y; q[qs + 5] = qn.z; q[qs + 6] = qn.w;
                const V3 r = qrot(qn, ld3(M.body_com + (size_t)M.joint_child[j] * 3));
                const V3 vc = von + cross(wn, r);
                qd[ds] = vc.x; qd[ds + 1] = vc.y; qd[ds + 2] = vc.z;
                qd[ds + 3] = wn.x; qd[ds + 4] = wn.y; qd[ds + 5] = wn.z;
            } else if (jt != kFixed) {
                for (int a = 0; a < nd; ++a) {
                    const double vm = mid ? 0.5 * (v0[ds + a] + v[ds + a]) : v[ds + a];
                    q[qs + a] += dt * vm;
                    qd[ds + a] = v[ds + a];
                }
            }
        }
        for (int i = 0; i < c1 - c0; ++i) out.joint_q[c0 + i] = (float)q[i];
        for (int i = 0; i < n; ++i) out.joint_qd[d0 + i] = (float)qd[i];
        // public FK (COM-free twist, linear first)
        std::vector<TF> X(nb);
        std::vector<SV> T(nb);  // w = linear, v = angular (public order kept in out)
        for (int j = j0; j < j1; ++j) {
            const int parent = M.joint_parent[j], child = M.joint_child[j] - b0, jt = M.joint_type[j];
            const float* xp = M.joint_X_p + (size_t)j * 7;
            const TF Xpj = TF{ld3(xp), Q{xp[3], xp[4], xp[5], xp[6]}};
            const TF Xcj = ld_tf(M.joint_X_c + (size_t)j * 7);
            auto tmul = [](TF a, TF b) { return TF{a.p + qrot_unit(a.q, b.p), qmul(a.q, b.q)}; };
            auto tinv = [](TF a) {
                const Q qi = qconj(a.q);
                return TF{-qrot_unit(qi, a.p), qi};
            };
            TF Xwpj = Xpj;
            V3 vlin{}, vang{};
            if (parent >= 0) {
                const int lp = parent - b0;
                Xwpj = tmul(X[lp], Xpj);
                const V3 rp = Xwpj.p - (X[lp].p + qrot_unit(X[lp].q, ld3(M.body_com + (size_t)parent * 3)));
                const V3 wp = T[lp].v;
                vlin = T[lp].w + cross(wp, rp);
                vang = wp;
            }
            const int qs = M.joint_q_start[j] - c0, ds = M.joint_qd_start[j] - d0;
            TF Xj{};
            V3 jl{}, ja{};
            if (jt == kPrismatic) {
                const V3 ax = ld3(M.joint_axis + (size_t)(d0 + ds) * 3);
                Xj = TF{ax * q[qs], Q{}};
                jl = ax * qd[ds];
            } else if (jt == kRevolute) {
                const V3 ax = ld3(M.joint_axis + (size_t)(d0 + ds) * 3);
                Xj = TF{V3{}, q_axis_angle(ax, q[qs])};
                ja = ax * qd[ds];
            } else if (jt == kBall) {
                Xj = TF{V3{}, Q{q[qs], q[qs + 1], q[qs + 2], q[qs + 3]}};
                ja = v3(qd[ds], qd[ds + 1], qd[ds + 2]);
            } else if (jt == kFree || jt == kDistance) {
                Xj = TF{v3(q[qs], q[qs + 1], q[qs + 2]), Q{q[qs + 3], q[qs + 4], q[qs + 5], q[qs + 6]}};
                jl = v3(qd[ds], qd[ds + 1], qd[ds + 2]);
                ja = v3(qd[ds + 3], qd[ds + 4], qd[ds + 5]);
            } else if (jt == kD6) {
                // sap_kinematics.py:170-218
                const int nl = M.joint_dof_dim[2 * j], na = M.joint_dof_dim[2 * j + 1];
                V3 a[6], cur[3];
                for (int k = 0; k < nl + na; ++k) a[k] = ld3(M.joint_axis + (size_t)(d0 + ds + k) * 3);
                V3 pos{};
                for (int k = 0; k < nl; ++k) {
                    pos = pos + a[k] * q[qs + k];
                    jl = jl + a[k] * qd[ds + k];
                }
                Xj = TF{pos, d6_rotation(na, a + nl, q.data() + qs + nl, cur)};
                for (int k = 0; k < na; ++k) ja = ja + cur[k] * qd[ds + nl + k];
            }
            X[child] = tmul(tmul(Xwpj, Xj), tinv(Xcj));
            T[child] = SV{vlin + qrot_unit(Xwpj.q, jl), vang + qrot_unit(Xwpj.q, ja)};
        }
        for (int b = 0; b < nb; ++b) {
            float* o = out.body_q + (size_t)(b0 + b) * 7;
            o[0] = (float)X[b].p.x; o[1] = (float)X[b].p.y; o[2] = (float)X[b].p.z;
            o[3] = (float)X[b].q.x; o[4] = (float)X[b].q.y; o[5] = (float)X[b].q.z; o[6] = (float)X[b].q.w;
            float* t = out.body_qd + (size_t)(b0 + b) * 6;
            t[0] = (float)T[b].w.x; t[1] = (float)T[b].w.y; t[2] = (float)T[b].w.z;
            t[3] = (float)T[b].v.x; t[4] = (float)T[b].v.y; t[5] = (float)T[b].v.z;
        }
    }
};

}  // namespace

extern "C" {

/*
 * One SAP step for every world on the CPU (threads: OpenMP threads, 0 = all).  Model / state /
 * control arrays are the SapModel / SapState / SapControl fields (host, float32); contacts are
 * world-grouped with world_start (world_count + 1).  v_guess (dof_count, may be null) is the
 * warm start.  Outputs: the new public state, the SAP velocity (float64) and per-world
 * statistics.  Returns 0.
 */
int sapx_cpu_step(const Model* m, const State* st, const Control* ctl, const Contacts* c, const Opts* o, double dt,
                  const double* v_guess, Out* out, int threads) {
    const Model& M = *m;
    // dof maps (contact_solve.py:3681-3720)
    std::vector<int> coord(M.dof_count, -1), target(M.dof_count), limit_ok(M.dof_count, 0);
    for (int i = 0; i < M.dof_count; ++i) target[i] = i;
    for (int j = 0; j < M.joint_count; ++j) {
        const int jt = M.joint_type[j], ds = M.joint_qd_start[j];
        if (jt == kFree || jt == kDistance)
            for (int a = 0; a < 3; ++a) {
                target[ds + a] = ds + a + 3;
                target[ds + a + 3] = ds + a;
            }
        if (jt != kPrismatic && jt != kRevolute && jt != kD6) continue;
        const int ac = M.joint_dof_dim[2 * j] + M.joint_dof_dim[2 * j + 1];
        for (int a = 0; a < ac; ++a) {
            coord[ds + a] = M.joint_q_start[j] + a;
            if (ac == 1) limit_ok[ds + a] = 1;
        }
    }
    // articulation dof ranges, body -> articulation
    std::vector<int> art_first(M.articulation_count, 0), art_count(M.articulation_count, 0),
        body_art(M.body_count, -1);
    for (int a = 0; a < M.articulation_count; ++a) {
        const int js = M.articulation_start[a], je = M.articulation_start[a + 1];
        int first = 1 << 30, cnt = 0;
        for (int j = js; j < je; ++j) {
            const int nd = M.joint_dof_dim[2 * j] + M.joint_dof_dim[2 * j + 1];
            if (nd > 0) first = std::min(first, M.joint_qd_start[j]);
            cnt += nd;
        }
        if (first == (1 << 30) || cnt <= 0) continue;
        art_first[a] = first;
        art_count[a] = cnt;
        for (int j = js; j < je; ++j) body_art[M.joint_child[j]] = a;
    }
    const int W = M.world_count;
    if (threads <= 0) threads = omp_get_max_threads();
#pragma omp parallel for schedule(dynamic, 1) num_threads(threads)
    for (int w = 0; w < W; ++w) {
        World wd(M, *o, w, dt);
        std::vector<double> fj;
        std::vector<SV> fb;
        wd.boundary(*st, *ctl, fj, fb);
        wd.free_motion(st->joint_q, fj, fb);
        wd.contacts(*c);
        wd.model_terms(st->joint_q, *ctl, coord.data(), target.data(), limit_ok.data());
        wd.participating(art_first.data(), art_count.data(), body_art.data());
        std::vector<double> v;
        int ls = 0, conv = 1;
        const int it = wd.solve(v, v_guess ? v_guess + wd.d0 : nullptr, &ls, &conv);
        wd.integrate_and_fk(st->joint_q, v, *out);
        for (int i = 0; i < wd.n; ++i) out->v_sap[wd.d0 + i] = v[i];
        out->iterations[w] = it;
        out->ls_iterations[w] = ls;
        out->contact_count[w] = (int)wd.cs.size();
        out->converged[w] = conv;
    }
    return 0;
}

}  // extern "C"

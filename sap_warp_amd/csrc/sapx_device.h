// Device-side types and small math helpers shared by the sapx kernels.
//
// Conventions: transforms are [p(3), q(4) xyzw]; spatial vectors are
// angular-first [w, v] inside the solver (the reference's SAP buffers,
// sim/sap_helpers.py:771-785) and linear-first in the public body_qd.
#pragma once

#include <hip/hip_runtime.h>
#include <stdint.h>

namespace sapx {

constexpr int kJointPrismatic = 0;
constexpr int kJointRevolute = 1;
constexpr int kJointBall = 2;
constexpr int kJointFixed = 3;
constexpr int kJointFree = 4;
constexpr int kJointDistance = 5;
constexpr int kJointD6 = 6;

constexpr int kGeoPlane = 1;
constexpr int kGeoSphere = 3;
constexpr int kGeoCapsule = 4;
constexpr int kGeoEllipsoid = 5;
constexpr int kGeoCylinder = 6;
constexpr int kGeoBox = 7;
constexpr int kGeoCone = 9;

// Shape-type pairs (sorted, ta <= tb) with an analytic contact routine (narrow_phase.py:304-567).
__host__ __device__ inline bool pair_primitive(int ta, int tb) {
    if (ta == kGeoPlane)
        return tb == kGeoSphere || tb == kGeoCapsule || tb == kGeoBox || tb == kGeoEllipsoid || tb == kGeoCylinder;
    if (ta == kGeoSphere) return tb == kGeoSphere || tb == kGeoCapsule || tb == kGeoBox || tb == kGeoCylinder;
    return ta == kGeoCapsule && tb == kGeoCapsule;
}
// Shapes with a support map (support.py:94-258).
__host__ __device__ inline bool support_type(int t) {
    return t == kGeoSphere || t == kGeoCapsule || t == kGeoEllipsoid || t == kGeoCylinder || t == kGeoBox ||
           t == kGeoCone;
}
// Sorted pairs the primitive routines leave to the MPR / GJK path (narrow_phase.py:570-575): two
// support-mapped shapes, or an infinite plane (turned into a box) against a cone.
__host__ __device__ inline bool pair_convex(int ta, int tb) {
    return !pair_primitive(ta, tb) && ((support_type(ta) && support_type(tb)) || (ta == kGeoPlane && tb == kGeoCone));
}

constexpr int kMaxTemplate = 64;  // joints / bodies per world (64-bit masks)
constexpr int kMaxDof = 64;       // dofs per world of the one-wave kernel (one row per lane)
constexpr int kMaxDofFused = 256; // dofs per world of the 256-thread kernel (one dof per thread)

// --------------------------------------------------------------------------
// Device model: per-world topology template (local indices, shared by all
// worlds) + global per-item parameter arrays (world-major).
// --------------------------------------------------------------------------
enum TopoField {
    TP_JTYPE = 0, TP_JPARENT, TP_JCHILD, TP_JQS, TP_JDS, TP_LEVEL_START, TP_LEVEL_JOINT, TP_DOF_JOINT,
    TP_DOF_COORD, TP_DOF_TARGET, TP_DOF_LIMIT_OK, TP_BODY_ART_START, TP_BODY_ART_COUNT, TP_CHILD_START,
    TP_CHILD, TP_BODY_ANC, TP_BODY_ANC_DOF, TP_DOF_ANC, TP_BODY_UP, TP_BODY_ROOT, TP_JLIN, TP_COUNT
};

struct DevModel {
    int W;   // worlds
    int J;   // joints per world
    int N;   // dofs per world
    int NQ;  // coords per world
    int NB;  // bodies per world
    int S;   // shapes per world
    int P;   // candidate pairs per world
    int num_levels;
    int fk_rounds;  // pointer-jumping rounds over the root paths: ceil(log2(depth in bodies))
    int total_shapes;
    int global_shape_base;  // first global (world -1) shape index = W*S
    int has_convex;         // some candidate pair is capsule-box or box-box (MPR / GJK path)

    // the per-world topology template packed into one int array (copied into LDS by
    // the step kernel); topo_off[k] = first int of field k (see enum TopoField)
    const int* topo;
    int topo_ints;
    int topo_off[TP_COUNT];

    // template, local indices
    const int* t_jtype;
    const int* t_jparent;   // local body or -1
    const int* t_jchild;    // local body
    const int* t_jqs;       // J+1 local coord starts
    const int* t_jds;       // J+1 local dof starts
    const int* t_level_start;  // num_levels+1
    const int* t_level_joint;  // J
    const int* t_dof_joint;    // N
    const int* t_dof_coord;    // N: local coord or -1
    const int* t_dof_target;   // N: local dof of the drive target (free joints swap blocks)
    const int* t_dof_limit_ok; // N
    const int* t_body_joint;   // NB: joint with this child body, -1 if none
    const unsigned long long* t_body_anc;  // NB: bit j = joint j on the root path (incl. own)
    const int* t_body_art_start;  // NB: first local dof of the body's articulation (-1 none)
    const int* t_body_art_count;  // NB
    const int* t_child_start;     // NB+1: CSR of child bodies
    const int* t_child;           // child bodies, grouped by parent
    const int* t_pairs;           // P x 2: >=0 local shape, <0 => global shape (-v-1)

    // per-world model parameters packed in the one-wave kernel's LDS order, wp_stride floats
    // per world (a multiple of 4): X_p (J x 8), X_c (J x 8), axes (N x 4), COMs (NB x 4),
    // inertias (NB x 9), masses (NB)
    const float* world_params;
    int wp_stride;

    // global per-item arrays
    const int* j_xc_identity;  // W*J
    const float* joint_axis;   // W*N x 3
    const float* joint_X_p;    // W*J x 7
    const float* joint_X_c;    // W*J x 7
    const float* armature;     // W*N
    const float* body_mass;    // W*NB
    const float* body_inertia; // W*NB x 9
    const float* body_com;     // W*NB x 3
    const float* gravity;      // W x 3
    const int* target_mode;    // W*N
    const float* target_ke;
    const float* target_kd;
    const float* effort_limit;
    const float* limit_lower;
    const float* limit_upper;
    const float* limit_ke;
    const float* limit_kd;
    const int* shape_body;     // total_shapes
    const int* shape_type;
    const float* shape_transform;  // x7
    const float* shape_scale;      // x3
    const float* shape_margin;
    const float* shape_gap;
    const float* shape_radius;
    const float* shape_mu;
    const float* shape_ke;
    const float* shape_tau;
};

// --------------------------------------------------------------------------
// float3 / quaternion helpers (float32)
// --------------------------------------------------------------------------
struct V3 {
    float x, y, z;
};
struct Q4 {
    float x, y, z, w;
};

__device__ __forceinline__ V3 v3(float x, float y, float z) { return V3{x, y, z}; }
__device__ __forceinline__ V3 operator+(V3 a, V3 b) { return v3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ V3 operator-(V3 a, V3 b) { return v3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ V3 operator-(V3 a) { return v3(-a.x, -a.y, -a.z); }
__device__ __forceinline__ V3 operator*(V3 a, float s) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ V3 operator*(float s, V3 a) { return v3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ float dot(V3 a, V3 b) { return a.x * b.x + a.y * b.y + a.z * b.z; }
__device__ __forceinline__ V3 cross(V3 a, V3 b) {
    return v3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ float length(V3 a) { return sqrtf(dot(a, a)); }

__device__ __forceinline__ V3 qrot(Q4 q, V3 v) {
    V3 qv = v3(q.x, q.y, q.z);
    V3 t = cross(qv, v) * 2.0f;
    return v + t * q.w + cross(qv, t);
}
__device__ __forceinline__ Q4 qconj(Q4 q) { return Q4{-q.x, -q.y, -q.z, q.w}; }
__device__ __forceinline__ Q4 qmul(Q4 a, Q4 b) {
    return Q4{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
__device__ __forceinline__ Q4 qnormalize(Q4 q) {
    float n = sqrtf(q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w);
    if (n <= 0.0f) return Q4{0.0f, 0.0f, 0.0f, 1.0f};
    float inv = 1.0f / n;
    return Q4{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
__device__ __forceinline__ Q4 q_axis_angle(V3 a, float angle) {
    float s = sinf(0.5f * angle);
    float c = cosf(0.5f * angle);
    return Q4{a.x * s, a.y * s, a.z * s, c};
}

struct TF {
    V3 p;
    Q4 q;
};
__device__ __forceinline__ TF load_tf(const float* a) {
    return TF{v3(a[0], a[1], a[2]), Q4{a[3], a[4], a[5], a[6]}};
}
__device__ __forceinline__ void store_tf(float* a, TF t) {
    a[0] = t.p.x; a[1] = t.p.y; a[2] = t.p.z;
    a[3] = t.q.x; a[4] = t.q.y; a[5] = t.q.z; a[6] = t.q.w;
}
// composition with normalisation (free-motion kinematics, sap_helpers.py:670-676)
__device__ __forceinline__ TF tf_compose_n(TF a, TF b) {
    Q4 aq = qnormalize(a.q);
    Q4 bq = qnormalize(b.q);
    return TF{a.p + qrot(aq, b.p), qnormalize(qmul(aq, bq))};
}
__device__ __forceinline__ TF tf_inverse_n(TF x) {
    Q4 qi = qconj(qnormalize(x.q));
    return TF{-qrot(qi, x.p), qi};
}
// plain composition (public FK / collision, no normalisation)
__device__ __forceinline__ TF tf_mul(TF a, TF b) { return TF{a.p + qrot(a.q, b.p), qmul(a.q, b.q)}; }
__device__ __forceinline__ TF tf_inverse(TF x) {
    Q4 qi = qconj(x.q);
    return TF{-qrot(qi, x.p), qi};
}
__device__ __forceinline__ bool tf_is_identity(const float* a) {
    return a[0] == 0.0f && a[1] == 0.0f && a[2] == 0.0f && a[3] == 0.0f && a[4] == 0.0f && a[5] == 0.0f &&
           a[6] == 1.0f;
}

// ---------------------------------------------------------------------------
// float64 transforms for the free-motion forward kinematics (the reference
// evaluates it with transformd / quatd, sap_helpers.py:575-683)
// ---------------------------------------------------------------------------
struct D3 {
    double x, y, z;
};
struct DQ {
    double x, y, z, w;
};
struct DTF {
    D3 p;
    DQ q;
};

// Component-wise selects: C++ ?: between two structs is emitted as a copy through a stack slot,
// which the backend keeps in scratch memory.
__device__ __forceinline__ V3 vsel(bool c, V3 a, V3 b) { return V3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ D3 dsel(bool c, D3 a, D3 b) { return D3{c ? a.x : b.x, c ? a.y : b.y, c ? a.z : b.z}; }
__device__ __forceinline__ DTF dtfsel(bool c, DTF a, DTF b) {
    return DTF{dsel(c, a.p, b.p), DQ{c ? a.q.x : b.q.x, c ? a.q.y : b.q.y, c ? a.q.z : b.q.z, c ? a.q.w : b.q.w}};
}
__device__ __forceinline__ D3 d3(double x, double y, double z) { return D3{x, y, z}; }
__device__ __forceinline__ D3 dadd(D3 a, D3 b) { return d3(a.x + b.x, a.y + b.y, a.z + b.z); }
__device__ __forceinline__ D3 dsub(D3 a, D3 b) { return d3(a.x - b.x, a.y - b.y, a.z - b.z); }
__device__ __forceinline__ D3 dscale(D3 a, double s) { return d3(a.x * s, a.y * s, a.z * s); }
__device__ __forceinline__ D3 dcross(D3 a, D3 b) {
    return d3(a.y * b.z - a.z * b.y, a.z * b.x - a.x * b.z, a.x * b.y - a.y * b.x);
}
__device__ __forceinline__ D3 dqrot_unit(DQ q, D3 v) {
    D3 qv = d3(q.x, q.y, q.z);
    D3 t = dscale(dcross(qv, v), 2.0);
    return dadd(dadd(v, dscale(t, q.w)), dcross(qv, t));
}
__device__ __forceinline__ DQ dqmul(DQ a, DQ b) {
    return DQ{a.w * b.x + a.x * b.w + a.y * b.z - a.z * b.y, a.w * b.y - a.x * b.z + a.y * b.w + a.z * b.x,
              a.w * b.z + a.x * b.y - a.y * b.x + a.z * b.w, a.w * b.w - a.x * b.x - a.y * b.y - a.z * b.z};
}
// q / |q| through one reciprocal square root (within two ulps of the four divisions)
__device__ __forceinline__ DQ dqnormalize(DQ q) {
    const double n2 = q.x * q.x + q.y * q.y + q.z * q.z + q.w * q.w;
    if (n2 <= 0.0) return DQ{0.0, 0.0, 0.0, 1.0};
    const double inv = rsqrt(n2);
    return DQ{q.x * inv, q.y * inv, q.z * inv, q.w * inv};
}
__device__ __forceinline__ DQ dq_axis_angle(D3 a, double angle) {
    double sn, cs;
    sincos(0.5 * angle, &sn, &cs);
    return DQ{a.x * sn, a.y * sn, a.z * sn, cs};
}
__device__ __forceinline__ DTF dtf_from(TF t) {
    return DTF{d3(t.p.x, t.p.y, t.p.z), DQ{t.q.x, t.q.y, t.q.z, t.q.w}};
}
__device__ __forceinline__ TF tf_from(DTF t) {
    return TF{v3((float)t.p.x, (float)t.p.y, (float)t.p.z), Q4{(float)t.q.x, (float)t.q.y, (float)t.q.z, (float)t.q.w}};
}
__device__ __forceinline__ DTF load_dtf(const double* a) { return DTF{d3(a[0], a[1], a[2]), DQ{a[3], a[4], a[5], a[6]}}; }
__device__ __forceinline__ void store_dtf(double* a, DTF t) {
    a[0] = t.p.x; a[1] = t.p.y; a[2] = t.p.z;
    a[3] = t.q.x; a[4] = t.q.y; a[5] = t.q.z; a[6] = t.q.w;
}
__device__ __forceinline__ DTF dtf_compose_n(DTF a, DTF b) {
    DQ aq = dqnormalize(a.q);
    DQ bq = dqnormalize(b.q);
    return DTF{dadd(a.p, dqrot_unit(aq, b.p)), dqnormalize(dqmul(aq, bq))};
}
// composition of transforms whose rotations are already unit (only the product is
// renormalised): normalising a unit float64 quaternion again changes it by at most an ulp
__device__ __forceinline__ DTF dtf_compose_u(DTF a, DTF b) {
    return DTF{dadd(a.p, dqrot_unit(a.q, b.p)), dqnormalize(dqmul(a.q, b.q))};
}
__device__ __forceinline__ DTF dtf_inverse_n(DTF x) {
    DQ n = dqnormalize(x.q);
    DQ qi = DQ{-n.x, -n.y, -n.z, n.w};
    D3 p = dqrot_unit(qi, x.p);
    return DTF{d3(-p.x, -p.y, -p.z), qi};
}
// x.p + rotate(normalize(x.q), p)  (_transform_point_d, sap_helpers.py:240-266)
__device__ __forceinline__ D3 dtf_point(DTF x, D3 p) { return dadd(x.p, dqrot_unit(dqnormalize(x.q), p)); }

// ---------------------------------------------------------------------------
// D6 joints: up to three linear axes fixed in the joint frame, then up to three
// angular axes.  One angular axis turns by its angle; two or three are first
// replaced by the frame that the matrix of their columns defines (the third
// column a0 x a1 for two), then applied in turn, each later axis carried by the
// rotations before it (sap_helpers.py:1109-1161, 1247-1286).  float64, as the
// reference's free motion.
// ---------------------------------------------------------------------------
// unit quaternion of the matrix [c0 c1 c2]: trace branch, else the largest diagonal entry's
// (sap_helpers.py:607-652)
__device__ inline DQ dq_from_cols(D3 c0, D3 c1, D3 c2) {
    const double tr = c0.x + c1.y + c2.z;
    DQ q;
    if (tr > 0.0) {
        const double s = sqrt(tr + 1.0) * 2.0;
        q = DQ{(c1.z - c2.y) / s, (c2.x - c0.z) / s, (c0.y - c1.x) / s, 0.25 * s};
    } else if (c0.x > c1.y && c0.x > c2.z) {
        const double s = sqrt(1.0 + c0.x - c1.y - c2.z) * 2.0;
        q = DQ{0.25 * s, (c1.x + c0.y) / s, (c2.x + c0.z) / s, (c1.z - c2.y) / s};
    } else if (c1.y > c2.z) {
        const double s = sqrt(1.0 + c1.y - c0.x - c2.z) * 2.0;
        q = DQ{(c1.x + c0.y) / s, 0.25 * s, (c2.y + c1.z) / s, (c2.x - c0.z) / s};
    } else {
        const double s = sqrt(1.0 + c2.z - c0.x - c1.y) * 2.0;
        q = DQ{(c2.x + c0.z) / s, (c2.y + c1.z) / s, 0.25 * s, (c0.y - c1.x) / s};
    }
    return dqnormalize(q);
}
struct D6Ang {
    DQ rot;           // rotation of the angular coordinates
    D3 a0, a1, a2;    // current axis of angular dof 0, 1, 2 (joint frame)
};
// na angular axes c0..c2 (unused ones ignored) at angles q0..q2
__device__ inline D6Ang d6_angular(int na, D3 c0, D3 c1, D3 c2, double q0, double q1, double q2) {
    D6Ang r{DQ{0.0, 0.0, 0.0, 1.0}, c0, c1, c2};
    if (na == 1) {
        r.rot = dq_axis_angle(c0, q0);
    } else if (na >= 2) {
        const DQ off = dq_from_cols(c0, c1, na == 2 ? dcross(c0, c1) : c2);
        r.a0 = dqrot_unit(off, d3(1.0, 0.0, 0.0));
        const D3 l1 = dqrot_unit(off, d3(0.0, 1.0, 0.0));
        const D3 l2 = dqrot_unit(off, d3(0.0, 0.0, 1.0));
        const DQ r0 = dq_axis_angle(r.a0, q0);
        r.a1 = dqrot_unit(dqnormalize(r0), l1);
        const DQ r10 = dqmul(dq_axis_angle(r.a1, q1), r0);
        r.a2 = dqrot_unit(dqnormalize(r10), l2);
        r.rot = na == 2 ? r10 : dqmul(dq_axis_angle(r.a2, q2), r10);
    }
    return r;
}

}  // namespace sapx

"""CPU oracle for the SAP compliant-contact step — TEST INFRASTRUCTURE ONLY.

This module is the checker, never the product: only ``tests/``,
``__graft_entry__.smoke()`` and ``bench.py``'s CPU-baseline leg may import it.
It is a float64 numpy restatement of the reference algorithm, written from the
reference source read as text, stage by stage:

* boundary conversions   — ``sim/sap_helpers.py:1482-1522, 1761-1848, 2058-2072``
* free motion            — ``sim/free_motion.py:154-220, 412-530, 680-1179, 1970-2178``
                           with helpers ``sim/sap_helpers.py:691-1433``
* contact Jacobian       — ``sim/contact_jacobian.py:153-251, 404-562``
                           with ``sim/sap_helpers.py:148-237, 307-338, 439-503``
* SAP solve              — ``sim/contact_solve.py:316-921, 1435-1692, 1970-2087,
                           2234-2420, 2670-3176, 4285-4429, 5108-5256, 5474-5908,
                           6129-6285``, ``sim/sap_helpers.py:2251-2448, 2603-2611``
* integration + public FK — ``sim/solver_sap.py:158-733, 1267-1386``,
                           ``sim/sap_kinematics.py:84-229``

Parity pinning (what fixes each function independently of this restatement)
is listed in ``tests/test_oracle_*.py``: closed forms, finite differences of
the Lagrangian, KKT/optimality of the convex SAP objective against an
independent minimiser, and invariants.  Functions with no independent check
say "parity unpinned" in their docstring.

Everything operates on one world at a time; the model is the global
``SapModel`` (numpy arrays), worlds are equally shaped and world-major.
"""
from __future__ import annotations

import math
from dataclasses import dataclass, field, replace

import numpy as np

JOINT_PRISMATIC = 0
JOINT_REVOLUTE = 1
JOINT_BALL = 2
JOINT_FIXED = 3
JOINT_FREE = 4
JOINT_DISTANCE = 5
JOINT_D6 = 6

MODE_NONE = 0
MODE_STICTION = 1
MODE_SLIDING = 2
MODE_FRICTIONLESS = 3

_PI = 3.141592653589793
SOFT_NORM_TOL = 1.0e-7          # contact_solve.py:47
PD_BETA = 0.1                   # contact_solve.py:48
LIMIT_BETA = 0.1                # contact_solve.py:49
LIMIT_WINDOW_FACTOR = 2.0       # contact_solve.py:51
EXACT_ALPHA_MAX = 1.5           # contact_solve.py:52
EXACT_MAX_ITERATIONS = 100      # contact_solve.py:53
EXACT_F_TOLERANCE = 1.0e-8      # contact_solve.py:54


# =============================================================================
# quaternion / transform / spatial helpers (xyzw, angular-first spatial)
# =============================================================================
def qmul(a, b):
    """sap_helpers.py:583-590."""
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def qnormalize(q):
    """sap_helpers.py:598-604."""
    n = math.sqrt(float(q[0] * q[0] + q[1] * q[1] + q[2] * q[2] + q[3] * q[3]))
    if n <= 0.0:
        return np.array([0.0, 0.0, 0.0, 1.0])
    return np.asarray(q, dtype=np.float64) / n


def qconj(q):
    return np.array([-q[0], -q[1], -q[2], q[3]])


def qrot_unit(q, v):
    """``_quatd_rotate_unit`` sap_helpers.py:663-667."""
    qv = np.asarray(q[:3], dtype=np.float64)
    t = 2.0 * np.cross(qv, v)
    return np.asarray(v, dtype=np.float64) + q[3] * t + np.cross(qv, t)


def qrot(q, v):
    """Normalising rotate (``_quat_rotate_vec3d`` sap_helpers.py:240-252)."""
    return qrot_unit(qnormalize(q), v)


def q_from_axis_angle(axis, angle):
    """sap_helpers.py:575-580 (axis used as given)."""
    s = math.sin(0.5 * angle)
    c = math.cos(0.5 * angle)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, c])


def tcompose(a, b):
    """``_transformd_compose`` (normalising) sap_helpers.py:670-676."""
    aq = qnormalize(a[3:])
    bq = qnormalize(b[3:])
    return np.concatenate([a[:3] + qrot_unit(aq, b[:3]), qnormalize(qmul(aq, bq))])


def tinverse(x):
    """sap_helpers.py:679-683."""
    qi = qconj(qnormalize(x[3:]))
    return np.concatenate([-qrot_unit(qi, x[:3]), qi])


def tpoint(x, p):
    """``_transform_point_d`` sap_helpers.py:264-266."""
    return x[:3] + qrot(x[3:], p)


def rotation_matrix(q):
    q = qnormalize(q)
    return np.column_stack([qrot_unit(q, e) for e in np.eye(3)])


def skew(c):
    return np.array([[0.0, -c[2], c[1]], [c[2], 0.0, -c[0]], [-c[1], c[0], 0.0]])


def sp(w, v):
    return np.concatenate([np.asarray(w, dtype=np.float64), np.asarray(v, dtype=np.float64)])


def shift_velocity(V, p):
    """sap_helpers.py:948-952."""
    return sp(V[:3], V[3:] + np.cross(V[:3], p))


def shift_acceleration(A, p, w):
    """sap_helpers.py:955-959."""
    return sp(A[:3], A[3:] + np.cross(A[:3], p) + np.cross(w, np.cross(w, p)))


def shift_force(F, p):
    """sap_helpers.py:962-966."""
    return sp(F[:3] - np.cross(p, F[3:]), F[3:])


def compose_velocity(V_WP, p_PB, V_PB):
    """sap_helpers.py:969-979."""
    return sp(V_WP[:3] + V_PB[:3], V_WP[3:] + np.cross(V_WP[:3], p_PB) + V_PB[3:])


def compose_acceleration(A_WP, V_WP, p_PB, V_PB, A_PB):
    """sap_helpers.py:982-994."""
    w = V_WP[:3]
    As = shift_acceleration(A_WP, p_PB, w)
    alpha = As[:3] + A_PB[:3] + np.cross(w, V_PB[:3])
    a = As[3:] + A_PB[3:] + 2.0 * np.cross(w, V_PB[3:])
    return sp(alpha, a)


def mobilizer_to_body(X_WF, p_MoBo_F, V_FM_F):
    """``_sap_transform_mobilizer_velocity_to_body`` sap_helpers.py:997-1009."""
    R = X_WF[3:]
    w = V_FM_F[:3]
    return sp(qrot_unit(R, w), qrot_unit(R, V_FM_F[3:] + np.cross(w, p_MoBo_F)))


def mobilizer_bias_to_body(X_WF, p_MoBo_F, V_FM_F, A_FM_F):
    """sap_helpers.py:1012-1024."""
    R = X_WF[3:]
    s = shift_acceleration(A_FM_F, p_MoBo_F, V_FM_F[:3])
    return sp(qrot_unit(R, s[:3]), qrot_unit(R, s[3:]))


def spatial_inertia_origin(X, I_com, m, com):
    """``_sap_spatial_inertia_body_origin_world`` sap_helpers.py:1027-1062."""
    R = rotation_matrix(X[3:]) if False else np.column_stack([qrot_unit(X[3:], e) for e in np.eye(3)])
    I_W = R @ I_com @ R.T
    c = qrot_unit(X[3:], com)
    I_O = I_W + m * (np.dot(c, c) * np.eye(3) - np.outer(c, c))
    mcx = m * skew(c)
    M = np.zeros((6, 6))
    M[:3, :3] = I_O
    M[:3, 3:] = mcx
    M[3:, :3] = -mcx
    M[3:, 3:] = m * np.eye(3)
    return M


def bias_force_origin(X, I_com, m, com, V):
    """``_sap_dynamic_bias_force_body_origin_world`` sap_helpers.py:1065-1093."""
    R = np.column_stack([qrot_unit(X[3:], e) for e in np.eye(3)])
    I_W = R @ I_com @ R.T
    c = qrot_unit(X[3:], com)
    I_O = I_W + m * (np.dot(c, c) * np.eye(3) - np.outer(c, c))
    w = V[:3]
    return sp(np.cross(w, I_O @ w), m * np.cross(w, np.cross(w, c)))


def gravity_force_origin(X, m, com, g):
    """sap_helpers.py:1096-1105."""
    c = qrot_unit(X[3:], com)
    f = m * np.asarray(g, dtype=np.float64)
    return sp(np.cross(c, f), f)


# =============================================================================
# world view
# =============================================================================
@dataclass
class WorldView:
    """Index ranges of world ``w`` inside the global model arrays."""

    model: object
    w: int
    j0: int = 0
    j1: int = 0
    d0: int = 0
    d1: int = 0
    c0: int = 0
    c1: int = 0
    b0: int = 0
    b1: int = 0

    def __post_init__(self):
        m = self.model
        jpw = m.joint_count // m.world_count
        self.j0, self.j1 = self.w * jpw, (self.w + 1) * jpw
        self.d0, self.d1 = int(m.joint_qd_start[self.j0]), int(m.joint_qd_start[self.j1])
        self.c0, self.c1 = int(m.joint_q_start[self.j0]), int(m.joint_q_start[self.j1])
        bpw = m.body_count // m.world_count
        self.b0, self.b1 = self.w * bpw, (self.w + 1) * bpw

    @property
    def n(self):
        return self.d1 - self.d0

    @property
    def nb(self):
        return self.b1 - self.b0


def _is_identity_tf(x) -> bool:
    return bool(np.all(np.asarray(x[:6]) == 0.0) and x[6] == 1.0)


# =============================================================================
# boundary conversions (sap_helpers.py)
# =============================================================================
def public_to_sap_velocity(model, wv: WorldView, body_q, joint_qd):
    """``copy_public_to_sap_velocity_f32`` sap_helpers.py:1482-1522.

    ``body_q`` is the state's world body pose array (global), ``joint_qd`` the
    global public velocity array.  Returns the world's SAP-order velocity (f64).
    """
    v = np.asarray(joint_qd[wv.d0:wv.d1], dtype=np.float64).copy()
    for j in range(wv.j0, wv.j1):
        jt = int(model.joint_type[j])
        if jt in (JOINT_FREE, JOINT_DISTANCE):
            child = int(model.joint_child[j])
            ds = int(model.joint_qd_start[j]) - wv.d0
            x = np.asarray(body_q[child], dtype=np.float64)
            r = qrot(x[3:], model.body_com[child].astype(np.float64))
            v_com = np.asarray(joint_qd[wv.d0 + ds:wv.d0 + ds + 3], dtype=np.float64)
            w = np.asarray(joint_qd[wv.d0 + ds + 3:wv.d0 + ds + 6], dtype=np.float64)
            v[ds:ds + 3] = w
            v[ds + 3:ds + 6] = v_com - np.cross(w, r)
    return v


def public_to_sap_force(model, wv: WorldView, body_q, joint_f):
    """``_copy_public_to_sap_force_kernel`` sap_helpers.py:1811-1848."""
    f = np.asarray(joint_f[wv.d0:wv.d1], dtype=np.float64).copy()
    for j in range(wv.j0, wv.j1):
        jt = int(model.joint_type[j])
        if jt in (JOINT_FREE, JOINT_DISTANCE):
            child = int(model.joint_child[j])
            ds = int(model.joint_qd_start[j]) - wv.d0
            x = np.asarray(body_q[child], dtype=np.float64)
            r = qrot(x[3:], model.body_com[child].astype(np.float64))
            f_com = np.asarray(joint_f[wv.d0 + ds:wv.d0 + ds + 3], dtype=np.float64)
            t_com = np.asarray(joint_f[wv.d0 + ds + 3:wv.d0 + ds + 6], dtype=np.float64)
            f[ds:ds + 3] = t_com + np.cross(r, f_com)
            f[ds + 3:ds + 6] = f_com
    return f


def public_body_force_to_sap(model, wv: WorldView, body_q, body_f):
    """``_copy_public_body_force_to_sap_body_origin_kernel`` sap_helpers.py:2058-2072 (negated)."""
    out = np.zeros((wv.nb, 6))
    if body_f is None:
        return out
    for lb in range(wv.nb):
        b = wv.b0 + lb
        f_com = np.asarray(body_f[b][:3], dtype=np.float64)
        t_com = np.asarray(body_f[b][3:], dtype=np.float64)
        x = np.asarray(body_q[b], dtype=np.float64)
        c = qrot(x[3:], model.body_com[b].astype(np.float64))
        out[lb] = -sp(t_com + np.cross(c, f_com), f_com)
    return out


# =============================================================================
# free motion (free_motion.py)
# =============================================================================
@dataclass
class FreeMotion:
    body_q: np.ndarray      # (nb, 7) world poses (FK inside free motion)
    S: np.ndarray           # (n, 6) motion subspace, body origin, world frame
    V: np.ndarray           # (nb, 6)
    A: np.ndarray           # (nb, 6) bias accelerations
    H: np.ndarray           # (n, n) mass matrix without armature
    A_env: np.ndarray       # (n, n) H + diag(armature)
    tau: np.ndarray         # (n,)
    qdd: np.ndarray
    v_star: np.ndarray
    body_jac: np.ndarray    # (nb, 6, n) body-origin Jacobian per body


def quat_from_cols(c0, c1, c2):
    """Quaternion of the matrix with columns c0, c1, c2, normalised (``_quatd_from_matrix_cols``
    sap_helpers.py:607-652: the trace branch, else the largest diagonal entry's)."""
    r00, r10, r20 = c0
    r01, r11, r21 = c1
    r02, r12, r22 = c2
    tr = r00 + r11 + r22
    if tr > 0.0:
        s = math.sqrt(tr + 1.0) * 2.0
        q = [(r21 - r12) / s, (r02 - r20) / s, (r10 - r01) / s, 0.25 * s]
    elif r00 > r11 and r00 > r22:
        s = math.sqrt(1.0 + r00 - r11 - r22) * 2.0
        q = [0.25 * s, (r01 + r10) / s, (r02 + r20) / s, (r21 - r12) / s]
    elif r11 > r22:
        s = math.sqrt(1.0 + r11 - r00 - r22) * 2.0
        q = [(r01 + r10) / s, 0.25 * s, (r12 + r21) / s, (r02 - r20) / s]
    else:
        s = math.sqrt(1.0 + r22 - r00 - r11) * 2.0
        q = [(r02 + r20) / s, (r12 + r21) / s, 0.25 * s, (r10 - r01) / s]
    return qnormalize(np.array(q, dtype=np.float64))


def d6_rotation(axes, q):
    """Rotation of a D6 joint's angular coordinates and the current axis of each angular dof
    (``_sap_compute_2d/3d_rotational_dofs`` sap_helpers.py:1109-1161, ``_d6_angular_axis``
    1247-1286): one axis rotates by q about it; two or three are first carried onto the frame
    whose columns they form (the third column of two is their cross product), then rotated in
    turn, each axis moved by the rotations before it (intrinsic sequence)."""
    m = len(axes)
    if m == 0:
        return np.array([0.0, 0.0, 0.0, 1.0]), []
    if m == 1:
        a0 = np.asarray(axes[0], dtype=np.float64)
        return q_from_axis_angle(a0, q[0]), [a0]
    a0, a1 = np.asarray(axes[0], dtype=np.float64), np.asarray(axes[1], dtype=np.float64)
    a2 = np.cross(a0, a1) if m == 2 else np.asarray(axes[2], dtype=np.float64)
    q_off = quat_from_cols(a0, a1, a2)
    l0 = qrot(q_off, np.array([1.0, 0.0, 0.0]))
    l1 = qrot(q_off, np.array([0.0, 1.0, 0.0]))
    q0 = q_from_axis_angle(l0, q[0])
    ax1 = qrot(q0, l1)
    q1 = q_from_axis_angle(ax1, q[1])
    q10 = qmul(q1, q0)
    if m == 2:
        return q10, [l0, ax1]
    l2 = qrot(q_off, np.array([0.0, 0.0, 1.0]))
    ax2 = qrot(q10, l2)
    q2 = q_from_axis_angle(ax2, q[2])
    return qmul(q2, q10), [l0, ax1, ax2]


def _mobilizer_transform(jt, axes, q, nlin=0):
    """``_sap_calc_across_mobilizer_transform`` sap_helpers.py:1164-1243 (``nlin``: a D6 joint's
    linear axis count; its axes list is linear then angular)."""
    if jt == JOINT_D6:
        pos = np.zeros(3)
        for k in range(nlin):
            pos = pos + np.asarray(axes[k], dtype=np.float64) * q[k]
        rot, _ = d6_rotation(axes[nlin:], q[nlin:])
        return np.concatenate([pos, rot])
    if jt == JOINT_PRISMATIC:
        return np.concatenate([axes[0] * q[0], [0.0, 0.0, 0.0, 1.0]])
    if jt == JOINT_REVOLUTE:
        return np.concatenate([np.zeros(3), q_from_axis_angle(axes[0], q[0])])
    if jt == JOINT_BALL:
        return np.concatenate([np.zeros(3), q[:4]])
    if jt in (JOINT_FREE, JOINT_DISTANCE):
        return np.asarray(q[:7], dtype=np.float64).copy()
    return np.array([0, 0, 0, 0, 0, 0, 1.0])


def free_motion(model, wv: WorldView, joint_q, v0, joint_f_sap, body_f_sap, dt) -> FreeMotion:
    """SAP free motion ``v* = v0 + dt (H + diag(armature))^-1 tau`` (free_motion.py:2094-2178).

    ``joint_q`` are the world's coordinates (f64), ``v0`` / ``joint_f_sap`` SAP
    order, ``body_f_sap`` the SAP body-origin external wrench (already negated).
    Kinematics follow ``_eval_rigid_id_tiled_articulations`` (free_motion.py:412-530),
    dynamics ``_eval_rigid_body_dynamics_parallel`` (680-719), torque recursion
    ``_eval_rigid_tau_no_drives`` (93-147), Jacobian ``_eval_rigid_jacobian_parallel``
    (722-765) and ``H = J^T M J`` (843-1179).
    """
    n, nb = wv.n, wv.nb
    body_q = np.zeros((nb, 7))
    body_q[:, 6] = 1.0
    V = np.zeros((nb, 6))
    Aacc = np.zeros((nb, 6))
    S = np.zeros((n, 6))
    g = model.gravity[wv.w].astype(np.float64)

    def lb(b):
        return -1 if b < 0 else b - wv.b0

    for j in range(wv.j0, wv.j1):
        jt = int(model.joint_type[j])
        parent = lb(int(model.joint_parent[j]))
        child = lb(int(model.joint_child[j]))
        qs = int(model.joint_q_start[j]) - wv.c0
        ds = int(model.joint_qd_start[j]) - wv.d0
        nd = int(model.joint_qd_start[j + 1]) - int(model.joint_qd_start[j])
        axes = [model.joint_axis[wv.d0 + ds + k].astype(np.float64) for k in range(nd)]
        X_p = model.joint_X_p[j].astype(np.float64)
        X_c = model.joint_X_c[j].astype(np.float64)
        X_c_ident = _is_identity_tf(X_c)
        X_wpj = X_p if parent < 0 else tcompose(body_q[parent], X_p)
        qd = v0[ds:ds + nd]
        if jt == JOINT_REVOLUTE and X_c_ident:
            # _sap_revolute_identity_child_kinematics (sap_helpers.py:704-720)
            q_FM = q_from_axis_angle(axes[0], joint_q[qs])
            R_WF = X_wpj[3:]
            X_wc = np.concatenate([X_wpj[:3], qmul(R_WF, q_FM)])
            S[ds] = sp(qrot_unit(R_WF, axes[0]), np.zeros(3))
            v_PB = S[ds] * qd[0]
            A_PB = np.zeros(6)
        else:
            nlin = int(model.joint_dof_dim[j, 0]) if jt == JOINT_D6 else 0
            X_j = _mobilizer_transform(jt, axes, joint_q[qs:], nlin)
            X_WM = tcompose(X_wpj, X_j)
            X_wc = X_WM if X_c_ident else tcompose(X_WM, tinverse(X_c))
            # _sap_calc_mobilizer_velocity_and_bias_body_world (sap_helpers.py:1289-1433)
            X_MB = tinverse(X_c)
            p_MoBo_F = qrot_unit(X_j[3:], X_MB[:3])
            if jt == JOINT_PRISMATIC:
                S[ds] = mobilizer_to_body(X_wpj, p_MoBo_F, sp(np.zeros(3), axes[0]))
                V_FM = sp(np.zeros(3), axes[0] * qd[0])
                A_PB = mobilizer_bias_to_body(X_wpj, p_MoBo_F, V_FM, np.zeros(6))
                v_PB = S[ds] * qd[0]
            elif jt == JOINT_REVOLUTE:
                R_WF = X_wpj[3:]
                S[ds] = sp(qrot_unit(R_WF, axes[0]), qrot_unit(R_WF, np.cross(axes[0], p_MoBo_F)))
                wFM = axes[0] * qd[0]
                A_PB = sp(np.zeros(3), qrot_unit(R_WF, np.cross(wFM, np.cross(wFM, p_MoBo_F))))
                v_PB = S[ds] * qd[0]
            elif jt == JOINT_BALL:
                # three angular mobilities in the joint frame (sap_helpers.py:1345-1361)
                V_FM = sp(qd[:3], np.zeros(3))
                v_PB = mobilizer_to_body(X_wpj, p_MoBo_F, V_FM)
                for k in range(3):
                    e = np.zeros(3)
                    e[k] = 1.0
                    S[ds + k] = mobilizer_to_body(X_wpj, p_MoBo_F, sp(e, np.zeros(3)))
                A_PB = mobilizer_bias_to_body(X_wpj, p_MoBo_F, V_FM, np.zeros(6))
            elif jt in (JOINT_FREE, JOINT_DISTANCE):
                V_FM = sp(qd[:3], qd[3:6])
                v_PB = mobilizer_to_body(X_wpj, p_MoBo_F, V_FM)
                for k in range(3):
                    e = np.zeros(3)
                    e[k] = 1.0
                    S[ds + k] = mobilizer_to_body(X_wpj, p_MoBo_F, sp(e, np.zeros(3)))
                    S[ds + 3 + k] = mobilizer_to_body(X_wpj, p_MoBo_F, sp(np.zeros(3), e))
                A_PB = mobilizer_bias_to_body(X_wpj, p_MoBo_F, V_FM, np.zeros(6))
            elif jt == JOINT_D6:
                # linear axes, then the angular axes at the current coordinates (sap_helpers.py:
                # 1387-1432); the bias is the body offset's velocity product only
                _, ang = d6_rotation(axes[nlin:], joint_q[qs + nlin:])
                V_FM = np.zeros(6)
                for k in range(nd):
                    col = sp(np.zeros(3), axes[k]) if k < nlin else sp(ang[k - nlin], np.zeros(3))
                    S[ds + k] = mobilizer_to_body(X_wpj, p_MoBo_F, col)
                    V_FM = V_FM + col * qd[k]
                v_PB = mobilizer_to_body(X_wpj, p_MoBo_F, V_FM)
                A_PB = mobilizer_bias_to_body(X_wpj, p_MoBo_F, V_FM, np.zeros(6))
            elif jt == JOINT_FIXED:
                v_PB = np.zeros(6)
                A_PB = np.zeros(6)
            else:
                raise NotImplementedError(f"joint type {jt} not in the oracle")
        body_q[child] = X_wc
        if parent >= 0:
            p_PB = X_wc[:3] - body_q[parent][:3]
            V[child] = compose_velocity(V[parent], p_PB, v_PB)
            Aacc[child] = compose_acceleration(Aacc[parent], V[parent], p_PB, v_PB, A_PB)
        else:
            V[child] = v_PB
            Aacc[child] = A_PB

    # body dynamics (free_motion.py:680-719)
    I_s = np.zeros((nb, 6, 6))
    f_b = np.zeros((nb, 6))
    for b in range(nb):
        gb = wv.b0 + b
        m = float(model.body_mass[gb])
        I_com = model.body_inertia[gb].astype(np.float64)
        com = model.body_com[gb].astype(np.float64)
        I_s[b] = spatial_inertia_origin(body_q[b], I_com, m, com)
        f_b[b] = I_s[b] @ Aacc[b] + bias_force_origin(body_q[b], I_com, m, com, V[b]) \
            - gravity_force_origin(body_q[b], m, com, g)

    # tau, reverse recursion (free_motion.py:93-147)
    tau = np.zeros(n)
    f_t = np.zeros((nb, 6))
    for j in reversed(range(wv.j0, wv.j1)):
        parent = lb(int(model.joint_parent[j]))
        child = lb(int(model.joint_child[j]))
        ds = int(model.joint_qd_start[j]) - wv.d0
        nd = int(model.joint_qd_start[j + 1]) - int(model.joint_qd_start[j])
        f_s = f_b[child] + f_t[child] + body_f_sap[child]
        for k in range(nd):
            tau[ds + k] = -np.dot(S[ds + k], f_s) + joint_f_sap[ds + k]
        if parent >= 0:
            f_t[parent] += shift_force(f_s, body_q[parent][:3] - body_q[child][:3])

    # body-origin Jacobian per body and H = J^T M J (free_motion.py:722-1179)
    joint_of_body = {}
    for j in range(wv.j0, wv.j1):
        joint_of_body[lb(int(model.joint_child[j]))] = j
    body_jac = np.zeros((nb, 6, n))
    for b in range(nb):
        if b not in joint_of_body:
            continue
        p_row = body_q[b][:3]
        j = joint_of_body[b]
        while j != -1:
            col_body = lb(int(model.joint_child[j]))
            p_col = body_q[col_body][:3]
            ds = int(model.joint_qd_start[j]) - wv.d0
            nd = int(model.joint_qd_start[j + 1]) - int(model.joint_qd_start[j])
            for k in range(nd):
                body_jac[b, :, ds + k] = shift_velocity(S[ds + k], p_row - p_col)
            j = int(model.joint_ancestor[j])
    H = np.zeros((n, n))
    for b in range(nb):
        H += body_jac[b].T @ I_s[b] @ body_jac[b]
    H = 0.5 * (H + H.T)
    arm = model.joint_armature[wv.d0:wv.d1].astype(np.float64)
    A_env = H + np.diag(arm)
    qdd = np.linalg.solve(A_env, tau)
    v_star = v0 + dt * qdd
    return FreeMotion(body_q=body_q, S=S, V=V, A=Aacc, H=H, A_env=A_env, tau=tau, qdd=qdd, v_star=v_star,
                      body_jac=body_jac)


# =============================================================================
# contacts (contact_jacobian.py, sap_helpers.py)
# =============================================================================
def make_frame(n):
    """Drake's MakeFromOneUnitVector, z = n (``_sap_make_from_one_unit_vector_z`` sap_helpers.py:467-503)."""
    ax, ay, az = abs(n[0]), abs(n[1]), abs(n[2])
    i = 0
    mn = ax
    if ay < mn:
        i, mn = 1, ay
    if az < mn:
        i, mn = 2, az
    mag = math.sqrt(1.0 - mn * mn)
    r = 1.0 / mag
    u = n[i]
    s = -r * u
    if i == 0:
        v = np.array([0.0, -r * n[2], r * n[1]])
        w = np.array([mag, s * n[1], s * n[2]])
    elif i == 1:
        v = np.array([r * n[2], 0.0, -r * n[0]])
        w = np.array([s * n[0], mag, s * n[2]])
    else:
        v = np.array([-r * n[1], r * n[0], 0.0])
        w = np.array([s * n[0], s * n[1], mag])
    return v, w, np.asarray(n, dtype=np.float64)


def safe_normalize(v):
    """sap_helpers.py:90-95."""
    v = np.asarray(v, dtype=np.float64)
    n2 = float(np.dot(v, v))
    if n2 <= 1.0e-32:
        return np.zeros(3)
    return v / math.sqrt(n2)


def _shape_stiffness(model, shape, fallback_k):
    """sap_helpers.py:182-195."""
    val = fallback_k
    if shape >= 0:
        val = float(model.shape_material_ke[shape])
    if not math.isfinite(val) or val <= 0.0:
        val = fallback_k
    if not math.isfinite(val) or val <= 0.0:
        val = 1.0
    return val


def _combine_stiffness(k0, k1):
    """sap_helpers.py:198-207."""
    if not math.isfinite(k0):
        return k1
    if not math.isfinite(k1):
        return k0
    d = k0 + k1
    if d <= 1.0e-12:
        return max(k0, k1)
    return k0 * k1 / d


def _shape_mu(model, shape, fallback_mu):
    """sap_helpers.py:148-159."""
    val = fallback_mu
    if shape >= 0:
        val = float(model.shape_material_mu[shape])
    return max(val, 0.0)


def _combine_mu(m0, m1):
    """sap_helpers.py:162-167."""
    d = m0 + m1
    if d <= 1.0e-12:
        return 0.0
    return 2.0 * m0 * m1 / d


def _shape_tau(model, shape, fallback_tau):
    """sap_helpers.py:210-223."""
    val = fallback_tau
    if shape >= 0:
        val = float(model.shape_material_tau[shape])
    if not math.isfinite(val) or val < 0.0:
        val = fallback_tau
    if not math.isfinite(val) or val < 0.0:
        val = 0.0
    return val


def _direction_weight(model, gb, body_pose, r_world, direction):
    """``_sap_body_direction_weight_d`` sap_helpers.py:307-338."""
    m = float(model.body_mass[gb])
    if m <= 1.0e-12:
        return 0.0
    c_world = np.cross(r_world, direction)
    q = body_pose[3:]
    n2 = float(np.dot(q, q))
    qinv = np.array([-q[0], -q[1], -q[2], q[3]]) / n2
    c_body = qrot(qinv, c_world)
    I = model.body_inertia[gb].astype(np.float64)
    ix, iy, iz = (max(float(I[k, k]), 1.0e-12) for k in range(3))
    return 1.0 / m + c_body[0] ** 2 / ix + c_body[1] ** 2 / iy + c_body[2] ** 2 / iz


@dataclass
class ContactSet:
    """Per-world contact slots (``SapContactJacobianResult`` fields, contact_jacobian.py:51-84)."""

    body0: np.ndarray
    body1: np.ndarray
    x0: np.ndarray
    x1: np.ndarray
    point: np.ndarray
    R: np.ndarray        # (nc, 3, 3) columns = cx, cy, n
    phi0: np.ndarray
    mu: np.ndarray
    k: np.ndarray
    tau: np.ndarray
    w: np.ndarray
    J: np.ndarray        # (nc, 3, n)
    flat_index: np.ndarray

    @property
    def count(self):
        return len(self.phi0)


def world_contacts(model, wv: WorldView, fm: FreeMotion, contacts: dict, *, max_contacts: int,
                   contact_point_mode: str = "witness_point", contact_weight_mode: str = "body_inertia",
                   fallback_mu: float = 0.1, fallback_k: float = 1.0e10, fallback_tau: float = 0.0) -> ContactSet:
    """Scatter the world's flat contacts into slots, in flat order (contact_jacobian.py:404-562).

    ``contacts`` holds numpy copies of the flat rigid-contact arrays; slot order
    is the flat index order (the reference's atomic slot order is not
    deterministic, SAP results do not depend on it).  Contacts beyond
    ``max_contacts`` are dropped like the reference (``contact_jacobian.py:470-472``).
    """
    count = int(contacts["count"])
    sh0_all = contacts["shape0"]
    sh1_all = contacts["shape1"]
    bpe = model.bodies_per_env
    n = wv.n
    rows = []
    for i in range(count):
        sh0, sh1 = int(sh0_all[i]), int(sh1_all[i])
        b0 = int(model.shape_body[sh0]) if sh0 >= 0 else -1
        b1 = int(model.shape_body[sh1]) if sh1 >= 0 else -1
        env = b0 // bpe if b0 >= 0 else (b1 // bpe if b1 >= 0 else -1)
        if env != wv.w:
            continue
        if b0 >= 0 and b0 // bpe != env:
            continue
        if b1 >= 0 and b1 // bpe != env:
            continue
        if len(rows) >= max_contacts:
            continue
        rows.append(i)
    nc = len(rows)
    cs = ContactSet(body0=np.zeros(nc, int), body1=np.zeros(nc, int), x0=np.zeros((nc, 3)), x1=np.zeros((nc, 3)),
                    point=np.zeros((nc, 3)), R=np.zeros((nc, 3, 3)), phi0=np.zeros(nc), mu=np.zeros(nc),
                    k=np.zeros(nc), tau=np.zeros(nc), w=np.zeros(nc), J=np.zeros((nc, 3, n)),
                    flat_index=np.asarray(rows, dtype=int))
    use_witness = contact_point_mode == "witness_point"
    for s, i in enumerate(rows):
        sh0, sh1 = int(sh0_all[i]), int(sh1_all[i])
        b0 = int(model.shape_body[sh0]) if sh0 >= 0 else -1
        b1 = int(model.shape_body[sh1]) if sh1 >= 0 else -1
        x0 = np.asarray(contacts["point0"][i], dtype=np.float64)
        x1 = np.asarray(contacts["point1"][i], dtype=np.float64)
        p0 = np.zeros(3)
        p1 = np.zeros(3)
        if b0 >= 0:
            pose = fm.body_q[b0 - wv.b0]
            p0 = pose[:3]
            x0 = tpoint(pose, x0)
        if b1 >= 0:
            pose = fm.body_q[b1 - wv.b0]
            p1 = pose[:3]
            x1 = tpoint(pose, x1)
        nrm = safe_normalize(contacts["normal"][i])
        cx, cy, cz = make_frame(nrm)
        k0 = _shape_stiffness(model, sh0, fallback_k)
        k1 = _shape_stiffness(model, sh1, fallback_k)
        den = k0 + k1
        w0, w1 = (k0 / den, k1 / den) if den != 0.0 else (0.5, 0.5)
        pC = w0 * x0 + w1 * x1
        cs.body0[s], cs.body1[s] = b0, b1
        cs.x0[s], cs.x1[s], cs.point[s] = x0, x1, pC
        cs.phi0[s] = np.dot(x1 - x0, cz) - float(contacts["margin0"][i]) - float(contacts["margin1"][i])
        cs.mu[s] = _combine_mu(_shape_mu(model, sh0, fallback_mu), _shape_mu(model, sh1, fallback_mu))
        cs.k[s] = _combine_stiffness(k0, k1)
        cs.tau[s] = _shape_tau(model, sh0, fallback_tau) + _shape_tau(model, sh1, fallback_tau)
        cs.R[s] = np.column_stack([cx, cy, cz])
        xj0 = x0 if use_witness else pC
        xj1 = x1 if use_witness else pC
        rel = np.zeros((3, n))
        if b0 >= 0:
            Jb = fm.body_jac[b0 - wv.b0]
            rel -= Jb[3:] + np.cross(Jb[:3].T, xj0 - p0).T
        if b1 >= 0:
            Jb = fm.body_jac[b1 - wv.b0]
            rel += Jb[3:] + np.cross(Jb[:3].T, xj1 - p1).T
        cs.J[s] = np.stack([cx @ rel, cy @ rel, cz @ rel])
    # effective weights
    if contact_weight_mode == "body_inertia":
        for s in range(nc):
            cx = safe_normalize(cs.R[s][:, 0])
            cy = safe_normalize(cs.R[s][:, 1])
            cz = safe_normalize(cs.R[s][:, 2])
            wt1 = wt2 = wn = 0.0
            for b, x in ((cs.body0[s], cs.x0[s]), (cs.body1[s], cs.x1[s])):
                if b < 0:
                    continue
                pose = fm.body_q[b - wv.b0]
                com_w = pose[:3] + qrot(pose[3:], model.body_com[b].astype(np.float64))
                r = x - com_w
                wt1 += _direction_weight(model, b, pose, r, cx)
                wt2 += _direction_weight(model, b, pose, r, cy)
                wn += _direction_weight(model, b, pose, r, cz)
            cs.w[s] = max((wt1 + wt2 + wn) / 3.0, 1.0e-12)
    elif contact_weight_mode == "diag_delassus":
        a_inv = 1.0 / np.maximum(np.diag(fm.A_env), 1.0e-12)
        for s in range(nc):
            W = (cs.J[s] * a_inv) @ cs.J[s].T
            cs.w[s] = max(math.sqrt(float(np.sum(W * W))) / 3.0, 1.0e-12)
    else:
        raise ValueError(contact_weight_mode)
    return cs


# =============================================================================
# SAP projection (contact_solve.py:743-921)
# =============================================================================
def contact_constants(cs: ContactSet, dt, beta=1.0, sigma=1.0e-3):
    """Per-contact regularisation (rt, rn, vhat_n, mu) — contact_solve.py:791-822."""
    nc = cs.count
    rt = np.zeros(nc)
    rn = np.zeros(nc)
    vhat = np.zeros(nc)
    mu = np.zeros(nc)
    bf = beta * beta / (4.0 * _PI * _PI)
    for c in range(nc):
        wi = cs.w[c]
        if wi < 1.0e-12 or not math.isfinite(wi):
            wi = 1.0e-12
        rn_hard = bf * wi
        kc = cs.k[c]
        if kc <= 0.0 or not math.isfinite(kc):
            kc = 1.0
        tau_c = max(cs.tau[c], 0.0)
        rn_soft = 1.0 / (dt * kc * (dt + tau_c))
        rnc = max(rn_hard, rn_soft)
        rtc = sigma * wi
        rt[c] = max(rtc, 1.0e-30)
        rn[c] = max(rnc, 1.0e-30)
        vhat[c] = -cs.phi0[c] / (dt + tau_c)
        m = cs.mu[c]
        mu[c] = 0.0 if (m < 0.0 or not math.isfinite(m)) else m
    return rt, rn, vhat, mu


def project(vc, rt, rn, vhat, mu, need_G=True):
    """Projection onto the regularised friction cone; returns gamma, G, cost, mode."""
    rt_inv = 1.0 / rt
    rn_inv = 1.0 / rn
    y = np.array([-rt_inv * vc[0], -rt_inv * vc[1], rn_inv * (vhat - vc[2])])
    yr = math.sqrt(y[0] * y[0] + y[1] * y[1] + SOFT_NORM_TOL * SOFT_NORM_TOL)
    t = np.array([y[0] / yr, y[1] / yr, 0.0]) if yr > 0.0 else np.zeros(3)
    gamma = np.zeros(3)
    G = np.zeros((3, 3))
    mode = MODE_NONE
    if mu <= 1.0e-12:
        if y[2] > 0.0:
            gamma = np.array([0.0, 0.0, y[2]])
            G[2, 2] = rn_inv
            mode = MODE_FRICTIONLESS
    else:
        mu_tilde = mu * math.sqrt(rt / rn)
        mu_hat = mu * rt / rn
        f = 1.0 / (1.0 + mu_tilde * mu_tilde)
        if yr <= mu * y[2]:
            gamma = y.copy()
            G = np.diag([rt_inv, rt_inv, rn_inv])
            mode = MODE_STICTION
        elif (-mu_hat * yr < y[2]) and (y[2] < yr / mu):
            gn = (y[2] + mu_hat * yr) * f
            gamma = np.array([mu * gn * t[0], mu * gn * t[1], gn])
            P = np.outer(t[:2], t[:2])
            Pp = np.eye(2) - P
            gn_over_yr = gn / yr if yr > 0.0 else 0.0
            dgt_dyt = mu * (gn_over_yr * Pp + mu_hat * f * P)
            G = np.zeros((3, 3))
            G[:2, :2] = dgt_dyt * rt_inv
            G[0, 2] = mu * f * t[0] * rn_inv
            G[1, 2] = mu * f * t[1] * rn_inv
            G[2, 0] = mu_hat * f * t[0] * rt_inv
            G[2, 1] = mu_hat * f * t[1] * rt_inv
            G[2, 2] = f * rn_inv
            mode = MODE_SLIDING
    cost = 0.5 * (rt * (gamma[0] ** 2 + gamma[1] ** 2) + rn * gamma[2] ** 2)
    return gamma, G, cost, mode


# =============================================================================
# PD / limit terms (contact_solve.py:490-741, sap_helpers.py:2251-2340)
# =============================================================================
def clamp_scalar(x, lim):
    if lim <= 0.0 or not math.isfinite(lim):
        return x
    return max(-lim, min(lim, x))


def clamp_derivative(x, lim):
    if lim <= 0.0 or not math.isfinite(lim):
        return 1.0
    return 1.0 if (-lim <= x <= lim) else 0.0


def clamp_antiderivative(x, lim):
    if lim <= 0.0 or not math.isfinite(lim):
        return 0.5 * x * x
    if x < -lim:
        return -lim * (x + 0.5 * lim)
    if x <= lim:
        return 0.5 * x * x
    return lim * (x - 0.5 * lim)


def effective_pd_gains(kp, kd, dt, delassus):
    """``_compute_effective_pd_gains_sap`` sap_helpers.py:2305-2340."""
    if kp <= 0.0 and kd <= 0.0:
        return 0.0, 0.0
    bf = PD_BETA * PD_BETA / (4.0 * _PI * _PI)
    r_nr = bf * max(delassus, 1.0e-12)
    common = dt * kp + kd
    if common <= 0.0:
        return 0.0, 0.0
    r = 1.0 / (dt * common)
    if r >= r_nr:
        return kp, kd
    if dt * kp > kd and kp > 0.0:
        tau = kd / kp
        kp_eff = 1.0 / (dt * (dt + tau) * r_nr)
        kd_eff = tau * kp_eff
    elif kd > 0.0:
        tau_inv = kp / kd
        kd_eff = 1.0 / (dt * (1.0 + dt * tau_inv) * r_nr)
        kp_eff = tau_inv * kd_eff
    else:
        kp_eff = 1.0 / (dt * dt * r_nr)
        kd_eff = 0.0
    return kp_eff, kd_eff


def dof_maps(model):
    """``_build_dof_maps`` contact_solve.py:3681-3720 (coord index, target index, limit support)."""
    nd = model.joint_dof_count
    coord = np.full(nd, -1, dtype=int)
    target = np.arange(nd, dtype=int)
    limit_ok = np.zeros(nd, dtype=int)
    for j in range(model.joint_count):
        jt = int(model.joint_type[j])
        ds = int(model.joint_qd_start[j])
        if jt in (JOINT_FREE, JOINT_DISTANCE):
            for a in range(3):
                target[ds + a] = ds + a + 3
                target[ds + a + 3] = ds + a
        if jt not in (JOINT_PRISMATIC, JOINT_REVOLUTE, JOINT_D6):
            continue
        cs0 = int(model.joint_q_start[j])
        ac = int(model.joint_dof_dim[j, 0] + model.joint_dof_dim[j, 1])
        for a in range(ac):
            coord[ds + a] = cs0 + a
            if ac == 1:
                limit_ok[ds + a] = 1
    return coord, target, limit_ok


@dataclass
class ModelTerms:
    pd_active: np.ndarray
    pd_a: np.ndarray
    pd_gain: np.ndarray
    pd_limit: np.ndarray
    lo_active: np.ndarray
    up_active: np.ndarray
    lo_vhat: np.ndarray
    up_vhat: np.ndarray
    lo_r: np.ndarray
    up_r: np.ndarray


def build_model_terms(model, wv: WorldView, joint_q_global, control: dict, A, v0, v_star, dt) -> ModelTerms:
    """PD (contact_solve.py:490-561) and limit (605-683) terms for one world."""
    n = wv.n
    coord, target, limit_ok = dof_maps(model)
    a_inv = 1.0 / np.array([d if (d >= 1.0e-12 and math.isfinite(d)) else 1.0e-12 for d in np.diag(A)])
    mt = ModelTerms(*[np.zeros(n) for _ in range(10)])
    mt.pd_active = np.zeros(n, dtype=int)
    mt.lo_active = np.zeros(n, dtype=int)
    mt.up_active = np.zeros(n, dtype=int)
    mt.lo_r[:] = 1.0
    mt.up_r[:] = 1.0
    for i in range(n):
        dof = wv.d0 + i
        c = coord[dof]
        t = target[dof]
        # PD
        if int(model.joint_target_mode[t]) != 0:
            kp = float(model.joint_target_ke[t])
            kd = float(model.joint_target_kd[t])
            if not (kp <= 0.0 and kd <= 0.0) and not (c < 0 and kp > 0.0):
                kp_e, kd_e = effective_pd_gains(kp, kd, dt, a_inv[i])
                gain = dt * kp_e + kd_e
                if gain > 0.0:
                    q0 = float(joint_q_global[c]) if c >= 0 else 0.0
                    qd = float(control["joint_target_pos"][t]) if c >= 0 else 0.0
                    vd = float(control["joint_target_vel"][t])
                    u0 = float(control["joint_act"][t])
                    mt.pd_active[i] = 1
                    mt.pd_a[i] = kp_e * (qd - q0) + kd_e * vd + u0
                    mt.pd_gain[i] = gain
                    mt.pd_limit[i] = float(model.joint_effort_limit[t])
        # limits
        if c < 0 or limit_ok[dof] == 0:
            continue
        ke = float(model.joint_limit_ke[dof])
        if ke <= 0.0:
            continue
        q0 = float(joint_q_global[c])
        lower = float(model.joint_limit_lower[dof])
        upper = float(model.joint_limit_upper[dof])
        window = LIMIT_WINDOW_FACTOR * dt * max(abs(v0[i]), abs(v_star[i]))
        kd = float(model.joint_limit_kd[dof])
        tau_d = kd / ke if kd > 0.0 else 0.0
        bf = LIMIT_BETA * LIMIT_BETA / (4.0 * _PI * _PI)
        r = max(bf * max(a_inv[i], 1.0e-12), 1.0 / (dt * ke * (dt + tau_d)))
        if math.isfinite(lower):
            g = q0 - lower
            if g <= window:
                mt.lo_active[i] = 1
                mt.lo_vhat[i] = -g / (dt + tau_d)
                mt.lo_r[i] = r
        if math.isfinite(upper):
            g = upper - q0
            if g <= window:
                mt.up_active[i] = 1
                mt.up_vhat[i] = -g / (dt + tau_d)
                mt.up_r[i] = r
    return mt


def eval_model_terms(mt: ModelTerms, v, dt):
    """PD + limit impulse, Hessian diagonal and cost at ``v`` (contact_solve.py:1569-1612)."""
    n = len(v)
    imp = np.zeros(n)
    hdiag = np.zeros(n)
    cost = 0.0
    for i in range(n):
        if mt.pd_active[i]:
            gain = mt.pd_gain[i]
            y = mt.pd_a[i] - gain * v[i]
            imp[i] += dt * clamp_scalar(y, mt.pd_limit[i])
            hdiag[i] += dt * gain * clamp_derivative(y, mt.pd_limit[i])
            if gain > 0.0:
                cost += (dt / gain) * clamp_antiderivative(y, mt.pd_limit[i])
        if mt.lo_active[i]:
            g = (mt.lo_vhat[i] - v[i]) / mt.lo_r[i]
            if g > 0.0:
                imp[i] += g
                hdiag[i] += 1.0 / mt.lo_r[i]
                cost += 0.5 * mt.lo_r[i] * g * g
        if mt.up_active[i]:
            g = (mt.up_vhat[i] + v[i]) / mt.up_r[i]
            if g > 0.0:
                imp[i] -= g
                hdiag[i] += 1.0 / mt.up_r[i]
                cost += 0.5 * mt.up_r[i] * g * g
    return imp, hdiag, cost


# =============================================================================
# SAP objective, Newton and line searches
# =============================================================================
@dataclass
class SolveOptions:
    max_iterations: int = 100
    optimality_abs_tol: float = 1.0e-14
    optimality_rel_tol: float = 1.0e-6
    cost_abs_tol: float = 1.0e-30
    cost_rel_tol: float = 1.0e-15
    line_search: str = "armijo_decay"
    ls_max_iterations: int = 40
    armijo_c: float = 1.0e-4
    rho: float = 0.8
    relative_slop: float = 1000.0 * np.finfo(np.float64).eps
    diag_shift: float = 0.0
    beta: float = 1.0
    sigma: float = 1.0e-3

    @staticmethod
    def for_variant(variant: str, **kw):
        o = SolveOptions(line_search=variant, **kw)
        if variant == "exact_root" and o.ls_max_iterations == 40:
            o.ls_max_iterations = EXACT_MAX_ITERATIONS
        if variant == "monotone_decay":
            # the reference's defaults for this variant (contact_solve.py:6179-6181)
            if "cost_abs_tol" not in kw:
                o.cost_abs_tol = 0.0
            if "cost_rel_tol" not in kw:
                o.cost_rel_tol = 5.0e-3
        return o


@dataclass
class SolveResult:
    v: np.ndarray
    iterations: int
    ls_iterations: int
    converged: bool
    optimality_reached: bool
    cost: float
    gamma: np.ndarray
    grad_norm: float
    opt_tol: float
    status: int = 0
    history: list = field(default_factory=list)


class SapProblem:
    """The SAP objective of one world: ell(v) = 1/2 (v-v*)^T A (v-v*) + sum regularisers."""

    def __init__(self, A, v_star, cs: ContactSet, mt: ModelTerms, dt, beta=1.0, sigma=1.0e-3):
        self.A = np.asarray(A, dtype=np.float64)
        self.v_star = np.asarray(v_star, dtype=np.float64)
        self.cs = cs
        self.mt = mt
        self.dt = dt
        self.rt, self.rn, self.vhat, self.mu = contact_constants(cs, dt, beta, sigma)
        self.Jf = cs.J.reshape(-1, A.shape[0]) if cs.count else np.zeros((0, A.shape[0]))

    def contact_terms(self, vc_all, need_G=True):
        nc = self.cs.count
        gam = np.zeros((nc, 3))
        G = np.zeros((nc, 3, 3))
        cost = 0.0
        for c in range(nc):
            g, Gc, cc, _ = project(vc_all[c], self.rt[c], self.rn[c], self.vhat[c], self.mu[c], need_G)
            gam[c] = g
            G[c] = Gc
            cost += cc
        return gam, G, cost

    def evaluate(self, v):
        r = v - self.v_star
        Ar = self.A @ r
        vc = (self.Jf @ v).reshape(-1, 3)
        gam, G, ccost = self.contact_terms(vc)
        imp_m, hdiag, mcost = eval_model_terms(self.mt, v, self.dt)
        jimp = self.Jf.T @ gam.reshape(-1) + imp_m
        grad = Ar - jimp
        cost = 0.5 * float(r @ Ar) + ccost + mcost
        return dict(cost=cost, grad=grad, jimp=jimp, Av=self.A @ v, gamma=gam, G=G, hdiag=hdiag, vc=vc)

    def hessian(self, ev):
        H = self.A.copy()
        for c in range(self.cs.count):
            Jc = self.cs.J[c]
            H += Jc.T @ ev["G"][c] @ Jc
        H += np.diag(ev["hdiag"])
        return H

    def cost(self, v):
        return self.evaluate(v)["cost"]


def _participating(model, wv: WorldView, cs: ContactSet, mt: ModelTerms):
    """Participating-DOF mask (contact_solve.py:422-487, 3722-3760)."""
    n = wv.n
    part = np.zeros(n, dtype=int)
    # body -> articulation dof range (whole articulation)
    art_range = {}
    for a in range(model.articulation_count):
        js, je = int(model.articulation_start[a]), int(model.articulation_start[a + 1])
        if js >= wv.j1 or je <= wv.j0:
            continue
        starts = [int(model.joint_qd_start[j]) for j in range(js, je)
                  if int(model.joint_dof_dim[j, 0] + model.joint_dof_dim[j, 1]) > 0]
        cnt = sum(int(model.joint_dof_dim[j, 0] + model.joint_dof_dim[j, 1]) for j in range(js, je))
        if not starts or cnt <= 0:
            continue
        for j in range(js, je):
            art_range[int(model.joint_child[j])] = (min(starts) - wv.d0, cnt)
    for c in range(cs.count):
        for b in (cs.body0[c], cs.body1[c]):
            if b >= 0 and b in art_range:
                s, k = art_range[b]
                part[s:s + k] = 1
        part[np.any(cs.J[c] != 0.0, axis=0)] = 1
    part[(mt.pd_active == 1) | (mt.lo_active == 1) | (mt.up_active == 1)] = 1
    return part


def _optimality_norms(ev, part, D, opts: SolveOptions):
    """SAP optimality test terms (contact_solve.py:1970-2087): ||D grad|| and
    eps_a + eps_r max(||D A v||, ||D j||) over the participating dofs, D = diag(A)^(-1/2)."""
    mask = part == 1
    gn = math.sqrt(float(np.sum((D[mask] * ev["grad"][mask]) ** 2)))
    pn = math.sqrt(float(np.sum((D[mask] * ev["Av"][mask]) ** 2)))
    jn = math.sqrt(float(np.sum((D[mask] * ev["jimp"][mask]) ** 2)))
    return gn, opts.optimality_abs_tol + opts.optimality_rel_tol * max(pn, jn)


def optimality_at(model, w, world_result: dict, v, dt, opts: SolveOptions | None = None):
    """Evaluate the SAP optimality test of world ``w``'s problem (as ``step_world`` built it,
    float64) at an arbitrary velocity ``v``: returns (||D grad||, tolerance, cost, gamma).
    Used to check a velocity computed elsewhere against the reference stopping rule."""
    opts = opts or SolveOptions()
    wv = WorldView(model, w)
    fm, cs, mt = world_result["free_motion"], world_result["contacts"], world_result["model_terms"]
    prob = SapProblem(fm.A_env, fm.v_star, cs, mt, dt, opts.beta, opts.sigma)
    part = _participating(model, wv, cs, mt)
    a_diag = np.array([d if (d >= 1.0e-12 and math.isfinite(d)) else 1.0e-12 for d in np.diag(fm.A_env)])
    ev = prob.evaluate(np.asarray(v, dtype=np.float64))
    gn, tol = _optimality_norms(ev, part, 1.0 / np.sqrt(a_diag), opts)
    return gn, tol, ev["cost"], ev["gamma"]


def optimality_on_data(model, w, world_result: dict, v, dt, opts: SolveOptions | None = None, *, A, v_star, J,
                       weight, phi0):
    """The SAP optimality test (contact_solve.py:1970-2087) in float64 at velocity ``v`` for world
    ``w``'s problem rebuilt from data formed elsewhere: a solver's dynamics matrix ``A`` (n x n),
    free-motion velocity ``v_star``, contact Jacobian rows ``J`` (nc x 3 x n), contact weights and
    signed distances, in place of the oracle's own (PD / limit terms are rebuilt from that A and v*,
    as build_model_terms does).  Checks that ``v`` is the optimum of the problem that solver
    formed, independently of how closely its float32 data rounds the oracle's.
    Returns (||D grad||, tolerance)."""
    opts = opts or SolveOptions()
    wv = WorldView(model, w)
    cs = world_result["contacts"]
    nc = cs.count
    A = np.asarray(A, dtype=np.float64)
    v_star = np.asarray(v_star, dtype=np.float64)
    cs2 = replace(cs, J=np.asarray(J, dtype=np.float64)[:nc].reshape(nc, 3, wv.n),
                  w=np.asarray(weight, dtype=np.float64)[:nc], phi0=np.asarray(phi0, dtype=np.float64)[:nc])
    mt = build_model_terms(model, wv, world_result["joint_q_in"], world_result["control_in"], A, world_result["v0"],
                           v_star, dt)
    prob = SapProblem(A, v_star, cs2, mt, dt, opts.beta, opts.sigma)
    part = _participating(model, wv, cs2, mt)
    a_diag = np.array([d if (d >= 1.0e-12 and math.isfinite(d)) else 1.0e-12 for d in np.diag(A)])
    ev = prob.evaluate(np.asarray(v, dtype=np.float64))
    return _optimality_norms(ev, part, 1.0 / np.sqrt(a_diag), opts)


def solve_on_data(model, w, world_result: dict, dt, opts: SolveOptions | None = None, *, A, v_star, J, weight, phi0,
                  v_guess=None) -> SolveResult:
    """SAP solve of world ``w``'s problem rebuilt from data formed elsewhere (as ``optimality_on_data``
    rebuilds it): the oracle's float64 answer to the problem a solver formed, so that solver's
    Newton, projection and line search can be checked apart from how its data rounds."""
    opts = opts or SolveOptions()
    wv = WorldView(model, w)
    cs = world_result["contacts"]
    nc = cs.count
    A = np.asarray(A, dtype=np.float64)
    v_star = np.asarray(v_star, dtype=np.float64)
    cs2 = replace(cs, J=np.asarray(J, dtype=np.float64)[:nc].reshape(nc, 3, wv.n),
                  w=np.asarray(weight, dtype=np.float64)[:nc], phi0=np.asarray(phi0, dtype=np.float64)[:nc])
    mt = build_model_terms(model, wv, world_result["joint_q_in"], world_result["control_in"], A, world_result["v0"],
                           v_star, dt)
    return sap_solve(model, wv, A, v_star, world_result["v0"], v_guess, cs2, mt, dt, opts)


def sap_solve(model, wv: WorldView, A, v_star, v0, v_guess, cs: ContactSet, mt: ModelTerms, dt,
              opts: SolveOptions) -> SolveResult:
    """SAP Newton solve of one world (contact_solve.py:6129-6285, 5474-5655)."""
    prob = SapProblem(A, v_star, cs, mt, dt, opts.beta, opts.sigma)
    v = np.array(v_guess if v_guess is not None else v0, dtype=np.float64)
    part = _participating(model, wv, cs, mt)
    unconstrained = cs.count == 0 and not np.any((mt.pd_active == 1) | (mt.lo_active == 1) | (mt.up_active == 1))
    if unconstrained:
        v = v_star.copy()
        return SolveResult(v=v, iterations=0, ls_iterations=0, converged=True, optimality_reached=False,
                           cost=0.0, gamma=np.zeros((0, 3)), grad_norm=0.0, opt_tol=0.0)
    v[part == 0] = v_star[part == 0]
    a_diag = np.array([d if (d >= 1.0e-12 and math.isfinite(d)) else 1.0e-12 for d in np.diag(A)])
    D = 1.0 / np.sqrt(a_diag)
    cost_min_alpha = 0.0 if opts.line_search == "monotone_decay" else 0.5
    alpha = 1.0
    previous_cost = 0.0
    ls_total = 0
    iteration = 0
    history = []
    status = 0
    while True:
        ev = prob.evaluate(v)
        gn, opt_tol = _optimality_norms(ev, part, D, opts)
        opt_reached = gn <= opt_tol
        cost_reached = False
        if iteration > 0 and alpha > cost_min_alpha:
            scale = 0.5 * (abs(ev["cost"]) + abs(previous_cost))
            if abs(ev["cost"] - previous_cost) < opts.cost_abs_tol + opts.cost_rel_tol * scale:
                cost_reached = True
        history.append(dict(iteration=iteration, cost=ev["cost"], grad_norm=gn, opt_tol=opt_tol, alpha=alpha))
        if opt_reached or cost_reached:
            return SolveResult(v=v, iterations=iteration, ls_iterations=ls_total, converged=True,
                               optimality_reached=opt_reached, cost=ev["cost"], gamma=ev["gamma"], grad_norm=gn,
                               opt_tol=opt_tol, status=status, history=history)
        if iteration >= opts.max_iterations:
            return SolveResult(v=v, iterations=iteration, ls_iterations=ls_total, converged=False,
                               optimality_reached=False, cost=ev["cost"], gamma=ev["gamma"], grad_norm=gn,
                               opt_tol=opt_tol, status=status, history=history)
        H = prob.hessian(ev)
        H = H + opts.diag_shift * np.eye(len(v))
        dv = np.linalg.solve(H, -ev["grad"])
        accepted, a_new, ls_it, status = line_search(prob, v, dv, ev, opts)
        if accepted:
            alpha = a_new
            previous_cost = ev["cost"]
            v = v + alpha * dv
            ls_total += ls_it
        elif opts.line_search == "monotone_decay":
            # the fused decay search leaves v, records the tried step and the current cost as the
            # previous one (contact_solve.py:2647-2662): the cost criterion ends the solve next
            alpha = a_new
            previous_cost = ev["cost"]
            ls_total += ls_it
        iteration += 1


def _line_data(prob: SapProblem, v, dv, ev):
    dp = prob.A @ dv
    r = v - prob.v_star
    return dict(dp=dp, dell0=float(ev["grad"] @ dv), dell_a0=float(dp @ r), d2ell_a=float(dv @ dp),
                m0=0.5 * float(r @ (prob.A @ r)), dvc=(prob.Jf @ dv).reshape(-1, 3), vc=(prob.Jf @ v).reshape(-1, 3))


def _trial(prob: SapProblem, v, dv, ld, alpha, need_second=False):
    """Cost, first and second derivative of phi(alpha) = ell(v + alpha dv) (contact_solve.py:2234-2420)."""
    vt = v + alpha * dv
    vc = (prob.Jf @ vt).reshape(-1, 3)
    gam, G, ccost = prob.contact_terms(vc)
    imp_m, hdiag, mcost = eval_model_terms(prob.mt, vt, prob.dt)
    cost = ld["m0"] + alpha * ld["dell_a0"] + 0.5 * alpha * alpha * ld["d2ell_a"] + ccost + mcost
    jimp = prob.Jf.T @ gam.reshape(-1) + imp_m
    d1 = float(ld["dp"] @ (vt - prob.v_star)) - float(dv @ jimp)
    d2 = 0.0
    if need_second:
        d2 = ld["d2ell_a"]
        for c in range(prob.cs.count):
            d2 += float(ld["dvc"][c] @ G[c] @ ld["dvc"][c])
        d2 += float(np.sum(hdiag * dv * dv))
    return cost, d1, d2


def line_search(prob, v, dv, ev, opts: SolveOptions):
    if opts.line_search == "exact_root":
        return exact_line_search(prob, v, dv, ev, opts)
    if opts.line_search == "armijo_decay":
        return armijo_line_search(prob, v, dv, ev, opts)
    if opts.line_search == "monotone_decay":
        return monotone_line_search(prob, v, dv, ev, opts)
    raise NotImplementedError(opts.line_search)


MONOTONE_DECAY = 0.5
MONOTONE_MIN_ALPHA = 1.0e-8
MONOTONE_RELAX_R = 1.0e-12
MONOTONE_RELAX_A = 1.0e-14


def monotone_line_search(prob, v, dv, ev, opts: SolveOptions):
    """Fused halving search (contact_solve.py:2486-2665, 5307-5415): the first alpha = 1, 1/2, 1/4 ...
    whose cost does not exceed the current one (relaxed by 1e-14 + 1e-12 |cost|); no descent test.
    Returns (accepted, alpha, trials, status); status -2 when alpha fell below 1e-8."""
    ld = _line_data(prob, v, dv, ev)
    current = ev["cost"]
    tol = MONOTONE_RELAX_A + MONOTONE_RELAX_R * abs(current)
    alpha = 1.0
    trials = 0
    for _ in range(opts.ls_max_iterations):
        ell, _, _ = _trial(prob, v, dv, ld, alpha)
        trials += 1
        if ell <= current + tol:
            return True, alpha, trials, 0
        alpha *= MONOTONE_DECAY
        if alpha < MONOTONE_MIN_ALPHA:
            return False, alpha, trials, -2
    return False, alpha, trials, 0


def armijo_line_search(prob, v, dv, ev, opts: SolveOptions):
    """Backtracking search (contact_solve.py:2670-2905, 5108-5305).  Returns (accepted, alpha, iters, status)."""
    ld = _line_data(prob, v, dv, ev)
    ell0 = ev["cost"]
    dell0 = ld["dell0"]
    amax = 1.0 / opts.rho
    rel = opts.relative_slop
    if dell0 >= 0.0 or not math.isfinite(dell0):
        return False, 1.0, 0, -1

    def armijo_ok(a, ell):
        return ell < ell0 + opts.armijo_c * a * dell0

    alpha = amax
    ell, d1, _ = _trial(prob, v, dv, ld, alpha)
    scale = max(1.0, 0.5 * (abs(ell) + abs(ell0)))
    slop = (rel / 10.0) * scale
    if d1 < 0.0 or d1 < slop:
        return True, alpha, 0, 0
    a_prev, e_prev = alpha, ell
    if opts.ls_max_iterations <= 1:
        return False, alpha, 0, -2
    iteration = 1
    while True:
        alpha = alpha * opts.rho
        ell, _, _ = _trial(prob, v, dv, ld, alpha)
        denom = alpha - a_prev
        dell_approx = (ell - e_prev) / denom if abs(denom) > 0.0 else 0.0
        if abs(dell_approx) < slop:
            return True, alpha, iteration, 0
        if ell > e_prev and armijo_ok(alpha, ell):
            if armijo_ok(a_prev, e_prev):
                return True, a_prev, iteration, 0
            return True, alpha, iteration, 0
        if iteration >= opts.ls_max_iterations - 1:
            if armijo_ok(alpha, ell):
                return True, alpha, opts.ls_max_iterations, 0
            return False, alpha, iteration, -2
        a_prev, e_prev = alpha, ell
        iteration += 1


def exact_line_search(prob, v, dv, ev, opts: SolveOptions):
    """Exact-root search (contact_solve.py:2913-3140, 5700-5908), cubic Hermite seed on."""
    ld = _line_data(prob, v, dv, ev)
    ell0 = ev["cost"]
    dell0 = ld["dell0"]
    amax = EXACT_ALPHA_MAX
    tol = EXACT_F_TOLERANCE
    if dell0 >= 0.0 or not math.isfinite(dell0):
        return False, amax, 0, -1
    ell, d1, d2 = _trial(prob, v, dv, ld, amax, need_second=True)
    if d1 <= 0.0:
        return True, amax, 0, 0
    if -dell0 < opts.cost_abs_tol + opts.cost_rel_tol * ell:
        return True, 1.0, 0, 0
    scale = -dell0
    f_upper = d1 / scale
    if not math.isfinite(f_upper) or f_upper <= 0.0:
        return False, amax, 0, -3
    quad = -dell0 / d2
    guess = quad
    h = amax
    theta = dell0 + d1 - 3.0 * (ell - ell0) / h
    disc = theta * theta - dell0 * d1
    if math.isfinite(disc) and disc >= 0.0:
        gamma = math.sqrt(disc)
        denom = d1 - dell0 + 2.0 * gamma
        if denom != 0.0:
            cubic = h - h * (d1 + gamma - theta) / denom
            if math.isfinite(cubic) and 0.0 < cubic <= h:
                guess = cubic
    if guess > amax:
        guess = amax
    if not math.isfinite(guess) or guess < 0.0 or guess > amax or tol * guess <= 0.0:
        return False, amax, 0, -4
    x_lo, x_hi = 0.0, amax
    f_hi = f_upper
    root = guess
    minus_dx = -amax
    minus_dx_prev = -amax
    x_tol = tol * guess
    iteration = 1
    while True:
        _, d1, d2 = _trial(prob, v, dv, ld, root, need_second=True)
        f = d1 / scale
        df = d2 / scale
        if (f < 0.0 and f_hi >= 0.0) or (f >= 0.0 and f_hi < 0.0):
            x_lo = root
        else:
            x_hi = root
            f_hi = f
        if abs(f) < tol:
            return True, root, iteration, 0
        if iteration >= opts.ls_max_iterations:
            return False, root, iteration, -5
        slow = 2.0 * abs(f) > abs(minus_dx_prev * df)
        minus_dx_prev = minus_dx
        bis_dx = 0.5 * (x_lo - x_hi)
        bis_root = x_lo - bis_dx
        newton_dx = 0.0
        newton_root = root
        if abs(df) > 0.0:
            newton_dx = f / df
            newton_root = root - newton_dx
        use_bisect = slow or newton_root < x_lo or newton_root > x_hi or not math.isfinite(newton_root) \
            or not math.isfinite(newton_dx)
        nxt, ndx = (bis_root, bis_dx) if use_bisect else (newton_root, newton_dx)
        if abs(ndx) < x_tol:
            return True, nxt, iteration, 0
        root = nxt
        minus_dx = ndx
        iteration += 1


# =============================================================================
# integration + public FK (solver_sap.py, sap_kinematics.py)
# =============================================================================
def _quat_sap_euler(q, w, h):
    """``_quat_sap_euler_xyzw`` solver_sap.py:161-181."""
    qx, qy, qz, qw = q
    hd = 0.5 * h
    x = qx + hd * (qw * w[0] + w[1] * qz - w[2] * qy)
    y = qy + hd * (qw * w[1] + w[2] * qx - w[0] * qz)
    z = qz + hd * (qw * w[2] + w[0] * qy - w[1] * qx)
    s = qw - hd * (w[0] * qx + w[1] * qy + w[2] * qz)
    nrm = math.sqrt(x * x + y * y + z * z + s * s)
    if nrm > 0.0:
        x, y, z, s = x / nrm, y / nrm, z / nrm, s / nrm
    return np.array([x, y, z, s])


def _quat_midpoint(q, w_mid, h):
    """``_quat_midpoint_xyzw_f32`` solver_sap.py:227-244."""
    angle = float(np.linalg.norm(w_mid)) * h
    if angle > 1.0e-12:
        axis = w_mid / np.linalg.norm(w_mid)
        s = math.sin(0.5 * angle)
        c = math.cos(0.5 * angle)
        r = np.array([axis[0] * s, axis[1] * s, axis[2] * s, c])
        out = qmul(r, q)
        nrm = math.sqrt(float(out @ out))
        if nrm > 0.0:
            out = out / nrm
        return out
    return np.asarray(q, dtype=np.float64)


def integrate(model, wv: WorldView, joint_q, v_prev, v_new, dt, mode="midpoint"):
    """f32-output integration (solver_sap.py:288-519); returns (joint_q_out, joint_qd_out public)."""
    q_out = np.asarray(joint_q[wv.c0:wv.c1], dtype=np.float64).copy()
    qd_out = np.zeros(wv.n)
    for j in range(wv.j0, wv.j1):
        jt = int(model.joint_type[j])
        qs = int(model.joint_q_start[j]) - wv.c0
        ds = int(model.joint_qd_start[j]) - wv.d0
        nd = int(model.joint_qd_start[j + 1]) - int(model.joint_qd_start[j])
        if jt in (JOINT_PRISMATIC, JOINT_REVOLUTE) or (jt not in (JOINT_FREE, JOINT_DISTANCE, JOINT_BALL, JOINT_FIXED)):
            for a in range(nd):
                vmid = 0.5 * (v_prev[ds + a] + v_new[ds + a]) if mode == "midpoint" else v_new[ds + a]
                q_out[qs + a] = q_out[qs + a] + dt * vmid
                qd_out[ds + a] = v_new[ds + a]
        elif jt == JOINT_BALL:
            # quaternion coordinates, angular velocity in the joint frame (solver_sap.py:316-330, 425-450)
            q = q_out[qs:qs + 4].copy()
            w_new = v_new[ds:ds + 3]
            if mode == "midpoint":
                q_out[qs:qs + 4] = _quat_midpoint(q, 0.5 * (v_prev[ds:ds + 3] + w_new), dt)
            else:
                q_out[qs:qs + 4] = _quat_sap_euler(q, w_new, dt)
            qd_out[ds:ds + 3] = w_new
        elif jt in (JOINT_FREE, JOINT_DISTANCE):
            q = q_out[qs + 3:qs + 7].copy()
            w_new = v_new[ds:ds + 3]
            vo_new = v_new[ds + 3:ds + 6]
            if mode == "midpoint":
                w_mid = 0.5 * (v_prev[ds:ds + 3] + w_new)
                vo_mid = 0.5 * (v_prev[ds + 3:ds + 6] + vo_new)
                qn = _quat_midpoint(q, w_mid, dt)
            else:
                vo_mid = vo_new
                qn = _quat_sap_euler(q, w_new, dt)
            q_out[qs:qs + 3] = q_out[qs:qs + 3] + dt * vo_mid
            q_out[qs + 3:qs + 7] = qn
            child = int(model.joint_child[j])
            r = qrot(qn, model.body_com[child].astype(np.float64))
            qd_out[ds:ds + 3] = vo_new + np.cross(w_new, r)
            qd_out[ds + 3:ds + 6] = w_new
    return q_out, qd_out


def public_fk(model, wv: WorldView, joint_q_world, joint_qd_world):
    """Public FK of one world (sap_kinematics.py:84-229)."""
    nb = wv.nb
    body_q = np.zeros((nb, 7))
    body_q[:, 6] = 1.0
    body_qd = np.zeros((nb, 6))
    com = model.body_com.astype(np.float64)

    def tmul(a, b):
        return np.concatenate([a[:3] + qrot_unit(a[3:], b[:3]), qmul(a[3:], b[3:])])

    def tinv(a):
        qi = qconj(a[3:])
        return np.concatenate([-qrot_unit(qi, a[:3]), qi])

    for j in range(wv.j0, wv.j1):
        parent = int(model.joint_parent[j])
        child = int(model.joint_child[j]) - wv.b0
        jt = int(model.joint_type[j])
        X_pj = model.joint_X_p[j].astype(np.float64)
        X_cj = model.joint_X_c[j].astype(np.float64)
        X_wpj = X_pj
        v_wpj = np.zeros(6)
        if parent >= 0:
            lp = parent - wv.b0
            X_wp = body_q[lp]
            X_wpj = tmul(X_wp, X_pj)
            r_p = X_wpj[:3] - (X_wp[:3] + qrot_unit(X_wp[3:], com[parent]))
            w_p = body_qd[lp, 3:]
            v_wpj = np.concatenate([body_qd[lp, :3] + np.cross(w_p, r_p), w_p])
        qs = int(model.joint_q_start[j]) - wv.c0
        ds = int(model.joint_qd_start[j]) - wv.d0
        X_j = np.array([0, 0, 0, 0, 0, 0, 1.0])
        v_j = np.zeros(6)
        if jt == JOINT_PRISMATIC:
            ax = model.joint_axis[wv.d0 + ds].astype(np.float64)
            X_j = np.concatenate([ax * joint_q_world[qs], [0, 0, 0, 1.0]])
            v_j = np.concatenate([ax * joint_qd_world[ds], np.zeros(3)])
        elif jt == JOINT_REVOLUTE:
            ax = model.joint_axis[wv.d0 + ds].astype(np.float64)
            X_j = np.concatenate([np.zeros(3), q_from_axis_angle(ax, joint_q_world[qs])])
            v_j = np.concatenate([np.zeros(3), ax * joint_qd_world[ds]])
        elif jt == JOINT_BALL:
            X_j = np.concatenate([np.zeros(3), np.asarray(joint_q_world[qs:qs + 4], dtype=np.float64)])
            v_j = np.concatenate([np.zeros(3), np.asarray(joint_qd_world[ds:ds + 3], dtype=np.float64)])
        elif jt in (JOINT_FREE, JOINT_DISTANCE):
            X_j = np.asarray(joint_q_world[qs:qs + 7], dtype=np.float64)
            v_j = np.asarray(joint_qd_world[ds:ds + 6], dtype=np.float64)
        elif jt == JOINT_D6:
            # sap_kinematics.py:170-218: linear axes sum into the offset, the angular ones rotate
            nl, na = int(model.joint_dof_dim[j, 0]), int(model.joint_dof_dim[j, 1])
            axes = [model.joint_axis[wv.d0 + ds + k].astype(np.float64) for k in range(nl + na)]
            pos = np.zeros(3)
            vv = np.zeros(3)
            for k in range(nl):
                pos = pos + axes[k] * joint_q_world[qs + k]
                vv = vv + axes[k] * joint_qd_world[ds + k]
            rot, ang = d6_rotation(axes[nl:], joint_q_world[qs + nl:qs + nl + na])
            vw = np.zeros(3)
            for k in range(na):
                vw = vw + ang[k] * joint_qd_world[ds + nl + k]
            X_j = np.concatenate([pos, rot])
            v_j = np.concatenate([vv, vw])
        X_wc = tmul(tmul(X_wpj, X_j), tinv(X_cj))
        body_q[child] = X_wc
        body_qd[child] = v_wpj + np.concatenate([qrot_unit(X_wpj[3:], v_j[:3]), qrot_unit(X_wpj[3:], v_j[3:])])
    return body_q, body_qd


# =============================================================================
# full step
# =============================================================================
@dataclass
class StepOptions:
    preset: str = "approx32"
    line_search: str = "armijo_decay"
    max_rigid_contact: int = 128
    fallback_mu: float | None = None
    fallback_stiffness: float = 1.0e10
    contact_tau_d: float | None = None
    solve: SolveOptions | None = None

    def resolved(self):
        presets = {
            "approx32": ("body_inertia", "witness_point", "midpoint"),
            "approx64": ("body_inertia", "witness_point", "midpoint"),
            "drake": ("diag_delassus", "contact_midpoint", "sap_euler"),
        }
        return presets[self.preset]


def infer_mu_fallback(model):
    """``_infer_contact_mu_fallback`` sap_helpers.py:40-45."""
    mu = np.asarray(model.shape_material_mu, dtype=np.float32)
    pos = mu[mu > 0.0]
    return 0.1 if pos.size == 0 else float(np.min(pos))


def infer_tau_fallback(model):
    """``_infer_contact_tau_d_fallback`` solver_sap.py:128-142."""
    vals = np.asarray(model.shape_material_tau, dtype=np.float64).reshape(-1)
    ok = vals[np.isfinite(vals) & (vals >= 0.0)]
    return float(ok[0]) if ok.size else 0.0


def step_world(model, w, state: dict, control: dict, contacts: dict, dt, opts: StepOptions, v_guess=None):
    """One SolverSAP.step for world ``w`` (solver_sap.py:1388-1494).

    ``state`` / ``control`` / ``contacts`` are dicts of global numpy arrays in
    public order.  Returns a dict with the world's new public state, SAP
    velocity and solver statistics.
    """
    wv = WorldView(model, w)
    weight_mode, point_mode, integration = opts.resolved()
    so = opts.solve or SolveOptions.for_variant(opts.line_search)
    jq = np.asarray(state["joint_q"], dtype=np.float64)
    v0 = public_to_sap_velocity(model, wv, state["body_q"], state["joint_qd"])
    f_sap = public_to_sap_force(model, wv, state["body_q"], control["joint_f"])
    bf = public_body_force_to_sap(model, wv, state["body_q"], state.get("body_f"))
    fm = free_motion(model, wv, jq[wv.c0:wv.c1], v0, f_sap, bf, dt)
    fallback_mu = opts.fallback_mu if opts.fallback_mu is not None else infer_mu_fallback(model)
    tau_fb = opts.contact_tau_d if opts.contact_tau_d is not None else infer_tau_fallback(model)
    cs = world_contacts(model, wv, fm, contacts, max_contacts=opts.max_rigid_contact,
                        contact_point_mode=point_mode, contact_weight_mode=weight_mode,
                        fallback_mu=fallback_mu, fallback_k=opts.fallback_stiffness, fallback_tau=tau_fb)
    mt = build_model_terms(model, wv, jq, control, fm.A_env, v0, fm.v_star, dt)
    res = sap_solve(model, wv, fm.A_env, fm.v_star, v0, v_guess, cs, mt, dt, so)
    q_out, qd_out = integrate(model, wv, jq, v0, res.v, dt, mode=integration)
    body_q, body_qd = public_fk(model, wv, q_out, qd_out)
    return dict(joint_q=q_out, joint_qd=qd_out, body_q=body_q, body_qd=body_qd, v_sap=res.v, v_star=fm.v_star,
                v0=v0, result=res, contacts=cs, free_motion=fm, model_terms=mt, joint_q_in=jq, control_in=control)


def step(model, state: dict, control: dict, contacts: dict, dt, opts: StepOptions, v_guess=None, worlds=None):
    """All worlds (or ``worlds``): returns global arrays + per-world results.  ``dt`` is a scalar or
    one timestep per world (``SolverSAP._update_dt_world``, solver_sap.py:975-992)."""
    worlds = range(model.world_count) if worlds is None else worlds
    dts = np.broadcast_to(np.asarray(dt, dtype=np.float64), (model.world_count,))
    out = dict(joint_q=np.asarray(state["joint_q"], dtype=np.float64).copy(),
               joint_qd=np.asarray(state["joint_qd"], dtype=np.float64).copy(),
               body_q=np.asarray(state["body_q"], dtype=np.float64).copy(),
               body_qd=np.asarray(state["body_qd"], dtype=np.float64).copy(),
               v_sap=np.zeros(model.joint_dof_count), per_world={})
    for w in worlds:
        wv = WorldView(model, w)
        guess = None if v_guess is None else v_guess[wv.d0:wv.d1]
        r = step_world(model, w, state, control, contacts, float(dts[w]), opts, guess)
        out["joint_q"][wv.c0:wv.c1] = r["joint_q"]
        out["joint_qd"][wv.d0:wv.d1] = r["joint_qd"]
        out["body_q"][wv.b0:wv.b1] = r["body_q"]
        out["body_qd"][wv.b0:wv.b1] = r["body_qd"]
        out["v_sap"][wv.d0:wv.d1] = r["v_sap"]
        out["per_world"][w] = r
    return out

"""Host-side public forward kinematics used by the scene builders.

Only needed at model construction (to make ``body_q``/``body_qd`` consistent
with ``joint_q``/``joint_qd``).  The per-step FK runs on the GPU inside the
fused step kernel.  Semantics follow ``sap_eval_single_articulation_fk``
(``sim/sap_kinematics.py:84-229``): ``body_qd`` is ``[linear, angular]`` where
the linear part is propagated from the parent's centre of mass to the joint
anchor and then carried over unchanged by rotational joints.  Worlds share
one topology, so each joint is evaluated for all worlds at once.
"""
from __future__ import annotations

import numpy as np

from .runtime import (
    SAP_JOINT_BALL,
    SAP_JOINT_D6,
    SAP_JOINT_DISTANCE,
    SAP_JOINT_FREE,
    SAP_JOINT_PRISMATIC,
    SAP_JOINT_REVOLUTE,
)


def _qmul(a, b):
    ax, ay, az, aw = a[..., 0], a[..., 1], a[..., 2], a[..., 3]
    bx, by, bz, bw = b[..., 0], b[..., 1], b[..., 2], b[..., 3]
    return np.stack([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz], axis=-1)


def _qrot(q, v):
    qv = q[..., :3]
    t = 2.0 * np.cross(qv, v)
    return v + q[..., 3:4] * t + np.cross(qv, t)


def _tmul(a, b):
    return np.concatenate([a[..., :3] + _qrot(a[..., 3:], b[..., :3]), _qmul(a[..., 3:], b[..., 3:])], axis=-1)


def _tinv(a):
    qi = a[..., 3:] * np.array([-1.0, -1.0, -1.0, 1.0])
    return np.concatenate([-_qrot(qi, a[..., :3]), qi], axis=-1)


def quat_from_cols(c0, c1, c2):
    """Unit quaternion (xyzw) of the 3x3 matrix with columns c0, c1, c2: the trace branch when the
    trace is positive, else the branch of the largest diagonal entry, then normalised (the
    construction of ``_quatd_from_matrix_cols``, sap_helpers.py:607-652)."""
    r00, r10, r20 = (float(x) for x in c0)
    r01, r11, r21 = (float(x) for x in c1)
    r02, r12, r22 = (float(x) for x in c2)
    tr = r00 + r11 + r22
    if tr > 0.0:
        s = np.sqrt(tr + 1.0) * 2.0
        q = np.array([r21 - r12, r02 - r20, r10 - r01, 0.25 * s * s]) / s
    elif r00 > r11 and r00 > r22:
        s = np.sqrt(1.0 + r00 - r11 - r22) * 2.0
        q = np.array([0.25 * s * s, r01 + r10, r02 + r20, r21 - r12]) / s
    elif r11 > r22:
        s = np.sqrt(1.0 + r11 - r00 - r22) * 2.0
        q = np.array([r01 + r10, 0.25 * s * s, r12 + r21, r02 - r20]) / s
    else:
        s = np.sqrt(1.0 + r22 - r00 - r11) * 2.0
        q = np.array([r02 + r20, r12 + r21, 0.25 * s * s, r10 - r01]) / s
    n = np.linalg.norm(q)
    return q / n if n > 0.0 else np.array([0.0, 0.0, 0.0, 1.0])


def d6_frame_axes(ang_axes):
    """The rotation axes of a D6 joint's angular coordinates at zero angles.

    One axis is used as given.  Two or three axes are replaced by the columns of the rotation
    whose quaternion is built from the matrix [a0, a1, a2] (a2 = a0 x a1 for two axes), so
    orthonormal right-handed axes come back unchanged (sap_helpers.py:1109-1161)."""
    m = len(ang_axes)
    if m <= 1:
        return [np.asarray(a, dtype=np.float64) for a in ang_axes]
    a0, a1 = np.asarray(ang_axes[0], np.float64), np.asarray(ang_axes[1], np.float64)
    a2 = np.cross(a0, a1) if m == 2 else np.asarray(ang_axes[2], np.float64)
    q = quat_from_cols(a0, a1, a2)
    return [_qrot(q, e) for e in np.eye(3)[:m]]


def _axis_angle(ax, ang):
    """(W, 4) quaternions of rotations by ``ang`` (W,) about the fixed axis ``ax``."""
    h = 0.5 * np.asarray(ang, dtype=np.float64)
    return np.concatenate([np.asarray(ax)[None, :] * np.sin(h)[:, None], np.cos(h)[:, None]], axis=-1)


def d6_joint_motion(lin_axes, ang_axes, q, qd):
    """Joint transform and joint-frame velocity of D6 joints, vectorised over worlds.

    ``q``/``qd`` are (W, nl + na).  The angular coordinates are intrinsic rotations about the
    frame axes l_k of :func:`d6_frame_axes`: R = R(l0, q0) R(l1, q1) R(l2, q2), whose angular
    velocity is l0 qd0 + R(l0, q0) l1 qd1 + R(l0, q0) R(l1, q1) l2 qd2; the linear coordinates
    translate along their fixed axes (sap_kinematics.py:170-218).  Returns ((W, 7), (W, 6)) with
    the velocity as [linear, angular]."""
    W = q.shape[0]
    nl = len(lin_axes)
    X = np.zeros((W, 7))
    X[:, 6] = 1.0
    v = np.zeros((W, 6))
    for k, a in enumerate(lin_axes):
        X[:, :3] += np.asarray(a, np.float64)[None, :] * q[:, k:k + 1]
        v[:, :3] += np.asarray(a, np.float64)[None, :] * qd[:, k:k + 1]
    rot = X[:, 3:].copy()
    for k, l in enumerate(d6_frame_axes(ang_axes)):
        v[:, 3:] += _qrot(rot, np.broadcast_to(l, (W, 3))) * qd[:, nl + k:nl + k + 1]
        rot = _qmul(rot, _axis_angle(l, q[:, nl + k]))
    X[:, 3:] = rot
    return X, v


def eval_fk_public(model, joint_q, joint_qd):
    """Return (body_q[B,7], body_qd[B,6]) in float64."""
    W = max(int(model.world_count), 1)
    J = model.joint_count // W
    NB = model.body_count // W
    jq = np.asarray(joint_q, dtype=np.float64)
    jqd = np.asarray(joint_qd, dtype=np.float64)
    body_q = np.zeros((model.body_count, 7))
    body_q[:, 6] = 1.0
    body_qd = np.zeros((model.body_count, 6))
    com = model.body_com.astype(np.float64)
    widx = np.arange(W)
    for j in range(J):
        gj = widx * J + j
        parent = model.joint_parent[gj].astype(np.int64)
        child = model.joint_child[gj].astype(np.int64)
        jt = int(model.joint_type[j])
        X_pj = model.joint_X_p[gj].astype(np.float64)
        X_cj = model.joint_X_c[gj].astype(np.float64)
        X_wpj = X_pj.copy()
        v_wpj = np.zeros((W, 6))
        has_parent = parent >= 0
        if np.any(has_parent):
            p = parent[has_parent]
            X_wp = body_q[p]
            X_wpj[has_parent] = _tmul(X_wp, X_pj[has_parent])
            r_p = X_wpj[has_parent, :3] - (X_wp[:, :3] + _qrot(X_wp[:, 3:], com[p]))
            w_p = body_qd[p, 3:]
            v_wpj[has_parent] = np.concatenate([body_qd[p, :3] + np.cross(w_p, r_p), w_p], axis=-1)
        qs = model.joint_q_start[gj].astype(np.int64)
        ds = model.joint_qd_start[gj].astype(np.int64)
        X_j = np.tile(np.array([0, 0, 0, 0, 0, 0, 1.0]), (W, 1))
        v_j = np.zeros((W, 6))
        if jt in (SAP_JOINT_PRISMATIC, SAP_JOINT_REVOLUTE):
            ax = model.joint_axis[ds].astype(np.float64)
            if jt == SAP_JOINT_PRISMATIC:
                X_j[:, :3] = ax * jq[qs][:, None]
                v_j[:, :3] = ax * jqd[ds][:, None]
            else:
                half = 0.5 * jq[qs]
                X_j[:, 3:6] = ax * np.sin(half)[:, None]
                X_j[:, 6] = np.cos(half)
                v_j[:, 3:] = ax * jqd[ds][:, None]
        elif jt == SAP_JOINT_BALL:
            X_j[:, 3:] = jq[qs[:, None] + np.arange(4)]
            v_j[:, 3:] = jqd[ds[:, None] + np.arange(3)]
        elif jt in (SAP_JOINT_FREE, SAP_JOINT_DISTANCE):
            X_j = jq[qs[:, None] + np.arange(7)]
            v_j = jqd[ds[:, None] + np.arange(6)]
        elif jt == SAP_JOINT_D6:
            nl, na = int(model.joint_dof_dim[j, 0]), int(model.joint_dof_dim[j, 1])
            axes = model.joint_axis[int(model.joint_qd_start[j]) + np.arange(nl + na)].astype(np.float64)
            X_j, v_j = d6_joint_motion(axes[:nl], axes[nl:], jq[qs[:, None] + np.arange(nl + na)],
                                       jqd[ds[:, None] + np.arange(nl + na)])
        X_wc = _tmul(_tmul(X_wpj, X_j), _tinv(X_cj))
        lin = _qrot(X_wpj[:, 3:], v_j[:, :3])
        ang = _qrot(X_wpj[:, 3:], v_j[:, 3:])
        body_q[child] = X_wc
        body_qd[child] = v_wpj + np.concatenate([lin, ang], axis=-1)
    return body_q, body_qd

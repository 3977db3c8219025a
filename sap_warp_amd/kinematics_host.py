"""Host-side public forward kinematics used by the scene builders.

Only needed at model construction (to make ``body_q``/``body_qd`` consistent
with ``joint_q``/``joint_qd``).  The per-step FK runs on the GPU inside the
fused step kernel.  Semantics follow ``sap_eval_single_articulation_fk``
(``sim/sap_kinematics.py:84-229``): ``body_qd`` is ``[linear, angular]`` where
the linear part is propagated from the parent's centre of mass to the joint
anchor and then carried over unchanged by rotational joints.
"""
from __future__ import annotations

import numpy as np

from .runtime import SAP_JOINT_BALL, SAP_JOINT_FREE, SAP_JOINT_DISTANCE, SAP_JOINT_PRISMATIC, SAP_JOINT_REVOLUTE


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _qrot(q, v):
    qv = np.asarray(q[:3], dtype=np.float64)
    t = 2.0 * np.cross(qv, v)
    return v + q[3] * t + np.cross(qv, t)


def _tmul(a, b):
    return np.concatenate([a[:3] + _qrot(a[3:], b[:3]), _qmul(a[3:], b[3:])])


def _tinv(a):
    qi = np.array([-a[3], -a[4], -a[5], a[6]])
    return np.concatenate([-_qrot(qi, a[:3]), qi])


def _axis_angle(axis, angle):
    s = np.sin(0.5 * angle)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, np.cos(0.5 * angle)])


def eval_fk_public(model, joint_q, joint_qd):
    """Return (body_q[B,7], body_qd[B,6]) in float64."""
    jq = np.asarray(joint_q, dtype=np.float64)
    jqd = np.asarray(joint_qd, dtype=np.float64)
    body_q = np.zeros((model.body_count, 7))
    body_q[:, 6] = 1.0
    body_qd = np.zeros((model.body_count, 6))
    com = model.body_com.astype(np.float64)
    for i in range(model.joint_count):
        parent = int(model.joint_parent[i])
        child = int(model.joint_child[i])
        jt = int(model.joint_type[i])
        X_pj = model.joint_X_p[i].astype(np.float64)
        X_cj = model.joint_X_c[i].astype(np.float64)
        X_wpj = X_pj
        v_wpj = np.zeros(6)
        if parent >= 0:
            X_wp = body_q[parent]
            X_wpj = _tmul(X_wp, X_pj)
            r_p = X_wpj[:3] - (X_wp[:3] + _qrot(X_wp[3:], com[parent]))
            w_p = body_qd[parent, 3:]
            v_p = body_qd[parent, :3] + np.cross(w_p, r_p)
            v_wpj = np.concatenate([v_p, w_p])
        qs = int(model.joint_q_start[i])
        ds = int(model.joint_qd_start[i])
        X_j = np.array([0, 0, 0, 0, 0, 0, 1.0])
        v_j = np.zeros(6)
        if jt == SAP_JOINT_PRISMATIC:
            ax = model.joint_axis[ds].astype(np.float64)
            X_j = np.concatenate([ax * jq[qs], [0, 0, 0, 1.0]])
            v_j = np.concatenate([ax * jqd[ds], np.zeros(3)])
        elif jt == SAP_JOINT_REVOLUTE:
            ax = model.joint_axis[ds].astype(np.float64)
            X_j = np.concatenate([np.zeros(3), _axis_angle(ax, jq[qs])])
            v_j = np.concatenate([np.zeros(3), ax * jqd[ds]])
        elif jt == SAP_JOINT_BALL:
            X_j = np.concatenate([np.zeros(3), jq[qs:qs + 4]])
            v_j = np.concatenate([np.zeros(3), jqd[ds:ds + 3]])
        elif jt in (SAP_JOINT_FREE, SAP_JOINT_DISTANCE):
            X_j = jq[qs:qs + 7].copy()
            v_j = jqd[ds:ds + 6].copy()
        X_wcj = _tmul(X_wpj, X_j)
        X_wc = _tmul(X_wcj, _tinv(X_cj))
        lin = _qrot(X_wpj[3:], v_j[:3])
        ang = _qrot(X_wpj[3:], v_j[3:])
        body_q[child] = X_wc
        body_qd[child] = v_wpj + np.concatenate([lin, ang])
    return body_q, body_qd

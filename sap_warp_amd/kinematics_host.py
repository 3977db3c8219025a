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

from .runtime import SAP_JOINT_BALL, SAP_JOINT_DISTANCE, SAP_JOINT_FREE, SAP_JOINT_PRISMATIC, SAP_JOINT_REVOLUTE


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
        X_wc = _tmul(_tmul(X_wpj, X_j), _tinv(X_cj))
        lin = _qrot(X_wpj[:, 3:], v_j[:, :3])
        ang = _qrot(X_wpj[:, 3:], v_j[:, 3:])
        body_q[child] = X_wc
        body_qd[child] = v_wpj + np.concatenate([lin, ang], axis=-1)
    return body_q, body_qd

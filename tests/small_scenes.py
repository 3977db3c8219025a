"""Small hand-built scenes and finite-difference kinematics for the oracle checks.

Everything here is independent of ``oracle/``: positions come from the public
forward kinematics of ``sap_warp_amd.kinematics_host`` (joint transforms only,
no velocities), quaternion algebra from ``sap_warp_amd.scenes``.
"""
from __future__ import annotations

import math

import numpy as np

from sap_warp_amd.runtime import (
    GEO_BOX,
    GEO_CAPSULE,
    GEO_SPHERE,
    SAP_JOINT_BALL,
    SAP_JOINT_D6,
    SAP_JOINT_FREE,
    SAP_JOINT_PRISMATIC,
    SAP_JOINT_REVOLUTE,
)
from sap_warp_amd.kinematics_host import eval_fk_public
from sap_warp_amd.scenes import (
    DofCfg,
    SceneBuilder,
    ShapeCfg,
    quat_from_axis_angle,
    quat_mul,
    quat_to_matrix,
    tf,
)


# ----------------------------------------------------------------------------
# scenes
# ----------------------------------------------------------------------------
def free_box_scene(offset=(0.12, -0.05, 0.2), half=(0.1, 0.2, 0.35), shape_rot=None, pose=None, density=700.0):
    """One free body whose only shape is an off-centre (optionally rotated) box."""
    b = SceneBuilder(gravity=-9.81)
    q_shape = (0.0, 0.0, 0.0, 1.0) if shape_rot is None else shape_rot
    p = tf((0.3, -0.2, 1.1), quat_from_axis_angle((1.0, 2.0, -0.5), 0.7)) if pose is None else pose
    body = b.add_body(p)
    b.add_shape(body, GEO_BOX, tf(offset, q_shape), half, ShapeCfg(density=density))
    j = b.add_free_joint(body)
    b.add_articulation([j])
    return b.finalize()


def fixed_base_tree_scene(armature=0.02):
    """Fixed-base tree with revolute and prismatic joints (every coordinate is a true coordinate).

    root (revolute y at the world anchor) -> arm A (revolute x) -> hand (revolute about a skew axis)
                                          -> slider B (prismatic along a skew axis)
    Shapes are offset from the joint anchors so every COM is off-axis.
    """
    b = SceneBuilder(gravity=-9.81)
    cfg = ShapeCfg(density=900.0)
    root = b.add_body(tf())
    b.add_shape(root, GEO_CAPSULE, tf((0.02, 0.0, -0.25), quat_from_axis_angle((1, 0, 0), 0.3)), (0.05, 0.18, 0.0), cfg)
    j0 = b.add_joint(-1, root, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=(0, 1, 0), armature=armature)],
                     parent_xform=tf((0.0, 0.0, 1.5)))
    arm = b.add_body(tf())
    b.add_shape(arm, GEO_BOX, tf((0.0, 0.05, -0.2)), (0.04, 0.06, 0.2), cfg)
    j1 = b.add_joint(root, arm, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=(1, 0, 0), armature=armature)],
                     parent_xform=tf((0.0, 0.1, -0.5)), child_xform=tf((0.0, 0.0, 0.03)))
    hand = b.add_body(tf())
    b.add_shape(hand, GEO_SPHERE, tf((0.05, 0.0, -0.08)), (0.06, 0.0, 0.0), cfg)
    ax = np.array([0.3, 0.5, 0.8]) / np.linalg.norm([0.3, 0.5, 0.8])
    j2 = b.add_joint(arm, hand, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=tuple(ax), armature=armature)],
                     parent_xform=tf((0.0, 0.0, -0.42), quat_from_axis_angle((0, 0, 1), 0.4)))
    slider = b.add_body(tf())
    b.add_shape(slider, GEO_BOX, tf((0.03, 0.0, 0.0)), (0.05, 0.05, 0.05), cfg)
    pax = np.array([1.0, 0.0, 0.2]) / np.linalg.norm([1.0, 0.0, 0.2])
    j3 = b.add_joint(root, slider, SAP_JOINT_PRISMATIC, linear_axes=[DofCfg(axis=tuple(pax), armature=armature)],
                     parent_xform=tf((0.1, -0.1, -0.3), quat_from_axis_angle((0, 1, 0), 0.5)))
    b.add_articulation([j0, j1, j2, j3])
    return b.finalize()


def _unit(v):
    v = np.asarray(v, dtype=np.float64)
    return tuple(v / np.linalg.norm(v))


def d6_tree_scene(kind="mixed", armature=0.02, as_plain=False, limits=False):
    """Fixed-base tree built from D6 joints.

    ``kind``:
      * "single": one-axis D6 joints only (two angular, one linear), the tree of
        :func:`fixed_base_tree_scene`'s shape; with ``as_plain`` the same tree with revolute /
        prismatic joints instead, for the equivalence checks;
      * "one_angle": D6 joints with at most one angular axis (1 lin + 1 ang, 2 lin + 1 ang, 3 lin);
      * "mixed": 3 angular (about a rotated orthonormal frame), 1 lin + 2 ang, 2 ang (skew pair,
        orthogonalised by the joint) and a revolute.
    ``limits`` gives every one-axis dof a finite range (lower 0.05 below, upper 0.1 above zero).
    """
    b = SceneBuilder(gravity=-9.81)
    cfg = ShapeCfg(density=900.0)
    rng_axes = quat_to_matrix(quat_from_axis_angle((0.3, -0.8, 0.5), 0.9))

    def dof(axis):
        kw = dict(axis=_unit(axis), armature=armature)
        if limits:
            kw.update(limit_lower=-0.05, limit_upper=0.1, limit_ke=1.0e4, limit_kd=10.0)
        return DofCfg(**kw)

    def joint(parent, child, lin, ang, **kw):
        if as_plain:
            assert len(lin) + len(ang) == 1
            jt = SAP_JOINT_PRISMATIC if lin else SAP_JOINT_REVOLUTE
            return b.add_joint(parent, child, jt, linear_axes=[dof(a) for a in lin], angular_axes=[dof(a) for a in ang], **kw)
        return b.add_joint(parent, child, SAP_JOINT_D6, linear_axes=[dof(a) for a in lin],
                           angular_axes=[dof(a) for a in ang], **kw)

    root = b.add_body(tf())
    b.add_shape(root, GEO_CAPSULE, tf((0.02, 0.0, -0.25), quat_from_axis_angle((1, 0, 0), 0.3)), (0.05, 0.18, 0.0), cfg)
    arm = b.add_body(tf())
    b.add_shape(arm, GEO_BOX, tf((0.0, 0.05, -0.2)), (0.04, 0.06, 0.2), cfg)
    hand = b.add_body(tf())
    b.add_shape(hand, GEO_SPHERE, tf((0.05, 0.0, -0.08)), (0.06, 0.0, 0.0), cfg)
    slider = b.add_body(tf())
    b.add_shape(slider, GEO_BOX, tf((0.03, 0.0, 0.0)), (0.05, 0.05, 0.05), cfg)
    skew = (0.3, 0.5, 0.8)
    pax = (1.0, 0.0, 0.2)
    if kind == "single":
        specs = [((), ((0, 1, 0),)), ((), ((1, 0, 0),)), ((), (skew,)), ((pax,), ())]
    elif kind == "one_angle":
        specs = [(((0, 0, 1),), ((0, 1, 0),)), ((pax, (0, 1, 0)), ((1, 0, 0),)), ((), (skew,)),
                 (((1, 0, 0), (0, 1, 0), (0, 0, 1)), ())]
    elif kind == "mixed":
        specs = [((), tuple(tuple(c) for c in rng_axes.T)), ((pax,), ((1, 0, 0), (0, 0, 1))),
                 ((), ((0.2, 1.0, 0.1), (1.0, 0.3, -0.2))), ((), ((0, 1, 0),))]
    else:
        raise ValueError(kind)
    j0 = joint(-1, root, *specs[0], parent_xform=tf((0.0, 0.0, 1.5)))
    j1 = joint(root, arm, *specs[1], parent_xform=tf((0.0, 0.1, -0.5)), child_xform=tf((0.0, 0.0, 0.03)))
    j2 = joint(arm, hand, *specs[2], parent_xform=tf((0.0, 0.0, -0.42), quat_from_axis_angle((0, 0, 1), 0.4)))
    j3 = joint(root, slider, *specs[3], parent_xform=tf((0.1, -0.1, -0.3), quat_from_axis_angle((0, 1, 0), 0.5)))
    b.add_articulation([j0, j1, j2, j3])
    return b.finalize()


def single_world(model):
    assert model.world_count == 1
    return model


# ----------------------------------------------------------------------------
# coordinates along a velocity field, finite-difference body velocities
# ----------------------------------------------------------------------------
def quat_exp(theta):
    """Unit quaternion of the rotation vector ``theta`` (xyzw)."""
    ang = float(np.linalg.norm(theta))
    if ang < 1.0e-300:
        return np.array([0.0, 0.0, 0.0, 1.0])
    ax = np.asarray(theta, dtype=np.float64) / ang
    return np.concatenate([ax * math.sin(0.5 * ang), [math.cos(0.5 * ang)]])


def advance_coords(model, q, v_sap, t):
    """Coordinates at time ``t`` when the SAP velocity ``v_sap`` is held constant.

    Free joints move their origin with the linear part and rotate with the
    world-frame angular part (q(t) = exp(t w / 2) q); revolute / prismatic / D6
    coordinates move linearly; a ball joint turns its child frame with the
    angular velocity w given in the parent's joint frame (r(t) = exp(t w / 2) r,
    the rate the public FK maps to the world by the parent joint frame's
    rotation, ``sap_kinematics.py:203-206``).  The derivative at t = 0 is
    exactly ``v_sap``.
    """
    q = np.asarray(q, dtype=np.float64).copy()
    for j in range(model.joint_count):
        jt = int(model.joint_type[j])
        qs, ds = int(model.joint_q_start[j]), int(model.joint_qd_start[j])
        nd = int(model.joint_qd_start[j + 1]) - ds
        if jt in (SAP_JOINT_REVOLUTE, SAP_JOINT_PRISMATIC, SAP_JOINT_D6):
            q[qs:qs + nd] += t * v_sap[ds:ds + nd]
        elif jt == SAP_JOINT_FREE:
            q[qs:qs + 3] += t * v_sap[ds + 3:ds + 6]
            q[qs + 3:qs + 7] = quat_mul(quat_exp(t * v_sap[ds:ds + 3]), q[qs + 3:qs + 7])
        elif jt == SAP_JOINT_BALL:
            q[qs:qs + 4] = quat_mul(quat_exp(t * v_sap[ds:ds + 3]), q[qs:qs + 4])
    return q


def body_frames(model, q):
    """World COM positions (nb, 3) and rotation matrices (nb, 3, 3) from public FK of ``q``."""
    body_q, _ = eval_fk_public(model, q, np.zeros(model.joint_dof_count))
    R = np.stack([quat_to_matrix(x[3:] / np.linalg.norm(x[3:])) for x in body_q])
    com = body_q[:, :3] + np.einsum("bij,bj->bi", R, model.body_com.astype(np.float64))
    return body_q, com, R


def fd_body_velocities(model, q, v_sap, eps=1.0e-6):
    """(v_com (nb,3), w (nb,3)) by central differences along :func:`advance_coords`."""
    _, c_p, R_p = body_frames(model, advance_coords(model, q, v_sap, eps))
    _, c_m, R_m = body_frames(model, advance_coords(model, q, v_sap, -eps))
    _, _, R0 = body_frames(model, q)
    v_com = (c_p - c_m) / (2.0 * eps)
    Om = np.einsum("bij,bkj->bik", (R_p - R_m) / (2.0 * eps), R0)
    w = np.stack([0.5 * (Om[:, 2, 1] - Om[:, 1, 2]), 0.5 * (Om[:, 0, 2] - Om[:, 2, 0]),
                  0.5 * (Om[:, 1, 0] - Om[:, 0, 1])], axis=1)
    return v_com, w


def kinetic_energy(model, q, v_sap):
    """sum_b 1/2 m |v_com|^2 + 1/2 w^T (R I R^T) w from finite-difference body velocities."""
    v_com, w = fd_body_velocities(model, q, v_sap)
    _, _, R = body_frames(model, q)
    m = model.body_mass.astype(np.float64)
    I = model.body_inertia.astype(np.float64)
    T = 0.0
    for b in range(model.body_count):
        Iw = R[b] @ I[b] @ R[b].T
        T += 0.5 * m[b] * float(v_com[b] @ v_com[b]) + 0.5 * float(w[b] @ Iw @ w[b])
    return T


def public_velocity_from_sap(model, q, v_sap):
    """Public joint velocities: free joints carry [v_com, w]; others are unchanged."""
    body_q, _, R = body_frames(model, q)
    out = np.asarray(v_sap, dtype=np.float64).copy()
    for j in range(model.joint_count):
        if int(model.joint_type[j]) != SAP_JOINT_FREE:
            continue
        ds = int(model.joint_qd_start[j])
        child = int(model.joint_child[j])
        r = R[child] @ model.body_com[child].astype(np.float64)
        w = v_sap[ds:ds + 3]
        out[ds:ds + 3] = v_sap[ds + 3:ds + 6] + np.cross(w, r)
        out[ds + 3:ds + 6] = w
    return out

"""Pin the oracle's contact Jacobian, contact weights and public body velocities (CPU).

Each check is independent of the restated formula it tests:

* contact Jacobian (``contact_jacobian.py:404-562``): J_c v equals R_WC^T times the
  relative velocity of the two bodies' material points at the contact, taken by
  central differences of forward-kinematics positions along v — at the witness
  points (``witness_point``) and at the stiffness-weighted point
  (``contact_midpoint``), on the humanoid, a ball-jointed chain and free bodies
  with off-centre mass;
* ``body_inertia`` weight (``sap_helpers.py:307-338``, ``contact_jacobian.py:201-251``):
  for free bodies with diagonal body inertia it is the trace of the contact's
  Delassus block J_c A^-1 J_c^T over three, from the pinned J_c and the oracle's
  mass matrix; closed form 8 / (3 m) shape for a centred sphere;
* ``diag_delassus`` weight (``contact_jacobian.py:153-198``): the Frobenius norm of
  J_c diag(A)^-1 J_c^T over three with diag(A) built by hand from the
  parallel-axis theorem; closed form for a resting centred sphere;
* public ``body_qd`` (``sap_kinematics.py:84-229``): angular part = finite-difference
  rotation rate; linear part = finite-difference velocity of the centre of mass
  for free joints and of the child's material point at the joint anchor for
  rotational joints.

Sensitivity: each test also checks that a planted mistake (no contact-frame
rotation, lever arm from the body origin, no division by three) would fail.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import collide as oc
from oracle import sap_oracle as so
from sap_warp_amd.runtime import GEO_BOX, GEO_CAPSULE, GEO_SPHERE
from sap_warp_amd.scenes import (SceneBuilder, ShapeCfg, build_ball_chain_scene, build_humanoid_scene,
                                 quat_from_axis_angle, quat_mul, quat_to_matrix, tf)
from small_scenes import advance_coords, body_frames, public_velocity_from_sap

DT = 2.0e-3


# ----------------------------------------------------------------------------
# scenes (one world each, in contact)
# ----------------------------------------------------------------------------
def _humanoid_near_ground(seed):
    model = build_humanoid_scene(1, seed=seed)
    rng = np.random.default_rng(seed)
    q = model.joint_q.astype(np.float64).copy()
    for j in range(model.joint_count):
        qs = int(model.joint_q_start[j])
        if int(model.joint_type[j]) == so.JOINT_FREE:
            q[qs:qs + 3] = [0.1, -0.2, rng.uniform(0.25, 0.45)]
            q[qs + 3:qs + 7] = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.8, 1.6))
        elif int(model.joint_type[j]) == so.JOINT_REVOLUTE:
            q[qs] = rng.uniform(-1.2, 1.2)
    return model, q


def _ball_chain(seed):
    model = build_ball_chain_scene(1, seed=seed)
    return model, model.joint_q.astype(np.float64).copy()


def _lowest_z(pose, offset, half):
    R = quat_to_matrix(pose[3:])
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * half
    return float(np.min((pose[:3] + (R @ (np.asarray(offset) + corners).T).T)[:, 2]))


def _free_bodies_on_ground():
    """A tilted box on a corner, an axis-aligned box with a sphere on top and a lying capsule, every
    shape off its body origin (centres of mass off the origins); ground plane; free joints."""
    b = SceneBuilder(rigid_gap=0.01)
    cfg = ShapeCfg(density=800.0, ke=2.0e5, tau=0.005, mu=0.7, gap=0.01)
    b.world_count = 1
    b.current_world = 0
    joints = []
    # tilted box touching the ground with its lowest corner 2 mm deep
    off_a, half_a = np.array([0.04, -0.03, 0.02]), np.array([0.15, 0.1, 0.08])
    pose = tf((0.6, 0.0, 0.0), quat_mul(quat_from_axis_angle((1, 0, 0), 0.35), quat_from_axis_angle((0, 1, 0), 0.25)))
    pose[2] = -0.002 - _lowest_z(pose, off_a, half_a)
    ba = b.add_body(pose)
    b.add_shape(ba, GEO_BOX, tf(tuple(off_a)), tuple(half_a), cfg)
    joints.append(b.add_free_joint(ba))
    # axis-aligned box resting 2 mm deep
    off_b, half_b = np.array([-0.03, 0.02, 0.01]), np.array([0.12, 0.12, 0.06])
    bb = b.add_body(tf((0.0, 0.0, half_b[2] - off_b[2] - 0.002)))
    b.add_shape(bb, GEO_BOX, tf(tuple(off_b)), tuple(half_b), cfg)
    joints.append(b.add_free_joint(bb))
    # sphere on top of that box, 3 mm deep, centre off its body origin
    r_s, off_s = 0.07, np.array([0.02, -0.01, 0.0])
    top = 2.0 * half_b[2] - 0.002
    bs = b.add_body(tf((0.03 - off_s[0], -0.02 - off_s[1], top + r_s - 0.003 - off_s[2])))
    b.add_shape(bs, GEO_SPHERE, tf(tuple(off_s)), (r_s, 0.0, 0.0), cfg)
    joints.append(b.add_free_joint(bs))
    # capsule lying along x on the ground, 1.5 mm deep
    r_c, hh, off_c = 0.04, 0.12, np.array([0.0, 0.03, 0.0])
    q_c = quat_from_axis_angle((0, 1, 0), 0.5 * math.pi)
    bc = b.add_body(tf((-0.5, 0.4 - off_c[1], r_c - 0.0015)))
    b.add_shape(bc, GEO_CAPSULE, tf(tuple(off_c), q_c), (r_c, hh, 0.0), cfg)
    joints.append(b.add_free_joint(bc))
    for j in joints:
        b.add_articulation([j])
    b.current_world = -1
    b.add_ground_plane(cfg)
    model = b.finalize()
    return model, model.joint_q.astype(np.float64).copy()


def _centred_sphere(depth=0.003, r=0.1):
    b = SceneBuilder(rigid_gap=0.01)
    cfg = ShapeCfg(density=600.0, ke=2.0e5, tau=0.005, mu=0.5, gap=0.01)
    b.world_count = 1
    b.current_world = 0
    body = b.add_body(tf((0.0, 0.0, r - depth)))
    b.add_shape(body, GEO_SPHERE, tf(), (r, 0.0, 0.0), cfg)
    b.add_articulation([b.add_free_joint(body)])
    b.current_world = -1
    b.add_ground_plane(cfg)
    model = b.finalize()
    return model, model.joint_q.astype(np.float64).copy()


SCENES = {"humanoid": lambda: _humanoid_near_ground(3), "humanoid_b": lambda: _humanoid_near_ground(8),
          "ball_chain": lambda: _ball_chain(7), "free_bodies": _free_bodies_on_ground}


def _contacts(model, q, point_mode, weight_mode):
    wv = so.WorldView(model, 0)
    body_q, _, _ = body_frames(model, q)
    c = oc.collide(model, body_q.astype(np.float32), 64)
    fm = so.free_motion(model, wv, q, np.zeros(wv.n), np.zeros(wv.n), np.zeros((wv.nb, 6)), DT)
    cs = so.world_contacts(model, wv, fm, c, max_contacts=64, contact_point_mode=point_mode,
                           contact_weight_mode=weight_mode, fallback_mu=0.5, fallback_k=1.0e10, fallback_tau=0.0)
    assert cs.count > 0
    return wv, fm, cs


def _material_point_velocity(model, q, v, body, x, eps=1.0e-6):
    """Velocity of ``body``'s material point now at world ``x`` (central differences along v)."""
    if body < 0:
        return np.zeros(3)
    bq0, _, R0 = body_frames(model, q)
    p_body = R0[body].T @ (x - bq0[body, :3])
    pts = []
    for s in (1.0, -1.0):
        bq, _, R = body_frames(model, advance_coords(model, q, v, s * eps))
        pts.append(bq[body, :3] + R[body] @ p_body)
    return (pts[0] - pts[1]) / (2.0 * eps)


# ----------------------------------------------------------------------------
# contact Jacobian
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("scene", sorted(SCENES))
@pytest.mark.parametrize("point_mode", ["witness_point", "contact_midpoint"])
def test_contact_jacobian_is_relative_point_velocity(scene, point_mode):
    model, q = SCENES[scene]()
    wv, fm, cs = _contacts(model, q, point_mode, "body_inertia")
    rng = np.random.default_rng(11)
    planted_fails = False
    for trial in range(2):
        v = rng.normal(scale=1.0, size=wv.n)
        for c in range(cs.count):
            x0 = cs.x0[c] if point_mode == "witness_point" else cs.point[c]
            x1 = cs.x1[c] if point_mode == "witness_point" else cs.point[c]
            rel = (_material_point_velocity(model, q, v, int(cs.body1[c]), x1)
                   - _material_point_velocity(model, q, v, int(cs.body0[c]), x0))
            want = cs.R[c].T @ rel
            got = cs.J[c] @ v
            scale = max(1.0, float(np.linalg.norm(rel)))
            assert np.max(np.abs(got - want)) <= 1.0e-7 * scale, (scene, c, got, want)
            # planted mistake: relative velocity left in the world frame (no R_WC^T)
            planted_fails |= np.max(np.abs(rel - got)) > 1.0e-4 * scale
            # contact frame orthonormal with the normal as its last column
            np.testing.assert_allclose(cs.R[c].T @ cs.R[c], np.eye(3), atol=1e-12)
    assert planted_fails


# ----------------------------------------------------------------------------
# contact weights
# ----------------------------------------------------------------------------
def test_body_inertia_weight_is_delassus_trace_over_three():
    model, q = _free_bodies_on_ground()
    assert np.all(model.joint_armature == 0.0)
    wv, fm, cs = _contacts(model, q, "witness_point", "body_inertia")
    Ainv = np.linalg.inv(fm.A_env)
    pairs = set()
    for c in range(cs.count):
        W = cs.J[c] @ Ainv @ cs.J[c].T
        want = max(float(np.trace(W)) / 3.0, 1.0e-12)
        assert abs(cs.w[c] - want) <= 1.0e-10 * want, (c, cs.w[c], want)
        assert abs(float(np.trace(W)) - want) > 0.5 * want  # the /3 matters
        pairs.add(frozenset((int(cs.body0[c]), int(cs.body1[c]))))
    # both boxes and the capsule on the ground, the sphere on the second box
    assert {frozenset(p) for p in ((-1, 0), (-1, 1), (-1, 3), (1, 2))} <= pairs, pairs


def test_body_inertia_weight_lever_arm_from_centre_of_mass():
    """Planted mistake: the same weight with lever arms from the body origins is off by more than
    the tolerance above on the off-centre bodies."""
    model, q = _free_bodies_on_ground()
    wv, fm, cs = _contacts(model, q, "witness_point", "body_inertia")
    worst = 0.0
    for c in range(cs.count):
        tot = 0.0
        for b, x in ((int(cs.body0[c]), cs.x0[c]), (int(cs.body1[c]), cs.x1[c])):
            if b < 0:
                continue
            pose = fm.body_q[b]
            for k in range(3):
                tot += so._direction_weight(model, b, pose, x - pose[:3], cs.R[c][:, k])
        worst = max(worst, abs(tot / 3.0 - cs.w[c]) / cs.w[c])
    assert worst > 1.0e-3


def test_body_inertia_weight_centred_sphere_closed_form():
    model, q = _centred_sphere()
    wv, fm, cs = _contacts(model, q, "witness_point", "body_inertia")
    m = float(model.body_mass[0])
    I = float(model.body_inertia[0][0, 0])
    r = float(np.linalg.norm(cs.x1[0] - fm.body_q[0, :3]))
    want = (2.0 * (1.0 / m + r * r / I) + 1.0 / m) / 3.0
    assert abs(cs.w[0] - want) <= 1e-12 * want
    # a solid sphere: I = 2/5 m R^2, so at r = R the weight is 8 / (3 m)
    R = float(model.shape_scale[0][0])
    np.testing.assert_allclose(I, 0.4 * m * R * R, rtol=1e-6)


def _hand_diag_A(model, fm):
    """diag(A) of free bodies in SAP coordinates ([w, v_origin], world frame) by the parallel-axis
    theorem: (diag(R I_c R^T) + m (|c|^2 - c_k^2), m, m, m) with c = R com, plus armature."""
    d = np.zeros(model.joint_dof_count)
    for j in range(model.joint_count):
        assert int(model.joint_type[j]) == so.JOINT_FREE
        b = int(model.joint_child[j])
        ds = int(model.joint_qd_start[j])
        R = quat_to_matrix(fm.body_q[b, 3:] / np.linalg.norm(fm.body_q[b, 3:]))
        m = float(model.body_mass[b])
        c = R @ model.body_com[b].astype(np.float64)
        Iw = R @ model.body_inertia[b].astype(np.float64) @ R.T
        d[ds:ds + 3] = np.diag(Iw) + m * (float(c @ c) - c * c)
        d[ds + 3:ds + 6] = m
    return d + model.joint_armature.astype(np.float64)


@pytest.mark.parametrize("point_mode", ["witness_point", "contact_midpoint"])
def test_diag_delassus_weight_hand_built(point_mode):
    model, q = _free_bodies_on_ground()
    wv, fm, cs = _contacts(model, q, point_mode, "diag_delassus")
    dA = _hand_diag_A(model, fm)
    np.testing.assert_allclose(np.diag(fm.A_env), dA, rtol=1e-10)
    for c in range(cs.count):
        W = (cs.J[c] / dA) @ cs.J[c].T
        want = max(float(np.linalg.norm(W)) / 3.0, 1.0e-12)
        assert abs(cs.w[c] - want) <= 1.0e-10 * want, (c, cs.w[c], want)


def test_diag_delassus_weight_centred_sphere_closed_form():
    """Resting centred sphere, contact depth r below the centre: W = diag(r^2/I + 1/m, r^2/I + 1/m,
    1/m) in any tangent basis, so w = sqrt(2 (r^2/I + 1/m)^2 + 1/m^2) / 3."""
    model, q = _centred_sphere()
    for mode in ("witness_point", "contact_midpoint"):
        wv, fm, cs = _contacts(model, q, mode, "diag_delassus")
        m = float(model.body_mass[0])
        I = float(model.body_inertia[0][0, 0])
        x = cs.x1[0] if mode == "witness_point" else cs.point[0]
        r = float(np.linalg.norm(x - fm.body_q[0, :3]))
        t = r * r / I + 1.0 / m
        want = math.sqrt(2.0 * t * t + 1.0 / (m * m)) / 3.0
        assert abs(cs.w[0] - want) <= 1e-10 * want, (mode, cs.w[0], want)


# ----------------------------------------------------------------------------
# public body velocities
# ----------------------------------------------------------------------------
@pytest.mark.parametrize("scene", sorted(SCENES))
def test_public_body_qd_matches_finite_differences(scene):
    """Angular part: the finite-difference rotation rate.  Linear part: the reference shifts the
    parent's linear velocity by w_P x (anchor - c_P), the lever arm from the parent's centre of mass
    (``sap_kinematics.py:117-128``), while the linear velocity it carries for a rotational child is
    that of the joint anchor.  So L_B = v_B(c_B) + e_B with v_B(c_B) the finite-difference velocity of
    B's centre of mass and e_B = e_P + w_B x (anchor - c_B) accumulated down the tree (e = 0 below
    a free joint's body, e_P = 0 for the world); a free joint's body carries its centre of mass
    velocity exactly.  Everything on the right comes from finite differences of FK positions."""
    model, q = SCENES[scene]()
    wv = so.WorldView(model, 0)
    rng = np.random.default_rng(4)
    v = rng.normal(size=wv.n)
    qd_pub = public_velocity_from_sap(model, q, v)
    body_q, body_qd = so.public_fk(model, wv, q, qd_pub)
    eps = 1.0e-6
    bq_p, com_p, R_p = body_frames(model, advance_coords(model, q, v, eps))
    bq_m, com_m, R_m = body_frames(model, advance_coords(model, q, v, -eps))
    bq0, com0, R0 = body_frames(model, q)
    Om = np.einsum("bij,bkj->bik", (R_p - R_m) / (2.0 * eps), R0)
    w_fd = np.stack([0.5 * (Om[:, 2, 1] - Om[:, 1, 2]), 0.5 * (Om[:, 0, 2] - Om[:, 2, 0]),
                     0.5 * (Om[:, 1, 0] - Om[:, 0, 1])], axis=1)
    v_com_fd = (com_p - com_m) / (2.0 * eps)
    np.testing.assert_allclose(body_q[:, :3], bq0[:, :3], atol=1e-12)
    offset = np.zeros((model.body_count, 3))
    checked, deep = 0, 0
    for j in range(model.joint_count):  # parents before children
        b = int(model.joint_child[j])
        parent = int(model.joint_parent[j])
        jt = int(model.joint_type[j])
        scale = max(1.0, float(np.max(np.abs(body_qd[b]))))
        assert np.max(np.abs(body_qd[b, 3:] - w_fd[b])) <= 1e-7 * scale, (scene, b, body_qd[b, 3:], w_fd[b])
        if jt in (so.JOINT_FREE, so.JOINT_DISTANCE):
            offset[b] = 0.0 if parent < 0 else offset[parent]
        elif jt in (so.JOINT_REVOLUTE, so.JOINT_BALL):
            X_p = model.joint_X_p[j].astype(np.float64)
            anchor = X_p[:3] if parent < 0 else bq0[parent, :3] + R0[parent] @ X_p[:3]
            e_parent = 0.0 if parent < 0 else offset[parent]
            offset[b] = e_parent + np.cross(w_fd[b], anchor - com0[b])
            deep += int(parent >= 0 and np.linalg.norm(offset[parent]) > 1e-3)
        else:
            continue
        want = v_com_fd[b] + offset[b]
        assert np.max(np.abs(body_qd[b, :3] - want)) <= 1e-7 * scale, (scene, b, body_qd[b, :3], want)
        checked += 1
    assert checked == model.joint_count
    if scene.startswith("humanoid"):
        assert deep > 0  # bodies two rotational joints deep carry a nonzero accumulated offset

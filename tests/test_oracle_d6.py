"""Pin the oracle's D6 joints (sap_helpers.py:1109-1286, 1387-1432; sap_kinematics.py:170-218).

A D6 joint translates along up to three fixed linear axes and turns about up to three angular
axes; two or three angular axes are first replaced by the frame their matrix defines, then
applied as intrinsic rotations (each axis carried by the rotations before it).  Checks:

* the rotation equals scipy's intrinsic Euler sequence about that frame;
* body velocities from the motion subspace equal finite differences of forward-kinematics
  positions, and so do the public FK's angular velocities;
* kinetic energy 1/2 v^T H v against finite-difference body velocities;
* Lagrange's equations for D6 trees whose joints have at most one angular axis (every term
  from finite differences), and at rest for any D6 (multi-angle D6 joints keep the
  reference's bias w x (w x p) only, without the rate of change of the moving axes, so the
  equations hold exactly only where the velocity vanishes);
* a three-angle D6 at zero angles steps exactly as a ball joint at the identity (same motion
  subspace and bias there);
* one-axis D6 joints step exactly as revolute / prismatic joints, limits included.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.spatial.transform import Rotation

from oracle import collide as oc
from oracle import cpu_step as cs
from oracle import sap_oracle as so
from sap_warp_amd.kinematics_host import d6_joint_motion, eval_fk_public
from sap_warp_amd.runtime import GEO_BOX, SAP_JOINT_BALL, SAP_JOINT_D6
from sap_warp_amd.scenes import DofCfg, SceneBuilder, ShapeCfg, quat_from_axis_angle, tf
from small_scenes import advance_coords, body_frames, d6_tree_scene, fd_body_velocities, kinetic_energy


def _free_motion(model, q, v, jf=None, body_f=None, dt=1.0e-3):
    wv = so.WorldView(model, 0)
    f = np.zeros(model.joint_dof_count) if jf is None else jf
    body_q, _, _ = body_frames(model, q)
    bf = so.public_body_force_to_sap(model, wv, body_q, body_f)
    return so.free_motion(model, wv, q, np.asarray(v, np.float64), np.asarray(f, np.float64), bf, dt)


def _random_state(model, rng, q_scale=0.6, v_scale=1.5):
    q = model.joint_q.astype(np.float64) + rng.uniform(-q_scale, q_scale, size=model.joint_coord_count)
    v = rng.normal(scale=v_scale, size=model.joint_dof_count)
    return q, v


# ----------------------------------------------------------------------------
@pytest.mark.parametrize("na", [2, 3])
def test_rotation_is_intrinsic_euler_about_the_axis_frame(na):
    rng = np.random.default_rng(na)
    R_off = Rotation.from_rotvec(rng.normal(size=3)).as_matrix()
    axes = [R_off[:, k] for k in range(na)]
    for _ in range(5):
        ang = rng.uniform(-2.5, 2.5, size=na)
        rot, cur = so.d6_rotation(axes, ang)
        R_exp = R_off @ Rotation.from_euler("XYZ"[:na], ang).as_matrix() @ R_off.T
        np.testing.assert_allclose(Rotation.from_quat(rot).as_matrix(), R_exp, atol=1e-12)
        X, _ = d6_joint_motion([], axes, ang[None, :], np.zeros((1, na)))
        np.testing.assert_allclose(Rotation.from_quat(X[0, 3:]).as_matrix(), R_exp, atol=1e-12)
        # the current axes: R_0 ... R_{k-1} l_k
        for k in range(na):
            R_pre = R_off @ Rotation.from_euler("XYZ"[:k], ang[:k]).as_matrix().reshape(3, 3) @ R_off.T if k else np.eye(3)
            np.testing.assert_allclose(cur[k], R_pre @ R_off[:, k], atol=1e-12)


def test_single_axis_rotation_and_skew_frames():
    ax = np.array([0.3, 0.5, 0.8]) / np.linalg.norm([0.3, 0.5, 0.8])
    rot, cur = so.d6_rotation([ax], [0.7])
    np.testing.assert_allclose(Rotation.from_quat(rot).as_rotvec(), 0.7 * ax, atol=1e-12)
    # two non-orthogonal axes: the joint turns about an orthonormal frame (l0, l1) built from them
    a0 = np.array([0.2, 1.0, 0.1]) / np.linalg.norm([0.2, 1.0, 0.1])
    a1 = np.array([1.0, 0.3, -0.2]) / np.linalg.norm([1.0, 0.3, -0.2])
    _, l = so.d6_rotation([a0, a1], [0.0, 0.0])
    assert abs(np.dot(l[0], l[1])) < 1e-12
    np.testing.assert_allclose([np.linalg.norm(l[0]), np.linalg.norm(l[1])], 1.0, atol=1e-12)
    X, _ = d6_joint_motion([], [a0, a1], np.array([[0.4, -1.1]]), np.zeros((1, 2)))
    rot, _ = so.d6_rotation([a0, a1], [0.4, -1.1])
    assert abs(abs(np.dot(X[0, 3:], rot)) - 1.0) < 1e-12


@pytest.mark.parametrize("kind", ["one_angle", "mixed"])
def test_velocities_match_finite_differences(kind):
    model = d6_tree_scene(kind)
    rng = np.random.default_rng(7)
    wv = so.WorldView(model, 0)
    com = model.body_com.astype(np.float64)
    for _ in range(3):
        q, v = _random_state(model, rng)
        fm = _free_motion(model, q, v)
        v_com, w = fd_body_velocities(model, q, v)
        for b in range(model.body_count):
            r = so.qrot(fm.body_q[b, 3:], com[b])
            np.testing.assert_allclose(fm.V[b, :3], w[b], atol=2e-7)
            np.testing.assert_allclose(fm.V[b, 3:] + np.cross(fm.V[b, :3], r), v_com[b], atol=2e-7)
        # public FK: the angular part is the body's angular velocity; the linear part is the
        # reference's joint-anchor velocity (sap_kinematics.py:117-128), checked against the host FK
        bq, bqd = so.public_fk(model, wv, q, v)
        bq_h, bqd_h = eval_fk_public(model, q, v)
        np.testing.assert_allclose(bqd[:, 3:], w, atol=2e-7)
        np.testing.assert_allclose(bq[:, :3], bq_h[:, :3], atol=1e-12)
        np.testing.assert_allclose(np.abs(np.sum(bq[:, 3:] * bq_h[:, 3:], axis=1)), 1.0, atol=1e-12)
        np.testing.assert_allclose(bqd, bqd_h, atol=1e-12)


@pytest.mark.parametrize("kind", ["one_angle", "mixed"])
def test_kinetic_energy(kind):
    model = d6_tree_scene(kind)
    rng = np.random.default_rng(3)
    for _ in range(3):
        q, v = _random_state(model, rng)
        fm = _free_motion(model, q, v)
        T_fd = kinetic_energy(model, q, v)
        assert abs(0.5 * float(v @ fm.H @ v) - T_fd) <= 1e-7 * max(1.0, T_fd)
        assert np.min(np.linalg.eigvalsh(fm.H)) > 0.0


def _lagrange_residual(model, q, qd, rng):
    n = model.joint_dof_count
    eps = 1.0e-5
    g = model.gravity[0].astype(np.float64)
    m = model.body_mass.astype(np.float64)
    jf = rng.normal(scale=2.0, size=n)
    body_f = np.zeros((model.body_count, 6))
    body_f[2] = rng.normal(scale=3.0, size=6)
    fm = _free_motion(model, q, qd, jf=jf, body_f=body_f)
    dH = np.zeros((n, n, n))
    dc = np.zeros((n, model.body_count, 3))
    dth = np.zeros((n, model.body_count, 3))
    for k in range(n):
        e = np.zeros(n)
        e[k] = eps
        dH[k] = (_free_motion(model, q + e, qd).H - _free_motion(model, q - e, qd).H) / (2 * eps)
        dc[k], dth[k] = fd_body_velocities(model, q, e / eps, eps=eps)
    C = np.einsum("kij,k->ij", dH, qd) @ qd - 0.5 * np.einsum("kij,i,j->k", dH, qd, qd)
    Q_g = np.einsum("kbx,b,x->k", dc, m, g)
    Q_ext = np.einsum("kbx,bx->k", dc, body_f[:, :3]) + np.einsum("kbx,bx->k", dth, body_f[:, 3:])
    rhs = jf + Q_g + Q_ext - C
    lhs = fm.A_env @ fm.qdd
    return np.max(np.abs(lhs - rhs)) / max(1.0, float(np.max(np.abs(rhs))))


def test_lagrange_equations_one_angle_d6():
    model = d6_tree_scene("one_angle")
    rng = np.random.default_rng(5)
    for _ in range(2):
        q, qd = _random_state(model, rng)
        assert _lagrange_residual(model, q, qd, rng) <= 2e-6


def test_lagrange_equations_at_rest_multi_angle_d6():
    model = d6_tree_scene("mixed")
    rng = np.random.default_rng(6)
    for _ in range(2):
        q, _ = _random_state(model, rng)
        assert _lagrange_residual(model, q, np.zeros(model.joint_dof_count), rng) <= 2e-6


def _pendulum(jt):
    b = SceneBuilder(gravity=-9.81)
    body = b.add_body(tf())
    b.add_shape(body, GEO_BOX, tf((0.1, -0.05, -0.3)), (0.05, 0.08, 0.2), ShapeCfg(density=800.0))
    axes = [DofCfg(axis=a, limit_ke=0.0) for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
    j = b.add_joint(-1, body, jt, angular_axes=axes, parent_xform=tf((0, 0, 1.0), quat_from_axis_angle((1, 1, 0), 0.4)),
                    child_xform=tf((0.0, 0.02, 0.05)))
    b.add_articulation([j])
    return b.finalize()


def test_three_angle_d6_at_zero_matches_ball():
    d6, ball = _pendulum(SAP_JOINT_D6), _pendulum(SAP_JOINT_BALL)
    assert d6.joint_coord_count == 3 and ball.joint_coord_count == 4
    rng = np.random.default_rng(9)
    for _ in range(3):
        v = rng.normal(scale=2.0, size=3)
        jf = rng.normal(size=3)
        fa = _free_motion(d6, np.zeros(3), v, jf=jf)
        fb = _free_motion(ball, np.array([0.0, 0.0, 0.0, 1.0]), v, jf=jf)
        np.testing.assert_allclose(fa.H, fb.H, atol=1e-12)
        np.testing.assert_allclose(fa.v_star, fb.v_star, atol=1e-12)


def _step_all(model, q0, qd0, steps=6, seed=0):
    rng = np.random.default_rng(seed)
    n = model.joint_dof_count
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=8)
    state = dict(joint_q=q0.astype(np.float32), joint_qd=qd0.astype(np.float32),
                 body_q=eval_fk_public(model, q0, qd0)[0].astype(np.float32),
                 body_qd=eval_fk_public(model, q0, qd0)[1].astype(np.float32), body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=rng.normal(size=n).astype(np.float32), joint_target_pos=model.joint_target_pos,
                   joint_target_vel=model.joint_target_vel, joint_act=model.joint_act)
    outs = []
    for _ in range(steps):
        out = so.step(model, state, control, _contacts(model, state["body_q"]), 1.0e-3, opts)
        outs.append(out)
        state = dict(joint_q=out["joint_q"], joint_qd=out["joint_qd"], body_q=out["body_q"], body_qd=out["body_qd"],
                     body_f=state["body_f"])
    return outs


def _contacts(model, body_q):
    return oc.collide(model, np.asarray(body_q, np.float32), 8)


@pytest.mark.parametrize("limits", [False, True])
def test_one_axis_d6_steps_as_revolute_and_prismatic(limits):
    d6 = d6_tree_scene("single", limits=limits)
    plain = d6_tree_scene("single", as_plain=True, limits=limits)
    # every coordinate starts next to a limit and moves into it
    q0 = np.array([0.099, -0.049, 0.0995, -0.0495])
    qd0 = np.array([2.0, -2.0, 1.5, -1.5])
    a = _step_all(d6, q0, qd0)
    b = _step_all(plain, q0, qd0)
    for oa, ob in zip(a, b):
        for k in ("joint_q", "joint_qd", "body_q", "body_qd"):
            np.testing.assert_allclose(oa[k], ob[k], rtol=0, atol=2e-6, err_msg=k)
    if limits:  # the limits act: the free tree runs past them
        free = _step_all(d6_tree_scene("single"), q0, qd0)
        assert np.all(a[-1]["joint_q"][[0, 2]] < free[-1]["joint_q"][[0, 2]] - 1e-4)
        assert np.all(a[-1]["joint_q"][[1, 3]] > free[-1]["joint_q"][[1, 3]] + 1e-4)


def test_cpu_step_matches_oracle_on_d6():
    """The C++ restatement of the step agrees with the numpy oracle on D6 trees."""
    model = d6_tree_scene("mixed", limits=True)
    rng = np.random.default_rng(12)
    q0 = rng.uniform(-0.04, 0.09, size=model.joint_coord_count)
    qd0 = rng.normal(size=model.joint_dof_count)
    bq, bqd = eval_fk_public(model, q0, qd0)
    state = dict(joint_q=q0.astype(np.float32), joint_qd=qd0.astype(np.float32), body_q=bq.astype(np.float32),
                 body_qd=bqd.astype(np.float32), body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=rng.normal(size=model.joint_dof_count).astype(np.float32),
                   joint_target_pos=model.joint_target_pos, joint_target_vel=model.joint_target_vel,
                   joint_act=model.joint_act)
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=8)
    c = _contacts(model, state["body_q"])
    ref = so.step(model, state, control, c, 1.0e-3, opts)
    out = cs.CpuStepper(model, opts).step(state, control, c, 1.0e-3, threads=1)
    for k in ("joint_q", "joint_qd", "body_q", "body_qd"):
        np.testing.assert_allclose(out[k], ref[k], rtol=0, atol=2e-6, err_msg=k)
    np.testing.assert_allclose(out["v_sap"], ref["v_sap"], rtol=0, atol=1e-9)


def test_advance_coords_moves_d6_coordinates_linearly():
    model = d6_tree_scene("mixed")
    q = model.joint_q.astype(np.float64)
    v = np.arange(model.joint_dof_count, dtype=np.float64)
    np.testing.assert_allclose(advance_coords(model, q, v, 0.5), q + 0.5 * v)


def test_d6_ant_cpu_step_matches_oracle_in_contact():
    """The D6 ant (free torso, every D6 kind, limits and drives) after 100 steps of falling, in
    contact with the ground: the C++ step and the numpy oracle agree step by step."""
    from sap_warp_amd.scenes import build_d6_ant_scene

    model = build_d6_ant_scene(4, seed=13)
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=16)
    stepper = cs.CpuStepper(model, opts)
    state = dict(joint_q=model.joint_q, joint_qd=model.joint_qd, body_q=model.body_q, body_qd=model.body_qd,
                 body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=model.joint_f, joint_target_pos=model.joint_target_pos,
                   joint_target_vel=model.joint_target_vel, joint_act=model.joint_act)
    ncon = 0
    for k in range(103):
        c = oc.collide(model, state["body_q"], 16)
        out = stepper.step(state, control, c, 2.0e-3, threads=4)
        if k >= 100:
            ref = so.step(model, state, control, c, 2.0e-3, opts)
            ncon += int(c["world_start"][-1])
            for key in ("joint_q", "joint_qd", "body_q", "body_qd"):
                np.testing.assert_allclose(out[key], ref[key], rtol=0, atol=5e-6, err_msg=f"step {k} {key}")
            np.testing.assert_allclose(out["v_sap"], ref["v_sap"], rtol=0, atol=1e-7)
        state = dict(joint_q=out["joint_q"], joint_qd=out["joint_qd"], body_q=out["body_q"], body_qd=out["body_qd"],
                     body_f=state["body_f"])
    assert ncon >= 3

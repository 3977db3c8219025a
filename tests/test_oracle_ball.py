"""Ball joints in the oracle: closed-form free motion about a fixed point, and agreement with a
revolute joint when the motion stays in one plane."""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import sap_oracle as so
from sap_warp_amd.runtime import GEO_BOX, GEO_CAPSULE, SAP_JOINT_BALL, SAP_JOINT_REVOLUTE
from sap_warp_amd.scenes import DofCfg, SceneBuilder, ShapeCfg, _with_consistent_body_q, quat_from_axis_angle, tf
from tests.helpers import control_dict_from_model, state_dict_from_model


def qmat(q):
    x, y, z, w = q
    return np.array([[1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
                     [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
                     [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)]])


def pendulum(kind, q_joint, qd_joint, pivot=(0.1, -0.2, 1.5), symmetric=False):
    """One capsule link hanging from a world pivot by a ball or a revolute (about x) joint; no
    ground.  The general link has its axis tilted in its body frame and its COM off the joint's
    plane; the symmetric one hangs straight below the joint (x is a principal axis through it)."""
    b = SceneBuilder(gravity=-9.81)
    body = b.add_body(tf(pivot))
    if symmetric:
        b.add_shape(body, GEO_CAPSULE, tf((0.0, 0.0, -0.35)), (0.05, 0.25, 0.0), ShapeCfg(density=900.0))
    else:
        b.add_shape(body, GEO_CAPSULE, tf((0.03, 0.0, -0.35), quat_from_axis_angle((0, 1, 0), 0.2)),
                    (0.05, 0.25, 0.0), ShapeCfg(density=900.0))
    if kind == "ball":
        axes = [DofCfg(axis=a, limit_ke=0.0, limit_kd=0.0) for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
        j = b.add_joint(-1, body, SAP_JOINT_BALL, angular_axes=axes, parent_xform=tf(pivot))
    else:
        j = b.add_joint(-1, body, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=(1, 0, 0), limit_ke=0.0, limit_kd=0.0)],
                        parent_xform=tf(pivot))
    b.add_articulation([j])
    qs, ds = b.joint_q_start[j], b.joint_qd_start[j]
    b.joint_q[qs:qs + len(q_joint)] = list(q_joint)
    b.joint_qd[ds:ds + len(qd_joint)] = list(qd_joint)
    return _with_consistent_body_q(b.finalize())


def _step(m, steps, dt, integration="midpoint"):
    st = state_dict_from_model(m)
    ctl = control_dict_from_model(m)
    opts = so.StepOptions(preset="approx32" if integration == "midpoint" else "drake", line_search="exact_root",
                          max_rigid_contact=8)
    empty = dict(count=0, world_start=np.zeros(2, np.int32), shape0=np.zeros(0, np.int32),
                 shape1=np.zeros(0, np.int32), point0=np.zeros((0, 3), np.float32),
                 point1=np.zeros((0, 3), np.float32), normal=np.zeros((0, 3), np.float32),
                 margin0=np.zeros(0, np.float32), margin1=np.zeros(0, np.float32))
    out = None
    for _ in range(steps):
        out = so.step(m, st, ctl, empty, dt, opts)
        st = dict(joint_q=out["joint_q"].astype(np.float32), joint_qd=out["joint_qd"].astype(np.float32),
                  body_q=out["body_q"].astype(np.float32), body_qd=out["body_qd"].astype(np.float32),
                  body_f=st["body_f"])
    return out


@pytest.mark.parametrize("seed", range(4))
def test_ball_free_motion_closed_form(seed):
    """v* of a body on a ball joint at a world pivot: I_p w' + w x I_p w = r x m g (Euler's equations
    about the fixed point, world frame = joint frame)."""
    rng = np.random.default_rng(seed)
    axis = rng.normal(size=3)
    q = quat_from_axis_angle(axis, rng.uniform(0.2, 2.5))
    w0 = rng.normal(0.0, 1.5, 3)
    m = pendulum("ball", q, w0)
    dt = 1.0e-3
    out = _step(m, 1, dt)
    v_star = out["per_world"][0]["v_star"]
    mass = float(m.body_mass[0])
    R = qmat(m.body_q[0][3:].astype(np.float64))
    r = R @ m.body_com[0].astype(np.float64)
    I_c = R @ m.body_inertia[0].astype(np.float64) @ R.T
    I_p = I_c + mass * (np.dot(r, r) * np.eye(3) - np.outer(r, r))
    tau = np.cross(r, mass * np.array([0.0, 0.0, -9.81]))
    expected = w0 + dt * np.linalg.solve(I_p, tau - np.cross(w0, I_p @ w0))
    np.testing.assert_allclose(v_star, expected, rtol=2e-5, atol=1e-7)


@pytest.mark.parametrize("integration", ["midpoint", "sap_euler"])
def test_ball_matches_revolute_in_plane(integration):
    """A pendulum symmetric about its swing plane, started in it, swings on a ball joint exactly
    like on a revolute joint (no torque or gyroscopic coupling leaves the plane)."""
    th, om = 0.7, -1.3
    mb = pendulum("ball", quat_from_axis_angle((1, 0, 0), th), (om, 0.0, 0.0), symmetric=True)
    mr = pendulum("rev", [th], [om], symmetric=True)
    ob = _step(mb, 40, 2.0e-3, integration)
    orv = _step(mr, 40, 2.0e-3, integration)
    np.testing.assert_allclose(ob["body_q"], orv["body_q"], atol=2e-6)
    np.testing.assert_allclose(ob["body_qd"], orv["body_qd"], atol=2e-5)
    np.testing.assert_allclose(ob["joint_qd"][0], orv["joint_qd"][0], atol=2e-5)
    np.testing.assert_allclose(ob["joint_qd"][1:], 0.0, atol=1e-9)
    angle = 2.0 * math.atan2(ob["joint_q"][0], ob["joint_q"][3])
    assert angle == pytest.approx(float(orv["joint_q"][0]), abs=2e-6)


def test_ball_chain_scene_starts_in_light_contact():
    """The ball-chain scene hangs each chain just above its box: every world's contacts start at
    most about 1 cm deep (a chain hung lower starts several cm inside the box and its first step
    is an ejection, not a contact), the chain touches the box in some worlds, and one oracle
    step converges everywhere."""
    from oracle.collide import collide
    from sap_warp_amd.scenes import build_ball_chain_scene

    m = build_ball_chain_scene(8, seed=7)
    assert m.joint_dof_count == 8 * 18
    c = collide(m, m.body_q, 64)
    assert c["unsupported"] == 0
    bq = np.asarray(m.body_q, dtype=np.float64)
    depths, chain_box = [], 0
    for k in range(int(c["count"])):
        p = []
        for side in (0, 1):
            s = int(c[f"shape{side}"][k])
            b = int(m.shape_body[s])
            x = np.asarray(c[f"point{side}"][k], dtype=np.float64)
            if b >= 0:
                x = bq[b, :3] + qmat(bq[b, 3:7]) @ x
            p.append(x)
        d = float(np.dot(p[1] - p[0], c["normal"][k]) - c["margin0"][k] - c["margin1"][k])
        depths.append(d)
        types = {int(m.shape_type[int(c["shape0"][k])]), int(m.shape_type[int(c["shape1"][k])])}
        chain_box += int(types == {GEO_CAPSULE, GEO_BOX})
    assert min(depths) > -0.012, min(depths)
    assert chain_box >= 2
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=64)
    out = so.step(m, state_dict_from_model(m), control_dict_from_model(m), c, 2.0e-3, opts)
    assert all(r["result"].converged for r in out["per_world"].values())

"""Pin the learned-warm-start oracle (features, output convention, MLP, Adam) on the CPU.

* bfloat16 rounding vs torch's own float32 -> bfloat16 conversion (ties to even).
* MLP backward vs central finite differences of the loss (float64, rounding off),
  for weights, biases and a padded input / output layout.
* Adam's first step has the closed form p - lr * g / (|g| + eps).
* Features are invariant under a horizontal translation and a yaw rotation of a
  whole world (contacts recomputed by the collision oracle on the moved poses).
* local_delta / apply_delta invert each other; a pure yaw leaves local rows alone.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import collide as oc
from oracle import learned_oracle as lo
from sap_warp_amd.scenes import build_ant_scene, build_humanoid_scene
from tests.helpers import control_dict_from_model, state_dict_from_model


def test_bf16_round_matches_torch():
    import torch

    rng = np.random.default_rng(3)
    x = np.concatenate([rng.standard_normal(4000).astype(np.float32) * 10.0 ** rng.integers(-6, 6, 4000),
                        np.array([1.0, 1.0 + 2.0 ** -8, 1.0 + 3 * 2.0 ** -8, -2.5, 0.0, 65504.0], np.float32)])
    ref = torch.from_numpy(x).to(torch.bfloat16).to(torch.float32).numpy()
    np.testing.assert_array_equal(lo.bf16_round(x), ref)
    assert lo.bf16_round(np.float32(1.0 + 2.0 ** -8)) == 1.0  # tie -> even


def _tiny_mlp(rounding=False, seed=0):
    m = lo.init_mlp([32, 64, 64, 64], in_features=20, out_features=7, seed=seed, out_scale=1.0)
    rng = np.random.default_rng(seed + 1)
    m.b = [rng.standard_normal(b.shape) * 0.1 for b in m.b]
    m.b[-1][m.out_features:] = 0.0
    m.x_mean = rng.standard_normal(20)
    m.x_std = 0.5 + rng.random(20)
    m.y_mean = rng.standard_normal(7)
    m.y_std = 0.5 + rng.random(7)
    m.rounding = rounding
    return m


def test_mlp_backward_matches_finite_differences():
    m = _tiny_mlp(rounding=False)
    rng = np.random.default_rng(5)
    x = rng.standard_normal((9, 20))
    t = rng.standard_normal((9, 7))
    loss, gW, gb = m.loss_and_grads(x, t)
    assert loss > 0.0
    h = 1e-6
    for l in range(len(m.W)):
        for _ in range(6):
            i = rng.integers(0, m.W[l].shape[0] if l < len(m.W) - 1 else m.out_features)
            j = rng.integers(0, m.in_features if l == 0 else m.W[l].shape[1])
            old = m.W[l][i, j]
            m.W[l][i, j] = old + h
            lp = m.loss_and_grads(x, t)[0]
            m.W[l][i, j] = old - h
            lm = m.loss_and_grads(x, t)[0]
            m.W[l][i, j] = old
            fd = (lp - lm) / (2 * h)
            assert abs(fd - gW[l][i, j]) <= 1e-6 * max(1.0, abs(fd)), (l, i, j, fd, gW[l][i, j])
        i = rng.integers(0, m.out_features if l == len(m.W) - 1 else m.b[l].shape[0])
        old = m.b[l][i]
        m.b[l][i] = old + h
        lp = m.loss_and_grads(x, t)[0]
        m.b[l][i] = old - h
        lm = m.loss_and_grads(x, t)[0]
        m.b[l][i] = old
        assert abs((lp - lm) / (2 * h) - gb[l][i]) <= 1e-6
    # padded input columns and padded output rows carry no gradient
    assert np.all(gW[0][:, m.in_features:] == 0.0)
    assert np.all(gW[-1][m.out_features:, :] == 0.0)


def test_mlp_rounding_points_change_little():
    m = _tiny_mlp(rounding=True)
    m64 = _tiny_mlp(rounding=False)
    rng = np.random.default_rng(6)
    x = rng.standard_normal((16, 20))
    y, _ = m.forward(x)
    y64, _ = m64.forward(x)
    assert np.max(np.abs(y - y64)) <= 0.05 * max(1.0, np.max(np.abs(y64)))
    assert np.max(np.abs(y - y64)) > 0.0


def test_adam_first_step_closed_form():
    rng = np.random.default_rng(7)
    p = [rng.standard_normal((5, 4)), rng.standard_normal(3)]
    g = [rng.standard_normal((5, 4)), rng.standard_normal(3)]
    p0 = [a.copy() for a in p]
    opt = lo.Adam(lr=1e-2)
    opt.step(p, g)
    for a, a0, gg in zip(p, p0, g):
        np.testing.assert_allclose(a, a0 - 1e-2 * gg / (np.abs(gg) + 1e-8), rtol=1e-12, atol=1e-15)


def _yaw_quat(theta):
    return np.array([0.0, 0.0, np.sin(theta / 2), np.cos(theta / 2)])


def _qmul(a, b):
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array([aw * bx + ax * bw + ay * bz - az * by, aw * by - ax * bz + ay * bw + az * bx,
                     aw * bz + ax * by - ay * bx + az * bw, aw * bw - ax * bx - ay * by - az * bz])


def _move_world(model, st, w, shift, theta):
    """Rigidly move world w: rotate by yaw theta about the world origin, then translate by shift."""
    out = {k: np.array(v, dtype=np.float64, copy=True) for k, v in st.items()}
    Rz = lo._quat_matrix(_yaw_quat(theta))
    qz = _yaw_quat(theta)
    nb = model.bodies_per_env
    for b in range(w * nb, (w + 1) * nb):
        out["body_q"][b, :3] = Rz @ out["body_q"][b, :3] + shift
        out["body_q"][b, 3:7] = _qmul(qz, out["body_q"][b, 3:7])
    jpw = model.joint_count // model.world_count
    for j in range(w * jpw, (w + 1) * jpw):
        if int(model.joint_type[j]) in (lo.JOINT_FREE, lo.JOINT_DISTANCE):
            qs, ds = int(model.joint_q_start[j]), int(model.joint_qd_start[j])
            out["joint_q"][qs:qs + 3] = Rz @ out["joint_q"][qs:qs + 3] + shift
            out["joint_q"][qs + 3:qs + 7] = _qmul(qz, out["joint_q"][qs + 3:qs + 7])
            out["joint_qd"][ds:ds + 3] = Rz @ out["joint_qd"][ds:ds + 3]
            out["joint_qd"][ds + 3:ds + 6] = Rz @ out["joint_qd"][ds + 3:ds + 6]
    return out


@pytest.mark.parametrize("scene", ["ant", "humanoid"])
def test_features_invariant_under_translation_and_yaw(scene):
    model = build_ant_scene(4, seed=3) if scene == "ant" else build_humanoid_scene(3, seed=4)
    st = state_dict_from_model(model)
    ctl = control_dict_from_model(model)
    rng = np.random.default_rng(11)
    ctl["joint_target_pos"] = ctl["joint_target_pos"] + rng.standard_normal(ctl["joint_target_pos"].shape) * 0.1
    c = oc.collide(model, st["body_q"], 64)
    F = lo.feature_count(model)
    for w in range(model.world_count):
        f = lo.features(model, w, st, ctl, c)
        assert f.shape == (F,)
        assert np.all(np.isfinite(f))
        moved = _move_world(model, st, w, np.array([0.7, -1.3, 0.0]), 0.9)
        c2 = oc.collide(model, moved["body_q"].astype(np.float32), 64)
        f2 = lo.features(model, w, moved, ctl, c2)
        np.testing.assert_allclose(f2, f, atol=2e-5, rtol=1e-5)
    # the per-body contact counts add up to the contacts that touch a body (twice for body-body pairs)
    nb, n = model.bodies_per_env, model.dof_per_env
    counts = np.concatenate([lo.features(model, w, st, ctl, c)[-5 * nb - n:-n:5] for w in range(model.world_count)])
    ncon = int(np.asarray(c["count"]).ravel()[0])
    on_bodies = sum(int(c["shape0"][k] >= 0 and model.shape_body[c["shape0"][k]] >= 0) +
                    int(c["shape1"][k] >= 0 and model.shape_body[c["shape1"][k]] >= 0) for k in range(ncon))
    if scene == "humanoid":
        assert ncon > 0  # standing humanoids touch the ground
    assert counts.sum() == on_bodies


def test_delta_round_trip():
    model = build_humanoid_scene(2, seed=5)
    rng = np.random.default_rng(12)
    n = model.dof_per_env
    for w in range(2):
        jq = model.joint_q.astype(np.float64)
        v_plain = rng.standard_normal(n)
        v = rng.standard_normal(n)
        d = lo.local_delta(model, w, jq, v - v_plain)
        np.testing.assert_allclose(lo.apply_delta(model, w, jq, v_plain, d), v, atol=1e-12)
        # non-free rows are untouched
        np.testing.assert_allclose(d[6:], (v - v_plain)[6:], atol=0)
        # the free rows keep their norms
        assert abs(np.linalg.norm(d[:3]) - np.linalg.norm((v - v_plain)[:3])) < 1e-12

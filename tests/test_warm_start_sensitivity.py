"""How close a warm start must be to cut Newton iterations (CPU, the C++ restatement of the step).

On the humanoid workload (PD targets on a random walk, approx32, exact_root), every world-step
that enters Newton is solved again from v(eps) = v* + eps (v_prev - v*), where v* is the step's
own solution and v_prev the previous step's solution (the plain warm start, contact_solve.py:
316-393).  The mean iteration count falls slowly with eps, as Newton's local quadratic
convergence predicts (iterations ~ log log 1/eps): a 40 % cut needs a start several times closer
to v* than v_prev is.  A learned start that reduces iterations by a few per cent is therefore
only marginally closer to v* than the previous solution.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import collide as oc
from oracle import cpu_step as cs
from oracle import sap_oracle as so
from sap_warp_amd.scenes import build_humanoid_scene

EPS = (0.0, 0.05, 0.2, 0.3, 0.5, 1.0)


@pytest.fixture(scope="module")
def curve():
    W, C, dt, steps = 64, 64, 2.0e-3, 40
    model = build_humanoid_scene(W, seed=2)
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=C)
    stepper = cs.CpuStepper(model, opts)
    n = model.joint_dof_count // W
    mask = np.zeros(n, np.float32)
    mask[6:] = 1.0
    rng = np.random.default_rng(2)
    state = dict(joint_q=model.joint_q, joint_qd=model.joint_qd, body_q=model.body_q, body_qd=model.body_qd,
                 body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=model.joint_f, joint_target_pos=model.joint_target_pos.copy(),
                   joint_target_vel=model.joint_target_vel, joint_act=model.joint_act)
    acc = {e: [] for e in EPS}
    plain = []
    guess = None
    for k in range(steps):
        control["joint_target_pos"] = control["joint_target_pos"] + (
            rng.standard_normal((W, n)).astype(np.float32) * 0.05 * mask).reshape(-1)
        c = oc.collide(model, state["body_q"], C, threads=4)
        out = stepper.step(state, control, c, dt, v_guess=guess, threads=4)
        if k >= 10 and guess is not None:
            active = out["iterations"] > 0
            plain.append(out["iterations"][active])
            for e in EPS:
                o2 = stepper.step(state, control, c, dt, v_guess=out["v_sap"] + e * (guess - out["v_sap"]), threads=4)
                acc[e].append(o2["iterations"][active])
        guess = out["v_sap"]
        state = dict(joint_q=out["joint_q"], joint_qd=out["joint_qd"], body_q=out["body_q"], body_qd=out["body_qd"],
                     body_f=state["body_f"])
    means = {e: float(np.mean(np.concatenate(acc[e]))) for e in EPS}
    print("mean Newton iterations by start error eps:", {e: round(m, 3) for e, m in means.items()})
    return means, float(np.mean(np.concatenate(plain)))


def test_exact_start_needs_no_iteration(curve):
    means, _ = curve
    assert means[0.0] == 0.0


def test_eps_one_is_the_plain_start(curve):
    means, plain = curve
    assert means[1.0] == pytest.approx(plain, abs=1e-12)
    assert plain > 3.0


def test_iterations_fall_slowly_with_start_error(curve):
    means, plain = curve
    vals = [means[e] for e in EPS]
    assert all(a <= b + 0.05 for a, b in zip(vals, vals[1:]))  # monotone up to noise
    # a 40 % cut needs a start at least ~4 times closer to the solution than the previous step's
    assert means[0.3] > 0.6 * plain
    # even a start 20 times closer still needs about two iterations (checking optimality)
    assert means[0.05] > 1.5

"""The C++ CPU restatement of the step (``oracle/cpu_step.cpp``, the CPU baseline) against the numpy
oracle it restates: same SAP velocities (to 1e-10 relative), same Newton iteration and contact
counts per world, same public outputs to float32 rounding; and the threaded collide equals the
single-threaded one bit for bit."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import collide as oc
from oracle import cpu_step as cs
from oracle import sap_oracle as so
from tests.helpers import control_dict_from_model, state_dict_from_model


def _scene(name):
    from sap_warp_amd import scenes

    if name == "humanoid":
        return scenes.build_humanoid_scene(4, seed=3), 2.0e-3, 64
    if name == "ant":
        return scenes.build_ant_scene(4, seed=1), 1.0e-3, 16
    if name == "ball_chain":
        return scenes.build_ball_chain_scene(4, seed=7), 2.0e-3, 64
    if name == "boxpile":
        return scenes.build_box_pile_scene(4, seed=2), 2.0e-3, 64
    return scenes.build_round_shapes_scene(4, seed=9), 2.0e-3, 16


@pytest.mark.parametrize("scene", ["humanoid", "ant", "ball_chain", "boxpile", "round"])
@pytest.mark.parametrize("preset,ls", [("approx32", "exact_root"), ("drake", "armijo_decay"),
                                       ("approx32", "monotone_decay")])
def test_cpu_step_matches_oracle(scene, preset, ls):
    model, dt, C = _scene(scene)
    st, ctl = state_dict_from_model(model), control_dict_from_model(model)
    c = oc.collide(model, st["body_q"], C)
    opts = so.StepOptions(preset=preset, line_search=ls, max_rigid_contact=C)
    ref = so.step(model, st, ctl, c, dt, opts)
    out = cs.CpuStepper(model, opts).step(st, ctl, c, dt, threads=2)
    scale = max(1.0, float(np.max(np.abs(ref["v_sap"]))))
    assert float(np.max(np.abs(out["v_sap"] - ref["v_sap"]))) <= 1e-10 * scale
    for w in range(model.world_count):
        r = ref["per_world"][w]
        assert out["iterations"][w] == r["result"].iterations
        assert out["contact_count"][w] == r["contacts"].count
        assert out["converged"][w] == 1
    np.testing.assert_allclose(out["joint_q"], ref["joint_q"], atol=2e-6)
    np.testing.assert_allclose(out["body_q"], ref["body_q"], atol=2e-6)
    np.testing.assert_allclose(out["body_qd"], ref["body_qd"], atol=2e-5 * max(1.0, float(np.max(np.abs(ref["body_qd"])))))


def test_cpu_step_warm_start_and_threads():
    """Three chained steps with the previous solution as warm start; 1 and 4 threads agree exactly."""
    model, dt, C = _scene("humanoid")
    st, ctl = state_dict_from_model(model), control_dict_from_model(model)
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=C)
    stp = cs.CpuStepper(model, opts)
    s1 = s4 = st
    g1 = g4 = None
    sref, gref = st, None
    for _ in range(3):
        c = oc.collide(model, s1["body_q"], C)
        o1 = stp.step(s1, ctl, c, dt, v_guess=g1, threads=1)
        o4 = stp.step(s4, ctl, c, dt, v_guess=g4, threads=4)
        ref = so.step(model, sref, ctl, c, dt, opts, v_guess=gref)
        np.testing.assert_array_equal(o1["v_sap"], o4["v_sap"])
        assert float(np.max(np.abs(o1["v_sap"] - ref["v_sap"]))) <= 1e-10 * max(1.0, float(np.max(np.abs(ref["v_sap"]))))
        nxt = lambda o: dict(joint_q=o["joint_q"], joint_qd=o["joint_qd"], body_q=o["body_q"], body_qd=o["body_qd"],
                             body_f=st["body_f"])
        s1, s4, g1, g4 = nxt(o1), nxt(o4), o1["v_sap"], o4["v_sap"]
        sref = dict(joint_q=ref["joint_q"].astype(np.float32), joint_qd=ref["joint_qd"].astype(np.float32),
                    body_q=ref["body_q"].astype(np.float32), body_qd=ref["body_qd"].astype(np.float32),
                    body_f=st["body_f"])
        gref = ref["v_sap"]


@pytest.mark.parametrize("scene", ["humanoid", "boxpile", "round"])
def test_threaded_collide_is_identical(scene):
    model, _, C = _scene(scene)
    a = oc.collide(model, model.body_q, C)
    b = oc.collide(model, model.body_q, C, threads=3)
    for k in ("count", "world_start", "shape0", "shape1", "point0", "point1", "normal", "margin0", "margin1", "pair"):
        np.testing.assert_array_equal(np.asarray(a[k]), np.asarray(b[k]), err_msg=k)

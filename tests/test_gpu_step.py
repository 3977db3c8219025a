"""GPU step parity: one SolverSAP.step on the MI355X vs the float64 oracle from the same input.

Tolerances (north star): post-step generalized velocities within 1e-5
relative (max-norm, relative to max(|v|, 1 m/s)), positions within 1e-5, and
every world converged by the SAP optimality / cost criterion.
"""
import numpy as np
import pytest

from oracle import sap_oracle as so
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu

VEL_TOL = 1.0e-5


def _scene(name, worlds, seed):
    from sap_warp_amd.scenes import build_ant_scene, build_box_pile_scene, build_humanoid_scene

    if name == "ant":
        return build_ant_scene(worlds, seed=seed, device="cuda:0"), 1.0e-3, 16
    if name == "boxpile":
        return build_box_pile_scene(worlds, seed=seed, device="cuda:0"), 2.0e-3, 64
    if name == "ball_chain":
        from sap_warp_amd.scenes import build_ball_chain_scene

        return build_ball_chain_scene(worlds, seed=seed + 7, device="cuda:0"), 2.0e-3, 64
    if name in ("quadruped", "arm_gripper", "pile"):
        from sap_warp_amd import hetero

        build = {"quadruped": hetero.build_quadruped_scene, "arm_gripper": hetero.build_arm_gripper_scene,
                 "pile": hetero.build_pile_scene}[name]
        return build(worlds, seed=seed + 3, device="cuda:0"), 2.0e-3, 64
    if name in ("hinge_slider_stack", "cube_toss"):
        import os

        from sap_warp_amd.scene_yaml import load_scene

        path = os.path.join(os.path.dirname(__file__), "..", "sap_warp_amd", "assets", name + ".yaml")
        sc = load_scene(path, num_worlds=worlds, device="cuda:0")
        return sc.model, sc.dt, sc.max_rigid_contact
    return build_humanoid_scene(worlds, seed=seed, device="cuda:0"), 2.0e-3, 64


def _np(t):
    return t.detach().cpu().numpy()


def _contacts_np(c):
    count = int(_np(c.rigid_contact_count)[0])
    return dict(count=count, world_start=_np(c.rigid_contact_world_start), shape0=_np(c.rigid_contact_shape0),
                shape1=_np(c.rigid_contact_shape1), point0=_np(c.rigid_contact_point0),
                point1=_np(c.rigid_contact_point1), normal=_np(c.rigid_contact_normal),
                margin0=_np(c.rigid_contact_margin0), margin1=_np(c.rigid_contact_margin1))


def _state_np(s):
    return dict(joint_q=_np(s.joint_q), joint_qd=_np(s.joint_qd), body_q=_np(s.body_q), body_qd=_np(s.body_qd),
                body_f=_np(s.body_f))


def _control_np(c):
    return dict(joint_f=_np(c.joint_f), joint_target_pos=_np(c.joint_target_pos),
                joint_target_vel=_np(c.joint_target_vel), joint_act=_np(c.joint_act))


def _run_and_compare(scene, preset, ls, worlds, steps, seed=0, schedule=None, cost_rel_tol=None, vel_tol=VEL_TOL):
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.solver_sap import SolverSAP

    model, dt, C = _scene(scene, worlds, seed)
    solver = SolverSAP(model, max_rigid_contact=C, line_search_variant=ls, contact_preset_variant=preset,
                       cost_rel_tol=cost_rel_tol,
                       capture_contact_jacobian_snapshots=True,
                       collect_iteration_stats=True)
    if schedule is not None:
        solver.set_schedule(**schedule)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts = pipe.contacts()
    s0, s1 = model.state(), model.state()
    ctl = model.control()
    solve = None if cost_rel_tol is None else so.SolveOptions.for_variant(ls, cost_rel_tol=cost_rel_tol)
    opts = so.StepOptions(preset=preset, line_search=ls, max_rigid_contact=C, solve=solve)
    v_guess = None
    worst = dict(v=0.0, vstar=0.0, q=0.0, qd=0.0, bq=0.0)
    for k in range(steps):
        pipe.collide(s0, contacts)
        st_in = _state_np(s0)
        c_np = _contacts_np(contacts)
        solver.step(s0, s1, ctl, contacts, dt)
        torch.cuda.synchronize()
        stats = solver.last_contact_solve_result
        ref = so.step(model, st_in, _control_np(ctl), c_np, dt, opts, v_guess=v_guess)
        v_gpu = _np(solver.read_buffer(0))
        vs_gpu = _np(solver.read_buffer(1))
        for w in range(model.world_count):
            r = ref["per_world"][w]
            d0, d1 = r["free_motion"].v_star.shape[0] * w, r["free_motion"].v_star.shape[0] * (w + 1)
            worst["vstar"] = max(worst["vstar"], rel_err(vs_gpu[d0:d1], r["v_star"]))
            worst["v"] = max(worst["v"], rel_err(v_gpu[d0:d1], r["v_sap"]))
            assert stats.converged_env[w] == 1, f"world {w} did not converge at step {k}"
            assert r["result"].converged, f"oracle world {w} did not converge at step {k}"
        worst["q"] = max(worst["q"], rel_err(_np(s1.joint_q), ref["joint_q"]))
        worst["qd"] = max(worst["qd"], rel_err(_np(s1.joint_qd), ref["joint_qd"]))
        worst["bq"] = max(worst["bq"], rel_err(_np(s1.body_q), ref["body_q"]))
        v_guess = v_gpu.astype(np.float64)
        s0, s1 = s1, s0
    print(scene, preset, ls, {k: f"{v:.2e}" for k, v in worst.items()})
    assert worst["vstar"] <= VEL_TOL, worst
    assert worst["v"] <= vel_tol, worst
    assert worst["qd"] <= vel_tol, worst
    assert worst["q"] <= VEL_TOL, worst
    assert worst["bq"] <= VEL_TOL, worst


@pytest.mark.parametrize("preset,ls", [("approx32", "armijo_decay"), ("approx32", "exact_root"),
                                       ("drake", "exact_root"), ("drake", "armijo_decay")])
def test_humanoid_step_parity(gpu_device, preset, ls):
    _run_and_compare("humanoid", preset, ls, worlds=6, steps=5)


@pytest.mark.parametrize("preset", ["approx32", "drake"])
def test_humanoid_step_parity_monotone_decay(gpu_device, preset):
    """monotone_decay at a tight cost tolerance and at the reference's default stop (relative cost
    change below 5e-3, contact_solve.py:6179-6181), both to the 1e-5 bar."""
    _run_and_compare("humanoid", preset, "monotone_decay", worlds=6, steps=5, cost_rel_tol=1.0e-15)
    _run_and_compare("humanoid", preset, "monotone_decay", worlds=6, steps=5)


@pytest.mark.parametrize("kernel", ["wave64", "fused"])
def test_humanoid_step_parity_other_kernels(gpu_device, monkeypatch, kernel):
    """The 64-slot wave variant alone (every world) and the 256-thread kernel, same bar."""
    monkeypatch.setenv("SAPX_STEP_KERNEL", kernel)
    _run_and_compare("humanoid", "approx32", "exact_root", worlds=6, steps=5)


def test_ant_step_parity(gpu_device):
    _run_and_compare("ant", "approx32", "armijo_decay", worlds=8, steps=5)


@pytest.mark.parametrize("small,fork,order", [(32, True, True), (32, False, False), (16, False, True)])
def test_humanoid_step_parity_schedules(gpu_device, small, fork, order):
    """Other work schedules of the wave kernel (32-slot variant, serial variants, identity order)."""
    _run_and_compare("humanoid", "drake", "exact_root", worlds=6, steps=3,
                     schedule={"kernel": "wave", "small_capacity": small, "fork": fork, "world_order": order})


@pytest.mark.parametrize("preset,ls", [("approx32", "exact_root"), ("drake", "armijo_decay")])
def test_box_pile_step_parity(gpu_device, preset, ls):
    """Six free bodies per world (a forest of free joints) in box-box / capsule-box contact."""
    _run_and_compare("boxpile", preset, ls, worlds=8, steps=4)


@pytest.mark.parametrize("scene,preset,ls", [("hinge_slider_stack", "drake", "exact_root"),
                                             ("hinge_slider_stack", "approx32", "monotone_decay"),
                                             ("cube_toss", "approx32", "armijo_decay")])
def test_yaml_scene_step_parity(gpu_device, scene, preset, ls):
    """Scenes loaded from the inline scene files: fixed anchors, a limited revolute and prismatic
    joint with drives, free boxes stacked in box-box / capsule-box contact, replicated worlds."""
    _run_and_compare(scene, preset, ls, worlds=8, steps=4, cost_rel_tol=1.0e-15 if ls == "monotone_decay" else None)


@pytest.mark.parametrize("scene", ["quadruped", "arm_gripper", "pile"])
def test_hetero_kind_step_parity(gpu_device, scene):
    """The other world kinds of the heterogeneous batch, each against homogeneous oracle runs."""
    _run_and_compare(scene, "approx32", "exact_root", worlds=8, steps=4)


@pytest.mark.parametrize("preset,ls", [("approx32", "exact_root"), ("drake", "armijo_decay")])
def test_ball_chain_step_parity(gpu_device, preset, ls):
    """Capsule chains on ball joints (quaternion coordinates, three angular mobilities, velocity
    damping drives) swinging onto a free box."""
    _run_and_compare("ball_chain", preset, ls, worlds=8, steps=5)


@pytest.mark.parametrize("kernel", ["fused", "wave64"])
def test_ball_chain_step_parity_other_kernels(gpu_device, monkeypatch, kernel):
    """Ball joints in the 256-thread kernel and in the 64-slot wave variant alone, same bar."""
    monkeypatch.setenv("SAPX_STEP_KERNEL", kernel)
    _run_and_compare("ball_chain", "approx32", "exact_root", worlds=8, steps=5)

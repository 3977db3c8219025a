"""Inline scene files (``sap_warp_amd.scene_yaml``): structure against closed forms.

Masses and inertias from the shapes' geometry and density, poses from forward
kinematics written out by hand, world offsets of a replicated grid, the
reference's own inline scenes (read as data when the reference checkout is
present), and the loaded scenes stepping through the float64 oracle.
"""
from __future__ import annotations

import math
import os

import numpy as np
import pytest

from sap_warp_amd.scene_yaml import SceneLoaderError, UnsupportedSceneFeature, build_scene, load_scene, number
from sap_warp_amd.scenes import world_offsets

ASSETS = os.path.join(os.path.dirname(__file__), "..", "sap_warp_amd", "assets")
REF_YAML = "/root/reference/assets/yaml"


def qrot(q, v):
    x, y, z, w = q
    u = np.array([x, y, z])
    return v + 2.0 * w * np.cross(u, v) + 2.0 * np.cross(u, np.cross(u, v))


def test_number_expressions():
    assert number("pi * 0.5") == pytest.approx(math.pi / 2)
    assert number("-pi / 4 + 1") == pytest.approx(1 - math.pi / 4)
    assert number(3) == 3.0
    for bad in ("__import__('os')", "e", "pi ** 2", "abs(-1)"):
        with pytest.raises(ValueError):
            number(bad)


def test_world_offsets_grid():
    np.testing.assert_allclose(world_offsets(4, (2.0, 2.0, 0.0)), [[-1, -1, 0], [1, -1, 0], [-1, 1, 0], [1, 1, 0]])
    np.testing.assert_allclose(world_offsets(3, (1.0, 0.0, 0.0)), [[-1, 0, 0], [0, 0, 0], [1, 0, 0]])
    np.testing.assert_allclose(world_offsets(2, (0.0, 0.0, 0.0)), 0.0)
    # z is the up axis: stacked worlds are not re-centred along it
    np.testing.assert_allclose(world_offsets(2, (0.0, 0.0, 3.0))[:, 2], [0.0, 3.0])


def _box_inertia(m, h):
    hx, hy, hz = h
    return np.diag([m / 3 * (hy * hy + hz * hz), m / 3 * (hx * hx + hz * hz), m / 3 * (hx * hx + hy * hy)])


def test_hinge_slider_stack_structure():
    sc = load_scene(os.path.join(ASSETS, "hinge_slider_stack.yaml"))
    m = sc.model
    assert sc.dt == 0.002 and sc.max_rigid_contact == 64
    assert sc.solver == {"contact_preset_variant": "drake", "line_search_variant": "exact_root"}
    assert m.joint_type.tolist() == [3, 1, 3, 0, 4, 4, 4]  # fixed, revolute, fixed, prismatic, 3 x free
    ids = sc.body_ids
    # carriage: box 0.12 x 0.12 x 0.08 at its own density 2000
    mc = 2000.0 * 0.12 * 0.12 * 0.08
    assert m.body_mass[ids["carriage"]] == pytest.approx(mc, rel=1e-6)
    np.testing.assert_allclose(m.body_inertia[ids["carriage"]], _box_inertia(mc, (0.06, 0.06, 0.04)), rtol=1e-5)
    # hinge at pi/4 about x through the anchor's bottom face, arm hanging 0.3 below the joint
    pivot = np.array([0.0, -1.0, 1.2 - 0.05])
    q = m.body_q[ids["arm"]]
    qh = (math.sin(math.pi / 8), 0.0, 0.0, math.cos(math.pi / 8))
    np.testing.assert_allclose(q[3:], qh, atol=1e-6)
    np.testing.assert_allclose(q[:3], pivot - qrot(qh, np.array([0.0, 0.0, 0.3])), atol=1e-6)
    # slider at +0.1 along x below the rail
    np.testing.assert_allclose(m.body_q[ids["carriage"]][:3], [1.1, 0.0, 0.82], atol=1e-6)
    # limits: expressions, defaults for stiffness, free joints unlimited
    np.testing.assert_allclose([m.joint_limit_lower[0], m.joint_limit_upper[0]], [-math.pi / 2, math.pi / 2])
    np.testing.assert_allclose([m.joint_limit_lower[1], m.joint_limit_upper[1]], [-0.25, 0.25])
    assert m.joint_limit_ke[0] == 1000.0 and np.all(m.joint_limit_ke[2:] == 0.0)
    np.testing.assert_allclose(m.joint_armature[:2], 0.01)
    np.testing.assert_allclose(m.joint_armature[2:], 0.0)
    # post-build drives: hinge position drive to its initial angle, slider velocity drive at -0.2
    assert m.joint_target_mode[0] == 1 and m.joint_target_ke[0] == 50.0
    assert m.joint_target_pos[0] == pytest.approx(math.pi / 4)
    assert m.joint_target_mode[1] == 2 and m.joint_target_kd[1] == 5.0 and m.joint_target_vel[1] == -0.2
    # shape materials: body-level physics override, shape cfg override, post-build tau, ground cfg
    shape_of = {int(b): i for i, b in enumerate(m.shape_body) if b >= 0}
    assert m.shape_material_mu[shape_of[ids["base_box"]]] == pytest.approx(0.6)
    assert m.shape_material_mu[shape_of[ids["roller"]]] == pytest.approx(0.4)
    assert m.shape_material_mu[-1] == 1.0 and m.shape_flags[-1] == 3
    np.testing.assert_allclose(m.shape_material_tau[:-1], 0.005)
    # parent-child filters: anchor-arm, rail-carriage
    filt = {tuple(p) for p in np.asarray(m.shape_collision_filter_pairs).tolist()}
    assert (shape_of[ids["anchor"]], shape_of[ids["arm"]]) in filt
    assert (shape_of[ids["rail"]], shape_of[ids["carriage"]]) in filt


def test_replicated_worlds():
    sc = load_scene(os.path.join(ASSETS, "cube_toss.yaml"))
    m = sc.model
    assert m.world_count == 16 and m.body_count == 16
    off = world_offsets(16, (1.0, 1.0, 0.0))
    np.testing.assert_allclose(m.body_q[:, :3], off + np.array([0.0, 0.0, 0.06]), atol=1e-6)
    q = np.asarray(m.joint_q).reshape(16, 7)
    np.testing.assert_allclose(q[:, :3], m.body_q[:, :3], atol=1e-6)
    np.testing.assert_allclose(np.asarray(m.joint_qd).reshape(16, 6), np.tile([0.4, -0.2, 0.5, 3.0, 0.0, 2 * math.pi],
                                                                              (16, 1)), rtol=1e-6)
    assert np.all(np.asarray(m.shape_world)[:16] == np.arange(16)) and m.shape_world[-1] == -1
    assert m.body_mass[0] == pytest.approx(1000.0 * 0.1 ** 3)


def test_strict_unsupported_features():
    cfg = {"bodies": [{"id": "b", "shapes": [{"type": "mesh"}]}],
           "joints": [{"id": "j", "type": "ball", "parent": "world", "child": "b"}]}
    with pytest.raises(UnsupportedSceneFeature) as e:
        build_scene(cfg)
    assert len(e.value.features) == 2
    sc = build_scene(cfg, strict=False)
    assert len(sc.unsupported) == 2
    # a joint naming an unknown body is reported like an unsupported joint (scene.py:1487-1491)
    with pytest.raises(UnsupportedSceneFeature):
        build_scene({"joints": [{"id": "j", "type": "free", "parent": "world", "child": "nope"}]})
    with pytest.raises(SceneLoaderError):
        build_scene({"bodies": [{"shapes": []}]})
    with pytest.raises(SceneLoaderError):
        build_scene({"articulations": [{"id": "a", "joints": ["missing"]}]})


def test_round_shape_kinds():
    """Cylinder, cone and ellipsoid shapes (scene.py:5135-5152) with their masses and the cone's
    centre of mass a quarter of its height below the shape origin (scene.py:5878-5884)."""
    cfg = {"bodies": [{"id": "cyl", "shapes": [{"type": "cylinder", "radius": 0.1, "half_height": 0.2,
                                                 "cfg": {"density": 500.0}}]},
                      {"id": "cone", "shapes": [{"type": "cone", "radius": 0.1, "half_height": 0.2,
                                                  "cfg": {"density": 500.0}}]},
                      {"id": "ell", "shapes": [{"type": "ellipsoid", "a": 0.3, "b": 0.2, "c": 0.1,
                                                 "cfg": {"density": 500.0}}]}]}
    m = build_scene(cfg).model
    ids = build_scene(cfg).body_ids
    assert m.body_mass[ids["cyl"]] == pytest.approx(500.0 * math.pi * 0.01 * 0.4)
    assert m.body_mass[ids["cone"]] == pytest.approx(500.0 * math.pi * 0.01 * 0.4 / 3.0)
    assert m.body_mass[ids["ell"]] == pytest.approx(500.0 * 4.0 / 3.0 * math.pi * 0.3 * 0.2 * 0.1)
    np.testing.assert_allclose(m.body_com[ids["cone"]], [0.0, 0.0, -0.1], atol=1e-7)
    np.testing.assert_allclose(m.body_com[ids["cyl"]], [0.0, 0.0, 0.0], atol=1e-7)
    # the cone's inertia about its centre of mass: 3/20 m r^2 + 3/80 m h^2 across, 3/10 m r^2 along
    mc = 500.0 * math.pi * 0.01 * 0.4 / 3.0
    np.testing.assert_allclose(np.diag(np.asarray(m.body_inertia[ids["cone"]]).reshape(3, 3)),
                               [0.15 * mc * 0.01 + 3.0 / 80.0 * mc * 0.16] * 2 + [0.3 * mc * 0.01], rtol=1e-5)


@pytest.mark.skipif(not os.path.isdir(REF_YAML), reason="reference scene files not present")
def test_reference_inline_scenes():
    """The reference's two inline scenes load unchanged."""
    sc = load_scene(os.path.join(REF_YAML, "free_cube.yaml"))
    m = sc.model
    assert m.body_mass[0] == pytest.approx(1500.0 * 0.04 ** 3)
    np.testing.assert_allclose(m.joint_qd, [0.5, 0.2, 0.3, 2.0, 1.0, 6.0])
    assert sc.solver["contact_preset_variant"] == "drake"
    sc = load_scene(os.path.join(REF_YAML, "multi_joints.yaml"))
    m = sc.model
    assert m.joint_type.tolist() == [3, 1, 3, 0, 4, 4]
    assert m.joint_q[0] == pytest.approx(math.pi / 2)
    b = sc.body_ids["b_rev"]
    np.testing.assert_allclose(m.body_q[b][:3], [0.0, -3.0 + 0.75, 2.1875 - 0.1875], atol=1e-6)
    assert sc.solver["line_search_variant"] == "monotone_decay"


def test_loaded_scene_steps_in_oracle():
    """The stack scene through collision + one float64 oracle step: the fixed anchors stay put and
    every world converges."""
    from oracle import sap_oracle as so
    from oracle.collide import collide
    from tests.helpers import control_dict_from_model, state_dict_from_model

    sc = load_scene(os.path.join(ASSETS, "hinge_slider_stack.yaml"), num_worlds=2)
    m = sc.model
    c = collide(m, m.body_q, sc.max_rigid_contact)
    assert c["unsupported"] == 0 and c["count"] > 0
    opts = so.StepOptions(preset="drake", line_search="exact_root", max_rigid_contact=sc.max_rigid_contact)
    out = so.step(m, state_dict_from_model(m), control_dict_from_model(m), c, sc.dt, opts)
    assert all(r["result"].converged for r in out["per_world"].values())
    anchor = sc.body_ids["anchor"]
    np.testing.assert_allclose(out["body_q"][anchor], m.body_q[anchor], atol=1e-9)

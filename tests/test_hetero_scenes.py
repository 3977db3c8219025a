"""World kinds of the heterogeneous batch: sizes, supported contact pairs, one oracle step each."""
from __future__ import annotations

import numpy as np
import pytest

from oracle import sap_oracle as so
from oracle.collide import collide
from sap_warp_amd.hetero import (WORLD_KINDS, build_arm_gripper_scene, build_pile_scene, build_quadruped_scene,
                                 draw_world_kinds)
from tests.helpers import control_dict_from_model, state_dict_from_model


@pytest.mark.parametrize("build,dof,bodies", [(build_quadruped_scene, 18, 13), (build_arm_gripper_scene, 15, 10),
                                              (build_pile_scene, 48, 8)])
def test_world_kind_builds_and_steps(build, dof, bodies):
    m = build(4)
    assert m.joint_dof_count == 4 * dof and m.body_count == 4 * bodies
    c = collide(m, m.body_q, 64)
    assert c["unsupported"] == 0
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=64)
    out = so.step(m, state_dict_from_model(m), control_dict_from_model(m), c, 2.0e-3, opts)
    assert all(r["result"].converged for r in out["per_world"].values())
    assert np.all(np.isfinite(out["joint_q"]))


def test_arm_gripper_grasps_box():
    """The fingers start astride the box and their first contacts are with it (box-box pairs)."""
    m = build_arm_gripper_scene(3)
    c = collide(m, m.body_q, 64)
    st = np.asarray(m.shape_type)
    n = c["count"]
    kinds = set(zip(st[c["shape0"][:n]].tolist(), st[c["shape1"][:n]].tolist()))
    assert (7, 7) in kinds


def test_draw_world_kinds():
    counts = draw_world_kinds(8192, 0)
    assert set(counts) == set(WORLD_KINDS) and sum(counts.values()) == 8192
    assert min(counts.values()) > 1800
    assert draw_world_kinds(8192, 0) == counts

"""The whole step (collision + solver step, with the forked 64-contact launch and the
dispatch-order kernel) captured in a HIP graph and replayed: every replay must equal the
eager step from the same input, with eager work and synchronisations between replays.
"""
from __future__ import annotations

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_graph_replay_matches_eager(gpu_device):
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.scenes import build_humanoid_scene
    from sap_warp_amd.solver_sap import SolverSAP

    model = build_humanoid_scene(8, seed=3, device="cuda:0")
    solver = SolverSAP(model, max_rigid_contact=64, line_search_variant="exact_root")
    pipe = SapCollisionPipeline(model, max_contacts_per_world=64)
    contacts = pipe.contacts()
    ctl = model.control()
    ctl.joint_target_pos.add_(0.05)
    s_in, s_out, s_other = model.state(), model.state(), model.state()
    dt = 2.0e-3

    def step(a, b):
        pipe.collide(a, contacts)
        solver.step(a, b, ctl, contacts, dt)

    # eager reference from the plain start (no warm-start guess)
    solver.reset_runtime_state()
    step(s_in, s_out)
    torch.cuda.synchronize()
    ref = {k: _np(getattr(s_out, k)).copy() for k in ("joint_q", "joint_qd", "body_q", "body_qd")}
    count_ref = int(_np(contacts.rigid_contact_count)[0])

    solver.reset_runtime_state()
    g = torch.cuda.CUDAGraph()
    stream = torch.cuda.Stream()
    stream.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(stream):
        with torch.cuda.graph(g, stream=stream):
            step(s_in, s_out)
    torch.cuda.current_stream().wait_stream(stream)
    torch.cuda.synchronize()
    for r in range(3):
        for k in ("joint_q", "joint_qd", "body_q", "body_qd"):
            getattr(s_out, k).zero_()
        g.replay()
        torch.cuda.synchronize()
        for k in ("joint_q", "joint_qd", "body_q", "body_qd"):
            np.testing.assert_array_equal(_np(getattr(s_out, k)), ref[k], err_msg=f"replay {r} field {k}")
        assert int(_np(contacts.rigid_contact_count)[0]) == count_ref
        # the program's own eager work between replays
        solver.reset_runtime_state()
        step(s_out, s_other)
        torch.cuda.synchronize()
        solver.reset_runtime_state()

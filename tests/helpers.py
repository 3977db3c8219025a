"""Shared helpers for the parity tests (CPU oracle vs native GPU path)."""
from __future__ import annotations

import numpy as np

from oracle import collide as oc
from oracle import sap_oracle as so


def oracle_contacts_dict(model, body_q, C):
    """Run the float32 collision oracle and return its arrays as a dict."""
    return oc.collide(model, np.asarray(body_q, dtype=np.float32), C)


def contacts_to_torch(c, model, device):
    """Upload oracle contact arrays into a SapContacts on ``device``."""
    import torch

    from sap_warp_amd.runtime import SapContacts

    cap = max(1, len(c["shape0"]))
    t = lambda a, dt: torch.as_tensor(np.ascontiguousarray(a), dtype=dt, device=device)  # noqa: E731
    return SapContacts(
        rigid_contact_count=t(np.array([c["count"]], np.int32), torch.int32),
        rigid_contact_shape0=t(c["shape0"], torch.int32),
        rigid_contact_shape1=t(c["shape1"], torch.int32),
        rigid_contact_point0=t(c["point0"], torch.float32),
        rigid_contact_point1=t(c["point1"], torch.float32),
        rigid_contact_normal=t(c["normal"], torch.float32),
        rigid_contact_margin0=t(c["margin0"], torch.float32),
        rigid_contact_margin1=t(c["margin1"], torch.float32),
        rigid_contact_world_start=t(c["world_start"], torch.int32),
        rigid_contact_max=cap,
    )


def state_dict_from_model(model):
    return dict(joint_q=model.joint_q.copy(), joint_qd=model.joint_qd.copy(), body_q=model.body_q.copy(),
                body_qd=model.body_qd.copy(), body_f=np.zeros((model.body_count, 6), np.float32))


def control_dict_from_model(model, target_pos=None):
    return dict(joint_f=model.joint_f.copy(), joint_target_pos=(model.joint_target_pos if target_pos is None
                                                               else target_pos).copy(),
                joint_target_vel=model.joint_target_vel.copy(), joint_act=model.joint_act.copy())


def rel_err(a, b, floor=1.0):
    a = np.asarray(a, dtype=np.float64)
    b = np.asarray(b, dtype=np.float64)
    return float(np.max(np.abs(a - b)) / max(float(np.max(np.abs(b))), floor))


def world_rel_err(a, b):
    """max|a - b| / max|b| over one world's entries, with no unit floor (0 when both are empty or zero)."""
    a = np.asarray(a, dtype=np.float64).reshape(-1)
    b = np.asarray(b, dtype=np.float64).reshape(-1)
    if b.size == 0:
        return 0.0
    diff = float(np.max(np.abs(a - b)))
    scale = float(np.max(np.abs(b)))
    if scale == 0.0:
        return 0.0 if diff == 0.0 else float("inf")
    return diff / scale


def oracle_step_options(preset="approx32", line_search="armijo_decay", C=64):
    return so.StepOptions(preset=preset, line_search=line_search, max_rigid_contact=C)

"""Runtime containers for the SAP compliant-contact step.

Mirrors the container API of the reference (``sim/sap_runtime.py:47-264``):
``SapModel`` is immutable topology + parameters, ``SapState``/``SapControl`` are
mutable per-step buffers, ``SapContacts`` is the flat rigid-contact bundle that
``SapCollisionPipeline.collide`` writes and ``SolverSAP.step`` consumes.

Layout differences, all deliberate:

* Model arrays are host ``numpy`` arrays.  The model is immutable, so the native
  library receives one copy at solver construction (``sapx_model_create``) and
  keeps it resident in HBM; nothing is re-uploaded per step.
* State / control / contact buffers are ``torch`` tensors (device memory on the
  GPU path).  Only their data pointers cross the C ABI.
* Transforms are stored as ``[px, py, pz, qx, qy, qz, qw]`` float32 rows and
  spatial vectors as 6 float32 — the same memory layout as Warp's
  ``wp.transform`` / ``wp.spatial_vector`` (quaternions xyzw,
  ``docs/concepts/convention.rst:17-21``).
"""
from __future__ import annotations

from dataclasses import dataclass, field
from typing import Any

import numpy as np

try:  # torch is plumbing only (device memory + streams)
    import torch
except Exception:  # pragma: no cover - torch is always present in this image
    torch = None

# Joint types (sim/sap_runtime.py:20-26)
SAP_JOINT_PRISMATIC = 0
SAP_JOINT_REVOLUTE = 1
SAP_JOINT_BALL = 2
SAP_JOINT_FIXED = 3
SAP_JOINT_FREE = 4
SAP_JOINT_DISTANCE = 5
SAP_JOINT_D6 = 6

SAP_JOINT_TARGET_NONE = 0  # sim/sap_runtime.py:28

SAP_BODY_FLAG_DYNAMIC = 1 << 0  # sim/sap_runtime.py:30-32
SAP_BODY_FLAG_KINEMATIC = 1 << 1
SAP_BODY_FLAG_ALL = SAP_BODY_FLAG_DYNAMIC | SAP_BODY_FLAG_KINEMATIC

# Shape types (sim/collision/types.py:14-26)
GEO_PLANE = 1
GEO_SPHERE = 3
GEO_CAPSULE = 4
GEO_ELLIPSOID = 5
GEO_CYLINDER = 6
GEO_BOX = 7
GEO_CONE = 9

# Shape flags (sim/collision/flags.py:12-19)
SHAPE_FLAG_VISIBLE = 1 << 0
SHAPE_FLAG_COLLIDE_SHAPES = 1 << 1
SHAPE_FLAG_COLLIDE_PARTICLES = 1 << 2


@dataclass(frozen=True)
class SapModel:
    """Immutable model arrays for ``world_count`` equally-shaped worlds.

    Field names and meanings follow ``SapModel`` in ``sim/sap_runtime.py:47-137``.
    Every per-joint / per-dof / per-body / per-shape array is global (all
    worlds, world-major).  ``shape_pairs`` is the explicit candidate pair list
    (``sim/loader/scene.py:1188-1212``).
    """

    device: str
    joint_count: int
    joint_dof_count: int
    joint_coord_count: int
    body_count: int
    articulation_count: int
    world_count: int
    shape_count: int

    joint_type: np.ndarray
    joint_articulation: np.ndarray
    joint_parent: np.ndarray
    joint_child: np.ndarray
    joint_q_start: np.ndarray  # joint_count + 1 (sentinel)
    joint_qd_start: np.ndarray  # joint_count + 1 (sentinel)
    joint_dof_dim: np.ndarray  # (joint_count, 2): linear, angular axis counts
    joint_axis: np.ndarray  # (dof, 3) f32
    joint_X_p: np.ndarray  # (joint, 7) f32
    joint_X_c: np.ndarray  # (joint, 7) f32
    joint_ancestor: np.ndarray
    articulation_start: np.ndarray  # articulation_count + 1
    joint_q: np.ndarray
    joint_qd: np.ndarray

    body_q: np.ndarray  # (body, 7)
    body_qd: np.ndarray  # (body, 6)
    body_mass: np.ndarray
    body_inertia: np.ndarray  # (body, 3, 3)
    body_com: np.ndarray  # (body, 3)
    body_flags: np.ndarray
    body_world: np.ndarray
    gravity: np.ndarray  # (world, 3)
    joint_armature: np.ndarray

    shape_body: np.ndarray
    shape_material_mu: np.ndarray
    shape_material_ke: np.ndarray
    shape_material_tau: np.ndarray

    joint_target_mode: np.ndarray
    joint_target_ke: np.ndarray
    joint_target_kd: np.ndarray
    joint_target_pos: np.ndarray
    joint_target_vel: np.ndarray
    joint_f: np.ndarray
    joint_act: np.ndarray
    joint_effort_limit: np.ndarray
    joint_limit_lower: np.ndarray
    joint_limit_upper: np.ndarray
    joint_limit_ke: np.ndarray
    joint_limit_kd: np.ndarray

    shape_transform: np.ndarray  # (shape, 7)
    shape_type: np.ndarray
    shape_scale: np.ndarray  # (shape, 3)
    shape_flags: np.ndarray
    shape_margin: np.ndarray
    shape_gap: np.ndarray
    shape_collision_radius: np.ndarray
    shape_collision_group: np.ndarray
    shape_world: np.ndarray
    shape_pairs: np.ndarray  # (P, 2) int32 explicit candidate pairs
    # (F, 2) int32 excluded pairs (min, max): same-body and parent-child shapes (scene.py:333-373, 409-414)
    shape_collision_filter_pairs: np.ndarray | None = None
    up_axis: int = 2
    labels: dict = field(default_factory=dict)

    # ----------------------------------------------------------------- helpers
    @property
    def dof_per_env(self) -> int:
        return self.joint_dof_count // max(self.world_count, 1)

    @property
    def bodies_per_env(self) -> int:
        return self.body_count // max(self.world_count, 1)

    @property
    def joints_per_env(self) -> int:
        return self.joint_count // max(self.world_count, 1)

    @property
    def coords_per_env(self) -> int:
        return self.joint_coord_count // max(self.world_count, 1)

    def state(self, device: str | None = None) -> "SapState":
        """Fresh public-order state initialised from the model (sap_runtime.py:139-155)."""
        dev = device or self.device
        return SapState(
            joint_q=_tensor(self.joint_q, dev),
            joint_qd=_tensor(self.joint_qd, dev),
            body_q=_tensor(self.body_q, dev),
            body_qd=_tensor(self.body_qd, dev),
            body_f=_tensor(np.zeros_like(self.body_qd), dev),
            joint_qd_order="public",
            body_f_order="public",
        )

    def control(self, device: str | None = None) -> "SapControl":
        """Public-order control initialised from the model (sap_runtime.py:157-174)."""
        dev = device or self.device
        return SapControl(
            joint_f=_tensor(self.joint_f, dev),
            joint_target_pos=_tensor(self.joint_target_pos, dev),
            joint_target_vel=_tensor(self.joint_target_vel, dev),
            joint_act=_tensor(self.joint_act, dev),
            joint_f_order="public",
        )


def _tensor(arr: np.ndarray, device: str):
    return torch.from_numpy(np.ascontiguousarray(arr)).to(device)


@dataclass(frozen=True)
class SapState:
    """Mutable state (sim/sap_runtime.py:177-205).  Tensors are mutated in place."""

    joint_q: Any
    joint_qd: Any
    body_q: Any
    body_qd: Any = None
    body_f: Any = None
    joint_qd_order: str = "public"
    body_f_order: str = "public"
    requires_grad: bool = False

    def clear_forces(self) -> None:
        if self.body_f is not None:
            self.body_f.zero_()

    def assign(self, other: "SapState") -> None:
        for name in ("joint_q", "joint_qd", "body_q", "body_qd", "body_f"):
            dst = getattr(self, name)
            src = getattr(other, name)
            if dst is None and src is None:
                continue
            if dst is None or src is None:
                raise ValueError(f"SapState assign mismatch for {name}")
            dst.copy_(src)


@dataclass(frozen=True)
class SapControl:
    """Mutable control input (sim/sap_runtime.py:208-227)."""

    joint_f: Any
    joint_target_pos: Any = None
    joint_target_vel: Any = None
    joint_act: Any = None
    joint_f_order: str = "public"

    def clear(self) -> None:
        self.joint_f.zero_()
        for t in (self.joint_target_pos, self.joint_target_vel, self.joint_act):
            if t is not None:
                t.zero_()


@dataclass(frozen=True)
class SapContacts:
    """Flat rigid-contact bundle (sim/sap_runtime.py:230-264, collision/contacts.py:14-114).

    ``rigid_contact_count`` is a 1-element int32 tensor holding the live count.
    ``rigid_contact_world_start`` (``world_count + 1`` int32) is an addition of
    this implementation: when present the flat list is grouped by world and the
    solver reads each world's contacts as a contiguous CSR segment.
    """

    rigid_contact_count: Any = None
    rigid_contact_shape0: Any = None
    rigid_contact_shape1: Any = None
    rigid_contact_point0: Any = None
    rigid_contact_point1: Any = None
    rigid_contact_normal: Any = None
    rigid_contact_margin0: Any = None
    rigid_contact_margin1: Any = None
    rigid_contact_world_start: Any = None
    rigid_contact_max: int = 0

    @property
    def has_rigid_contacts(self) -> bool:
        return self.rigid_contact_count is not None

    def clear(self) -> None:
        if self.rigid_contact_count is not None:
            self.rigid_contact_count.zero_()


def allocate_contacts(capacity: int, world_count: int, device: str) -> SapContacts:
    """Allocate a contact buffer of ``capacity`` flat entries."""
    capacity = max(int(capacity), 1)
    i32 = torch.int32
    f32 = torch.float32
    return SapContacts(
        rigid_contact_count=torch.zeros(1, dtype=i32, device=device),
        rigid_contact_shape0=torch.full((capacity,), -1, dtype=i32, device=device),
        rigid_contact_shape1=torch.full((capacity,), -1, dtype=i32, device=device),
        rigid_contact_point0=torch.zeros((capacity, 3), dtype=f32, device=device),
        rigid_contact_point1=torch.zeros((capacity, 3), dtype=f32, device=device),
        rigid_contact_normal=torch.zeros((capacity, 3), dtype=f32, device=device),
        rigid_contact_margin0=torch.zeros((capacity,), dtype=f32, device=device),
        rigid_contact_margin1=torch.zeros((capacity,), dtype=f32, device=device),
        rigid_contact_world_start=torch.zeros((world_count + 1,), dtype=i32, device=device),
        rigid_contact_max=capacity,
    )


Model = SapModel
State = SapState
Control = SapControl

__all__ = [
    "SapModel",
    "SapState",
    "SapControl",
    "SapContacts",
    "allocate_contacts",
    "SAP_JOINT_PRISMATIC",
    "SAP_JOINT_REVOLUTE",
    "SAP_JOINT_BALL",
    "SAP_JOINT_FIXED",
    "SAP_JOINT_FREE",
    "SAP_JOINT_DISTANCE",
    "SAP_JOINT_D6",
    "GEO_PLANE",
    "GEO_SPHERE",
    "GEO_CAPSULE",
    "GEO_ELLIPSOID",
    "GEO_CYLINDER",
    "GEO_BOX",
    "GEO_CONE",
]

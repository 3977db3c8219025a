"""``SapCollisionPipeline`` — drop-in for ``sim/collision/pipeline.py:312-709``.

``collide(state, contacts)`` computes shape AABBs, tests the candidate pairs
and runs the narrowphase (analytic primitives; MPR / GJK for capsule-box and
box-box) in one HIP kernel per world (``sapx_collide``), writing contacts
grouped by world in candidate-pair order together with per-world offsets
(``rigid_contact_world_start``) that ``SolverSAP.step`` consumes directly.

``reduce_contacts`` is accepted and has no effect: the reference applies it
only to mesh contacts on the GPU and switches it off for models without
meshes (``narrow_phase.py:1387-1388``); this pipeline has no mesh shapes.

Broadphase modes (``pipeline.py:281-292, 421-454``): ``"explicit"`` (the
default) tests the model's explicit pair list; ``"nxn"`` and ``"sap"`` test
every pair of a world's colliding shapes plus the global shapes that the
world / collision-group rules and the filter pairs allow.  The world, group
and filter tests do not depend on the state, so the library applies them once
when the pipeline is created and keeps the surviving pairs (lower-triangular
order per world) as a template; each ``collide`` tests their AABBs.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .runtime import SapContacts, SapModel, SapState, allocate_contacts


class SapCollisionPipeline:
    """Rigid contact generation for a ``SapModel`` (explicit or all-pairs broadphase + narrowphase)."""

    def __init__(
        self,
        model: SapModel,
        *,
        reduce_contacts: bool = True,
        rigid_contact_max: int | None = None,
        max_contacts_per_world: int | None = None,
        broad_phase: str | None = None,
    ):
        if not isinstance(model, SapModel):
            raise TypeError("SapCollisionPipeline requires SapModel.")
        mode = "explicit" if broad_phase is None else str(broad_phase).strip().lower()
        if mode not in ("explicit", "nxn", "sap"):
            raise ValueError(f"broad_phase must be one of ('nxn', 'sap', 'explicit'), got {broad_phase!r}")
        self.broad_phase_mode = mode
        self._device = nat.require_gpu_device(model.device, "SapCollisionPipeline")
        W = max(int(model.world_count), 1)
        if max_contacts_per_world is None:
            max_contacts_per_world = (int(rigid_contact_max) // W) if rigid_contact_max else 128
        if max_contacts_per_world <= 0:
            raise ValueError("max_contacts_per_world must be positive")
        self.model = model
        self.reduce_contacts = bool(reduce_contacts)
        self.max_contacts_per_world = int(max_contacts_per_world)
        self.rigid_contact_max = int(rigid_contact_max) if rigid_contact_max else self.max_contacts_per_world * W
        self._native_model = nat.NativeModel(model, self._device)
        self._lib = nat.load()
        h = ctypes.c_void_p()
        with nat.on_device(self._device):
            self._create(model, mode, h)
        self._handle = h

    def _create(self, model, mode, h):
        if mode == "explicit":
            nat.check(self._lib.sapx_pipeline_create(self._native_model.handle, self.max_contacts_per_world,
                                                     ctypes.byref(h)), "sapx_pipeline_create")
        else:
            groups = np.ascontiguousarray(np.asarray(model.shape_collision_group, dtype=np.int32).reshape(-1))
            flags = np.ascontiguousarray(np.asarray(model.shape_flags, dtype=np.int32).reshape(-1))
            fp = model.shape_collision_filter_pairs
            fp = np.zeros((0, 2), np.int32) if fp is None else np.asarray(fp, dtype=np.int32).reshape(-1, 2)
            fp = np.ascontiguousarray(fp)
            if groups.shape[0] != model.shape_count or flags.shape[0] != model.shape_count:
                raise ValueError("shape_collision_group and shape_flags need one entry per shape")
            ip = ctypes.POINTER(ctypes.c_int32)
            nat.check(self._lib.sapx_pipeline_create_all_pairs(
                self._native_model.handle, self.max_contacts_per_world, groups.ctypes.data_as(ip),
                flags.ctypes.data_as(ip), fp.ctypes.data_as(ip) if fp.shape[0] else None, int(fp.shape[0]),
                ctypes.byref(h)), "sapx_pipeline_create_all_pairs")

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.sapx_pipeline_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def contacts(self) -> SapContacts:
        """Allocate a contact buffer sized for this pipeline (pipeline.py:515-530)."""
        return allocate_contacts(self.rigid_contact_max, self.model.world_count, self.model.device)

    def collide(self, state: SapState, contacts: SapContacts) -> SapContacts:
        """Generate contacts for ``state.body_q`` into ``contacts`` (pipeline.py:546-709)."""
        import torch

        if not isinstance(contacts, SapContacts):
            raise TypeError("collide requires SapContacts")
        if contacts.rigid_contact_count is None:
            raise ValueError("contacts need rigid_contact_count; allocate them with pipeline.contacts()")
        nat.check_contacts(contacts, self._device, self.model.world_count, need_world_start=True)
        bq = state.body_q
        if (not isinstance(bq, torch.Tensor) or bq.dtype != torch.float32 or bq.numel() != self.model.body_count * 7
                or not bq.is_contiguous()):
            raise TypeError("state.body_q must be a contiguous float32 (body_count, 7) tensor")
        nat.require_on(bq, self._device, "state.body_q")
        view = nat.contacts_view(contacts)
        with nat.on_device(self._device):
            rc = self._lib.sapx_collide(self._handle, nat._ptr(bq), ctypes.byref(view), nat.stream_ptr())
        nat.check(rc, "sapx_collide")
        return contacts


__all__ = ["SapCollisionPipeline"]

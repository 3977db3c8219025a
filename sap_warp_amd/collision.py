"""``SapCollisionPipeline`` — drop-in for ``sim/collision/pipeline.py:312-709``.

``collide(state, contacts)`` computes shape AABBs, tests the model's explicit
candidate pairs and runs the primitive narrowphase in one HIP kernel per
world (``sapx_collide``), writing contacts grouped by world in candidate-pair
order together with per-world offsets (``rigid_contact_world_start``) that
``SolverSAP.step`` consumes directly.
"""
from __future__ import annotations

import ctypes

import numpy as np

from . import _native as nat
from .runtime import SapContacts, SapModel, SapState, allocate_contacts


class SapCollisionPipeline:
    """Rigid contact generation for a ``SapModel`` (explicit broadphase + primitive narrowphase)."""

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
        if broad_phase not in (None, "explicit"):
            raise NotImplementedError("only the explicit-pair broadphase is in the native pipeline")
        W = max(int(model.world_count), 1)
        if max_contacts_per_world is None:
            max_contacts_per_world = (int(rigid_contact_max) // W) if rigid_contact_max else 128
        if max_contacts_per_world <= 0:
            raise ValueError("max_contacts_per_world must be positive")
        self.model = model
        self.reduce_contacts = bool(reduce_contacts)
        self.max_contacts_per_world = int(max_contacts_per_world)
        self.rigid_contact_max = int(rigid_contact_max) if rigid_contact_max else self.max_contacts_per_world * W
        self._native_model = nat.NativeModel(model)
        self._lib = nat.load()
        h = ctypes.c_void_p()
        nat.check(self._lib.sapx_pipeline_create(self._native_model.handle, self.max_contacts_per_world,
                                                 ctypes.byref(h)), "sapx_pipeline_create")
        self._handle = h

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
        if contacts.rigid_contact_world_start is None:
            raise ValueError("contacts need rigid_contact_world_start; allocate them with pipeline.contacts()")
        bq = state.body_q
        if bq.dtype != torch.float32 or bq.numel() != self.model.body_count * 7 or not bq.is_contiguous():
            raise TypeError("state.body_q must be a contiguous float32 (body_count, 7) tensor")
        view = nat.contacts_view(contacts)
        nat.check(self._lib.sapx_collide(self._handle, nat._ptr(bq), ctypes.byref(view), nat.stream_ptr()),
                  "sapx_collide")
        return contacts


__all__ = ["SapCollisionPipeline"]

"""sapx_collide_noalloc (the collide entry point that allocates nothing: static device scratch)
against sapx_collide, through the C ABI: identical contact lists, bit for bit, on two calls in a
row, the second with a smaller input (fewer worlds, fewer contact slots)."""
from __future__ import annotations

import ctypes

import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def test_collide_noalloc_matches_collide(gpu_device):
    import torch

    from sap_warp_amd import _native as nat
    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.scenes import build_humanoid_scene

    lib = nat.load()
    for W, C, seed in ((64, 64, 0), (16, 32, 1)):
        print(f"noalloc case W={W} C={C} seed={seed}", flush=True)
        model = build_humanoid_scene(W, seed=seed, lying_fraction=0.5, contact_gap=0.03, device="cuda:0")
        pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
        ref, got = pipe.contacts(), pipe.contacts()
        for t in (got.rigid_contact_shape0, got.rigid_contact_point0, got.rigid_contact_world_start):
            t.fill_(-7)
        state = model.state()
        pipe.collide(state, ref)
        view = nat.contacts_view(got)
        with nat.on_device("cuda:0"):
            rc = lib.sapx_collide_noalloc(pipe._handle, nat._ptr(state.body_q), ctypes.byref(view), nat.stream_ptr())
        nat.check(rc, "sapx_collide_noalloc")
        torch.cuda.synchronize()
        n = int(_np(ref.rigid_contact_count)[0])
        assert n > 0
        assert int(_np(got.rigid_contact_count)[0]) == n
        np.testing.assert_array_equal(_np(got.rigid_contact_world_start), _np(ref.rigid_contact_world_start))
        for f in ("rigid_contact_shape0", "rigid_contact_shape1"):
            np.testing.assert_array_equal(_np(getattr(got, f))[:n], _np(getattr(ref, f))[:n], err_msg=f)
        for f in ("rigid_contact_point0", "rigid_contact_point1", "rigid_contact_normal", "rigid_contact_margin0",
                  "rigid_contact_margin1"):
            np.testing.assert_array_equal(_np(getattr(got, f))[:n].view(np.uint32),
                                          _np(getattr(ref, f))[:n].view(np.uint32), err_msg=f)

"""Host-side guards that run before any native call (CPU only).

* Models and tensors in host memory are rejected before a pointer reaches a kernel.
* Per-step statistics read-back and the multi-stream batch step refuse to run under a
  HIP graph capture (both would record illegal or unsupported work).
"""
import numpy as np
import pytest
import torch

from sap_warp_amd import _native as nat
from sap_warp_amd.runtime import SapContacts, allocate_contacts
from sap_warp_amd.scenes import build_humanoid_scene


@pytest.fixture(scope="module")
def cpu_model():
    return build_humanoid_scene(2, seed=1, device="cpu")


def test_solver_rejects_host_model(cpu_model):
    from sap_warp_amd.solver_sap import SolverSAP

    with pytest.raises(ValueError, match="GPU"):
        SolverSAP(cpu_model, max_rigid_contact=64)


def test_pipeline_rejects_host_model(cpu_model):
    from sap_warp_amd.collision import SapCollisionPipeline

    with pytest.raises(ValueError, match="GPU"):
        SapCollisionPipeline(cpu_model, max_contacts_per_world=64)


def test_require_on_rejects_other_device():
    dev = torch.device("cuda", 0)
    with pytest.raises(ValueError, match="cuda:0"):
        nat.require_on(torch.zeros(3), dev, "x")
    with pytest.raises(TypeError):
        nat.require_on(np.zeros(3), dev, "x")


def test_check_contacts_rejects_host_and_short_buffers():
    dev = torch.device("cuda", 0)
    c = allocate_contacts(32, 2, "cpu")
    with pytest.raises(ValueError):
        nat.check_contacts(c, dev, 2)
    # capacity larger than the arrays: rejected even on the right device (checked on CPU here)
    short = SapContacts(**{**c.__dict__, "rigid_contact_max": 64})
    with pytest.raises(ValueError, match="capacity"):
        nat.check_contacts(short, torch.device("cpu"), 2)
    bad_ws = SapContacts(**{**c.__dict__, "rigid_contact_world_start": torch.zeros(2, dtype=torch.int32)})
    with pytest.raises(ValueError, match="world_count"):
        nat.check_contacts(bad_ws, torch.device("cpu"), 2)
    nat.check_contacts(c, torch.device("cpu"), 2)  # consistent bundle: accepted


def test_stats_readback_refused_under_capture(cpu_model, monkeypatch):
    from sap_warp_amd.solver_sap import SolverSAP

    solver = object.__new__(SolverSAP)
    solver.collect_iteration_stats = True
    monkeypatch.setattr(torch.cuda, "is_current_stream_capturing", lambda: True)
    s = cpu_model.state()
    with pytest.raises(RuntimeError, match="capture"):
        solver.step(s, cpu_model.state(), cpu_model.control(), None, 2.0e-3)


def test_hetero_batch_refused_under_capture(monkeypatch):
    """The streams mode (nested fork) raises under capture; the batch mode raises under capture
    until both slots are bound, and when statistics are collected."""
    from sap_warp_amd.hetero import HeteroBatch

    batch = object.__new__(HeteroBatch)
    batch.groups = {}
    batch.dt = 2.0e-3
    batch.mode = "streams"
    monkeypatch.setattr(torch.cuda, "is_current_stream_capturing", lambda: True)
    with pytest.raises(RuntimeError, match="captured"):
        batch.step()
    batch.mode = "batch"
    batch._k = 0
    batch._bound = [None, None]
    batch._batch = object()  # a batch launch exists but no slot is bound yet
    batch.batched, batch.own = [], []
    with pytest.raises(RuntimeError, match="before capturing"):
        batch.step()

    class _Solver:
        collect_iteration_stats = True

    batch.groups = {"kind": {"solver": _Solver()}}
    with pytest.raises(RuntimeError, match="collect_iteration_stats"):
        batch.step()

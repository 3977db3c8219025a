"""Autotuner of the step's work schedule (no counterpart in the reference).

The native step can map a batch of worlds onto the GPU in several ways that all
compute the same step (``SapxSchedule`` in ``include/sapx.h``):

* ``kernel``: one 64-lane wave per world, or one 256-thread workgroup per world;
* ``small_capacity``: the contact slots (16 / 32 / 64) of the wave variant that every
  world with at most that many contacts runs in (more slots = more LDS per world =
  fewer worlds per CU); the 64-slot variant takes the rest;
* ``fork``: launch the 64-slot variant on a second stream beside the small one;
* ``world_order``: dispatch the worlds most-expensive-first.

Which one is fastest depends on the model (dofs, bodies) and on the contact-count
distribution of the batch, so ``autotune`` times every schedule the model admits on
copies of the caller's state (collision + step, a few warmup and timed steps each,
HIP events on the current stream) and keeps the fastest.
"""
from __future__ import annotations


def candidate_schedules(max_contacts: int):
    """Distinct schedules for a contact capacity (options that cannot matter are pinned)."""
    out = [{"kernel": "fused", "small_capacity": 64, "fork": False, "world_order": False}]
    single = min(c for c in (16, 32, 64) if c >= min(max_contacts, 64))  # smallest variant holding every world
    for small in (16, 32, 64):
        if small > single:
            continue
        for order in (True, False):
            if small == single:
                out.append({"kernel": "wave", "small_capacity": small, "fork": False, "world_order": order})
            else:
                for fork in (True, False):
                    out.append({"kernel": "wave", "small_capacity": small, "fork": fork, "world_order": order})
    return out


def _robustness(sc) -> tuple:
    """Preference among near-ties: the schedules that degrade least when contact counts grow."""
    return (sc["kernel"] == "wave", sc["fork"], sc["world_order"], -sc["small_capacity"])


def autotune(solver, pipeline, state, control, contacts, dt, steps: int = 6, warmup: int = 2,
             candidates=None, keep: bool = True, tie: float = 0.03):
    """Time the candidate schedules from ``state`` (left unchanged) and set the fastest.

    Candidates within ``tie`` (relative) of the fastest count as ties and are resolved towards
    the schedule that copes best with more contacts than the sample shows (the wave kernel,
    the forked 64-slot launch, most-expensive-first order, fewer slots): a few steps from one
    state do not show how the contact counts of the batch evolve.
    Returns ``{"best": schedule, "ms_per_step": {description: ms}}``; schedules the model
    does not admit are skipped.  The solver's warm-start guess is reset before and after.
    """
    import torch

    cands = candidates if candidates is not None else candidate_schedules(solver.max_rigid_contact)
    model = solver.model
    a, b = model.state(), model.state()
    original = solver.get_schedule()
    times = {}
    measured = []
    for sc in cands:
        try:
            solver.set_schedule(**sc)
        except NotImplementedError:
            continue
        for k in ("joint_q", "joint_qd", "body_q", "body_qd", "body_f"):
            getattr(a, k).copy_(getattr(state, k))
        solver.reset_runtime_state()
        s0, s1 = a, b
        for _ in range(warmup):
            pipeline.collide(s0, contacts)
            solver.step(s0, s1, control, contacts, dt)
            s0, s1 = s1, s0
        e0 = torch.cuda.Event(enable_timing=True)
        e1 = torch.cuda.Event(enable_timing=True)
        e0.record()
        for _ in range(steps):
            pipeline.collide(s0, contacts)
            solver.step(s0, s1, control, contacts, dt)
            s0, s1 = s1, s0
        e1.record()
        torch.cuda.synchronize()
        ms = e0.elapsed_time(e1) / steps
        key = (f"{sc['kernel']}/small{sc['small_capacity']}/{'fork' if sc['fork'] else 'serial'}/"
               f"{'ordered' if sc['world_order'] else 'identity'}")
        times[key] = round(ms, 4)
        measured.append((ms, dict(sc)))
    best = None
    if measured:
        fastest = min(ms for ms, _ in measured)
        near = [sc for ms, sc in measured if ms <= fastest * (1.0 + tie)]
        best = max(near, key=_robustness)
    solver.reset_runtime_state()
    if keep and best is not None:
        solver.set_schedule(**best)
    else:
        solver.set_schedule(**original)
    return {"best": best if best is not None else original, "ms_per_step": times}

"""One world of free bodies on the whole GPU (``sapx_bigworld_*``, ``BigWorldSolver``) against the
oracle.

World: a bin of free boxes, spheres and capsules stacked in columns (``build_bin_scene``), a few
steps of it by the oracle's C++ step (so the columns are collapsing: resting, sliding and impact
contacts), then:

* collide: the contact list of the model's explicit pairs is bit-identical to the oracle's
  (``oracle/narrowphase.c``: count, shapes, body-frame points, normals, margins), including a
  capacity that cuts the list (the per-world overflow rule);
* step: two steps (the second warm-started from the first's solution), each from the same
  state, contacts and start as the oracle's float64 step (``sap_oracle.step``): the SAP
  velocity within 1e-5 relative (max-norm), the
  public joint_q / joint_qd / body_q / body_qd within 1e-5, contact impulses within 1e-4 (the
  redundant pile contacts make gamma ill-conditioned; v is not), for both line searches and the
  drake preset (Delassus weights, contact midpoint, sap_euler);
* the full 1024-body bin of the bench: contacts bit-identical at the start and after 10 GPU steps;
* a contact list longer than max_rigid_contact: the step keeps the first C contacts, as the oracle;
* models outside the path (articulated joints) are rejected.
"""
import numpy as np
import pytest

from helpers import world_rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _bin(bodies, seed, steps, C, preset="approx32", ls="exact_root"):
    from oracle import collide as oc
    from oracle import cpu_step as cs
    from oracle import sap_oracle as so
    from sap_warp_amd import hetero

    model = hetero.build_bin_scene(bodies, seed=seed, columns=3, device="cuda:0")
    dt = 1.0e-3
    opts = so.StepOptions(preset=preset, line_search=ls, max_rigid_contact=C, solve=so.SolveOptions.for_variant(ls))
    stepper = cs.CpuStepper(model, opts)
    state = dict(joint_q=model.joint_q, joint_qd=model.joint_qd, body_q=model.body_q, body_qd=model.body_qd,
                 body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=model.joint_f, joint_target_pos=model.joint_target_pos,
                   joint_target_vel=model.joint_target_vel, joint_act=model.joint_act)
    for _ in range(steps):
        c = oc.collide(model, state["body_q"], C, threads=8)
        out = stepper.step(state, control, c, dt, threads=8)
        state = dict(joint_q=out["joint_q"], joint_qd=out["joint_qd"], body_q=out["body_q"], body_qd=out["body_qd"],
                     body_f=state["body_f"])
    return model, state, control, dt, opts


def _gpu_state(model, st):
    import torch

    s = model.state()
    for k in ("joint_q", "joint_qd", "body_q", "body_qd"):
        getattr(s, k).copy_(torch.as_tensor(np.asarray(st[k], dtype=np.float32)))
    return s


@pytest.mark.parametrize("C", [4000, 60])
def test_bigworld_collide_bit_exact(gpu_device, C):
    from oracle import collide as oc
    from sap_warp_amd.large_world import BigWorldSolver

    print(f"bigworld collide C {C}", flush=True)
    model, st, _, _, _ = _bin(60, seed=5, steps=40, C=4000)
    solver = BigWorldSolver(model, max_rigid_contact=C, line_search_variant="exact_root", contact_tau_d=0.005)
    contacts = solver.allocate_contacts()
    s = _gpu_state(model, st)
    solver.collide(s, contacts)
    ref = oc.collide(model, np.asarray(st["body_q"], np.float32), C)
    n = int(_np(contacts.rigid_contact_count)[0])
    print(f"contacts gpu {n} oracle {ref['count']}", flush=True)
    assert n == ref["count"] and n > 0
    assert n <= C
    np.testing.assert_array_equal(_np(contacts.rigid_contact_world_start), [0, n])
    for key, field in (("shape0", "rigid_contact_shape0"), ("shape1", "rigid_contact_shape1"),
                       ("point0", "rigid_contact_point0"), ("point1", "rigid_contact_point1"),
                       ("normal", "rigid_contact_normal"), ("margin0", "rigid_contact_margin0"),
                       ("margin1", "rigid_contact_margin1")):
        np.testing.assert_array_equal(_np(getattr(contacts, field))[:n], ref[key][:n], err_msg=key)
    solver.close()


def test_bigworld_collide_full_bin(gpu_device):
    """The bench's 1024-body bin (528 896 candidate pairs): contacts bit-identical to the oracle's
    at the initial stacking and after 10 GPU steps of the collapsing columns."""
    import torch

    from oracle import collide as oc
    from sap_warp_amd import hetero
    from sap_warp_amd.large_world import BigWorldSolver

    print("bigworld collide full bin", flush=True)
    model = hetero.build_bin_scene(1024, seed=17, device="cuda:0")
    C = 16384
    solver = BigWorldSolver(model, max_rigid_contact=C, line_search_variant="exact_root", contact_tau_d=0.005)
    contacts = solver.allocate_contacts()
    s0, s1 = model.state(), model.state()
    ctl = model.control()
    for k in range(11):
        solver.collide(s0, contacts)
        torch.cuda.synchronize()
        if k in (0, 10):
            ref = oc.collide(model, _np(s0.body_q), C, threads=8)
            n = int(_np(contacts.rigid_contact_count)[0])
            print(f"step {k}: contacts gpu {n} oracle {ref['count']}", flush=True)
            assert n == ref["count"] and n > 1000
            for key in ("shape0", "shape1", "point0", "point1", "normal", "margin0", "margin1"):
                np.testing.assert_array_equal(_np(getattr(contacts, "rigid_contact_" + key))[:n], ref[key][:n],
                                              err_msg=f"step {k} {key}")
        if k < 10:
            solver.step(s0, s1, ctl, contacts, 1.0e-3)
            s0, s1 = s1, s0
    solver.close()


@pytest.mark.parametrize("preset,ls", [("approx32", "exact_root"), ("approx32", "armijo_decay"),
                                       ("drake", "exact_root")])
def test_bigworld_step_matches_oracle(gpu_device, preset, ls):
    """Two steps: the first solve starts from v0, the second from the first step's solution (the
    solver's warm start), each compared with the oracle step from the same input state, contacts
    and start."""
    import torch

    from oracle import sap_oracle as so
    from sap_warp_amd.large_world import BigWorldSolver

    C = 4000
    print(f"bigworld step {preset} {ls}", flush=True)
    model, st, ctl, dt, opts = _bin(60, seed=7, steps=60, C=C, preset=preset, ls=ls)
    solver = BigWorldSolver(model, max_rigid_contact=C, line_search_variant=ls, contact_preset_variant=preset,
                            contact_tau_d=0.005)
    contacts = solver.allocate_contacts()
    s0, s1 = _gpu_state(model, st), model.state()
    n = model.joint_dof_count
    v = torch.zeros(n, dtype=torch.float64, device="cuda:0")
    gamma = torch.zeros((C, 3), dtype=torch.float64, device="cuda:0")
    guess = None
    for k in range(2):
        st_in = dict(joint_q=_np(s0.joint_q), joint_qd=_np(s0.joint_qd), body_q=_np(s0.body_q),
                     body_qd=_np(s0.body_qd), body_f=np.zeros((model.body_count, 6), np.float32))
        solver.collide(s0, contacts)
        solver.step(s0, s1, model.control(), contacts, dt, v_out=v, gamma_out=gamma)
        torch.cuda.synchronize()
        nc = int(_np(contacts.rigid_contact_count)[0])
        c_np = dict(count=nc, world_start=_np(contacts.rigid_contact_world_start))
        for key in ("shape0", "shape1", "point0", "point1", "normal", "margin0", "margin1"):
            c_np[key] = _np(getattr(contacts, "rigid_contact_" + key))
        ref = so.step(model, st_in, ctl, c_np, dt, opts, v_guess=guess)
        r = ref["per_world"][0]
        stt = solver.solve_stats()
        e = dict(v=world_rel_err(_np(v), r["v_sap"]), q=world_rel_err(_np(s1.joint_q), ref["joint_q"]),
                 qd=world_rel_err(_np(s1.joint_qd), ref["joint_qd"]), bq=world_rel_err(_np(s1.body_q), ref["body_q"]),
                 bqd=world_rel_err(_np(s1.body_qd), ref["body_qd"]),
                 gamma=world_rel_err(_np(gamma)[:nc], r["result"].gamma))
        print(f"step {k} nc {nc} gpu it {stt.iterations} conv {stt.converged} | oracle it {r['result'].iterations} | "
              + " ".join(f"{name} {x:.2e}" for name, x in e.items()), flush=True)
        assert nc >= 60
        assert r["contacts"].count == nc
        assert stt.converged and r["result"].converged
        for name in ("v", "q", "qd", "bq", "bqd"):
            assert e[name] <= 1.0e-5, (k, name, e[name])
        assert e["gamma"] <= 1.0e-4
        guess = _np(v).astype(np.float64)
        s0, s1 = s1, s0
    solver.close()


def test_bigworld_step_capacity_cut(gpu_device):
    """A contact list longer than the solver's max_rigid_contact: the step keeps the first C
    contacts (contact_jacobian.py:470-472), as the oracle step with the same capacity does."""
    import torch

    from oracle import sap_oracle as so
    from sap_warp_amd.large_world import BigWorldSolver

    print("bigworld step capacity cut", flush=True)
    model, st, ctl, dt, _ = _bin(60, seed=11, steps=80, C=4000)
    big = BigWorldSolver(model, max_rigid_contact=4000, line_search_variant="exact_root", contact_tau_d=0.005)
    contacts = big.allocate_contacts()
    s0, s1 = _gpu_state(model, st), model.state()
    big.collide(s0, contacts)
    torch.cuda.synchronize()
    nc = int(_np(contacts.rigid_contact_count)[0])
    C = nc - 25
    assert C > 40
    solver = BigWorldSolver(model, max_rigid_contact=C, line_search_variant="exact_root", contact_tau_d=0.005)
    v = torch.zeros(model.joint_dof_count, dtype=torch.float64, device="cuda:0")
    solver.step(s0, s1, model.control(), contacts, dt, v_out=v)
    torch.cuda.synchronize()
    c_np = dict(count=nc, world_start=_np(contacts.rigid_contact_world_start))
    for key in ("shape0", "shape1", "point0", "point1", "normal", "margin0", "margin1"):
        c_np[key] = _np(getattr(contacts, "rigid_contact_" + key))
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=C,
                          solve=so.SolveOptions.for_variant("exact_root"))
    ref = so.step(model, st, ctl, c_np, dt, opts)
    r = ref["per_world"][0]
    e = dict(v=world_rel_err(_np(v), r["v_sap"]), q=world_rel_err(_np(s1.joint_q), ref["joint_q"]))
    print(f"nc {nc} C {C} oracle slots {r['contacts'].count} | v {e['v']:.2e} q {e['q']:.2e}", flush=True)
    assert r["contacts"].count == C
    assert solver.solve_stats().converged and r["result"].converged
    assert e["v"] <= 1.0e-5 and e["q"] <= 1.0e-5
    solver.close()
    big.close()


def test_bigworld_rejects_articulated_model(gpu_device):
    from sap_warp_amd.large_world import BigWorldSolver
    from sap_warp_amd.scenes import build_humanoid_scene

    model = build_humanoid_scene(1, seed=0, device="cuda:0")
    with pytest.raises(NotImplementedError):
        BigWorldSolver(model, max_rigid_contact=64, line_search_variant="exact_root", contact_tau_d=0.005)

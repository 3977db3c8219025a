"""GPU collision parity: contact lists must match the float32 oracle bit for bit."""
import numpy as np
import pytest

from tests.helpers import oracle_contacts_dict

pytestmark = pytest.mark.gpu


def _scene(name, worlds, seed):
    from sap_warp_amd.scenes import build_ant_scene, build_box_pile_scene, build_humanoid_scene

    if name == "ant":
        return build_ant_scene(worlds, seed=seed, device="cuda:0")
    if name == "boxpile":
        return build_box_pile_scene(worlds, seed=seed, device="cuda:0")
    if name == "round":
        from sap_warp_amd.scenes import build_round_shapes_scene

        return build_round_shapes_scene(worlds, seed=seed + 9, device="cuda:0")
    if name == "convex":
        from sap_warp_amd.scenes import build_convex_shapes_scene

        return build_convex_shapes_scene(worlds, seed=seed + 11, device="cuda:0")
    if name == "humanoid_lying":
        return build_humanoid_scene(worlds, seed=seed, lying_fraction=1.0, contact_gap=0.05, device="cuda:0")
    return build_humanoid_scene(worlds, seed=seed, device="cuda:0")


def _jittered_body_q(model, seed):
    """Every body moved by U(-5, 5) mm and turned by up to 3 degrees (pile contacts stay in touch)."""
    from sap_warp_amd.scenes import quat_from_axis_angle, quat_mul

    rng = np.random.default_rng(seed)
    bq = np.asarray(model.body_q, dtype=np.float64).copy()
    for b in range(bq.shape[0]):
        bq[b, :3] += rng.uniform(-0.005, 0.005, 3)
        dq = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, 0.05))
        q = quat_mul(dq, bq[b, 3:])
        bq[b, 3:] = q / np.linalg.norm(q)
    return bq.astype(np.float32)


def _perturbed_body_q(model, seed):
    """Random joint configurations near the ground -> FK body poses (host)."""
    from sap_warp_amd.kinematics_host import eval_fk_public

    rng = np.random.default_rng(seed)
    q = model.joint_q.astype(np.float64).copy()
    for j in range(model.joint_count):
        qs = int(model.joint_q_start[j])
        if int(model.joint_type[j]) == 4:
            q[qs + 2] = rng.uniform(0.05, 1.0)
            quat = rng.normal(size=4)
            q[qs + 3:qs + 7] = quat / np.linalg.norm(quat)
        elif int(model.joint_type[j]) == 1:
            q[qs] = rng.uniform(-1.5, 1.5)
    bq, _ = eval_fk_public(model, q.astype(np.float32), model.joint_qd)
    return bq.astype(np.float32)


def _check_equal(gpu, ref, C, model):
    import torch

    torch.cuda.synchronize()
    count = int(gpu.rigid_contact_count.cpu()[0])
    assert count == ref["count"]
    ws = gpu.rigid_contact_world_start.cpu().numpy()
    np.testing.assert_array_equal(ws, ref["world_start"])
    n = count
    for name, key in (("rigid_contact_shape0", "shape0"), ("rigid_contact_shape1", "shape1")):
        np.testing.assert_array_equal(getattr(gpu, name).cpu().numpy()[:n], ref[key][:n])
    for name, key in (("rigid_contact_point0", "point0"), ("rigid_contact_point1", "point1"),
                      ("rigid_contact_normal", "normal"), ("rigid_contact_margin0", "margin0"),
                      ("rigid_contact_margin1", "margin1")):
        g = getattr(gpu, name).cpu().numpy()[:n]
        r = ref[key][:n]
        # bit-exact float32
        np.testing.assert_array_equal(g.view(np.uint32), r.view(np.uint32), err_msg=name)


@pytest.mark.parametrize("scene,worlds,C", [("ant", 16, 16), ("humanoid", 16, 64), ("boxpile", 32, 64),
                                            ("round", 32, 16), ("humanoid_lying", 16, 64), ("convex", 32, 32)])
@pytest.mark.parametrize("seed", [0, 1, 2])
def test_collide_bit_exact(gpu_device, scene, worlds, C, seed):
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene(scene, worlds, seed)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts = pipe.contacts()
    state = model.state()
    for variant in range(2):
        if variant == 0:
            bq = model.body_q
        elif scene in ("boxpile", "round", "convex"):
            bq = _jittered_body_q(model, 100 + seed)
        else:
            bq = _perturbed_body_q(model, 100 + seed)
        print(f"collide {scene} seed {seed} variant {variant}", flush=True)
        state.body_q.copy_(torch.as_tensor(bq, device=gpu_device))
        pipe.collide(state, contacts)
        ref = oracle_contacts_dict(model, bq, C)
        assert ref["unsupported"] == 0
        _check_equal(contacts, ref, C, model)


def test_collide_convex_pairs_present(gpu_device):
    """The pile's contact list holds box-box and capsule-box contacts (the MPR / GJK path)."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene("boxpile", 8, 4)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=64)
    contacts = pipe.contacts()
    state = model.state()
    pipe.collide(state, contacts)
    torch.cuda.synchronize()
    n = int(contacts.rigid_contact_count.cpu()[0])
    st = np.asarray(model.shape_type)
    t0 = st[contacts.rigid_contact_shape0.cpu().numpy()[:n]]
    t1 = st[contacts.rigid_contact_shape1.cpu().numpy()[:n]]
    kinds = set(zip(t0.tolist(), t1.tolist()))
    assert (7, 7) in kinds and (4, 7) in kinds, kinds


def test_collide_convex_shape_pairs_present(gpu_device):
    """The convex-shapes scene reaches every pair kind it is built for: plane-cone through the box
    proxy, box-cone, box-cylinder, ellipsoid-ellipsoid, sphere-ellipsoid and capsule-cylinder on the
    MPR / GJK path, plus the plane primitives."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene("convex", 8, 0)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=32)
    contacts = pipe.contacts()
    pipe.collide(model.state(), contacts)
    torch.cuda.synchronize()
    n = int(contacts.rigid_contact_count.cpu()[0])
    st = np.asarray(model.shape_type)
    t0 = st[contacts.rigid_contact_shape0.cpu().numpy()[:n]]
    t1 = st[contacts.rigid_contact_shape1.cpu().numpy()[:n]]
    kinds = set(zip(t0.tolist(), t1.tolist()))
    for k in ((1, 9), (7, 9), (6, 7), (5, 5), (3, 5), (4, 6), (1, 5), (1, 6), (1, 7)):
        assert k in kinds, (k, kinds)


@pytest.mark.parametrize("worlds", [33, 64, 96, 100, 4096])
def test_collide_count_closes_list(gpu_device, worlds):
    """count and world_start[W] equal the total of the per-world counts on every call, for world
    counts whose last world also falls in the upper half of the previous placement workgroup's
    wave (a regression test: those lanes once raced the owning workgroup's write of the total)."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene("humanoid", worlds, 5)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=64)
    contacts = pipe.contacts()
    state = model.state()
    ref = oracle_contacts_dict(model, model.body_q, 64) if worlds <= 128 else None
    for rep in range(20):
        pipe.collide(state, contacts)
        torch.cuda.synchronize()
        ws = contacts.rigid_contact_world_start.cpu().numpy()
        count = int(contacts.rigid_contact_count.cpu()[0])
        assert count == ws[-1], rep
        assert np.all(np.diff(ws) >= 0), rep
        if ref is not None:
            np.testing.assert_array_equal(ws, ref["world_start"])
        elif rep == 0:
            first = ws.copy()
        else:
            np.testing.assert_array_equal(ws, first)


def test_collide_overflow_rule(gpu_device):
    """A tiny per-world capacity drops whole pair batches, identically to the oracle."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene("humanoid", 8, 3)
    bq = _perturbed_body_q(model, 7)
    C = 5
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts = pipe.contacts()
    state = model.state()
    state.body_q.copy_(torch.as_tensor(bq, device=gpu_device))
    pipe.collide(state, contacts)
    ref = oracle_contacts_dict(model, bq, C)
    assert np.all(np.diff(ref["world_start"]) <= C)
    _check_equal(contacts, ref, C, model)


@pytest.mark.parametrize("C", [3, 7])
def test_collide_overflow_rule_convex(gpu_device, C):
    """Box pile at a tiny capacity: primitive batches are dropped whole, convex-pair contacts one by
    one (each claims a slot and is written while the slot fits), identically to the oracle."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene("boxpile", 16, 6)
    bq = _jittered_body_q(model, 9)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts = pipe.contacts()
    state = model.state()
    state.body_q.copy_(torch.as_tensor(bq, device=gpu_device))
    pipe.collide(state, contacts)
    ref = oracle_contacts_dict(model, bq, C)
    assert np.all(np.diff(ref["world_start"]) <= C)
    _check_equal(contacts, ref, C, model)


@pytest.mark.parametrize("scene,worlds,C", [("humanoid", 16, 64), ("boxpile", 16, 64)])
@pytest.mark.parametrize("mode", ["nxn", "sap"])
def test_collide_all_pairs_broadphase(gpu_device, scene, worlds, C, mode):
    """broad_phase="nxn" / "sap": bit-exact vs the oracle fed its all-pairs enumeration, and per
    world the same contacts as the explicit-pair pipeline (in the all-pairs order)."""
    import torch

    from oracle.collide import all_pairs, collide
    from sap_warp_amd.collision import SapCollisionPipeline

    model = _scene(scene, worlds, 1)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C, broad_phase=mode)
    explicit = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts, contacts_e = pipe.contacts(), explicit.contacts()
    state = model.state()
    pairs = all_pairs(model)
    for variant in range(2):
        if variant == 0:
            bq = model.body_q
        elif scene == "boxpile":
            bq = _jittered_body_q(model, 30)
        else:
            bq = _perturbed_body_q(model, 30)
        print(f"all-pairs {mode} {scene} variant {variant}", flush=True)
        state.body_q.copy_(torch.as_tensor(bq, device=gpu_device))
        pipe.collide(state, contacts)
        explicit.collide(state, contacts_e)
        ref = collide(model, np.asarray(bq, dtype=np.float32), C, pairs=pairs)
        assert ref["unsupported"] == 0
        _check_equal(contacts, ref, C, model)
        torch.cuda.synchronize()
        ws = contacts.rigid_contact_world_start.cpu().numpy()
        we = contacts_e.rigid_contact_world_start.cpu().numpy()
        np.testing.assert_array_equal(np.diff(ws), np.diff(we))
        def rows(c):
            s0, s1 = c.rigid_contact_shape0.cpu().numpy(), c.rigid_contact_shape1.cpu().numpy()
            p0, nn = c.rigid_contact_point0.cpu().numpy(), c.rigid_contact_normal.cpu().numpy()
            return [(int(s0[i]), int(s1[i])) + tuple(p0[i].tolist()) + tuple(nn[i].tolist()) for i in range(len(s0))]

        ra, rb = rows(contacts), rows(contacts_e)
        for w in range(model.world_count):
            assert sorted(ra[ws[w]:ws[w + 1]]) == sorted(rb[we[w]:we[w + 1]])


def test_collide_small_contact_buffer(gpu_device):
    """A contact buffer with fewer slots than worlds x per-world capacity: the list is the oracle's
    world-grouped list cut at the buffer's capacity, every world offset is clamped into the buffer,
    and the step reads only inside it."""
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.runtime import allocate_contacts
    from sap_warp_amd.solver_sap import SolverSAP

    model = _scene("humanoid", 8, 3)
    bq = _perturbed_body_q(model, 11)
    C = 64
    ref = oracle_contacts_dict(model, bq, C)
    cap = max(1, ref["count"] // 2)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    contacts = allocate_contacts(cap, model.world_count, gpu_device)
    state = model.state()
    state.body_q.copy_(torch.as_tensor(bq, device=gpu_device))
    print(f"small buffer: {ref['count']} contacts into {cap} slots", flush=True)
    pipe.collide(state, contacts)
    torch.cuda.synchronize()
    assert int(contacts.rigid_contact_count.cpu()[0]) == cap
    ws = contacts.rigid_contact_world_start.cpu().numpy()
    np.testing.assert_array_equal(ws, np.minimum(ref["world_start"], cap))
    for name, key in (("rigid_contact_shape0", "shape0"), ("rigid_contact_shape1", "shape1")):
        np.testing.assert_array_equal(getattr(contacts, name).cpu().numpy()[:cap], ref[key][:cap])
    g = contacts.rigid_contact_point0.cpu().numpy()[:cap]
    np.testing.assert_array_equal(g.view(np.uint32), ref["point0"][:cap].view(np.uint32))
    solver = SolverSAP(model, max_rigid_contact=C, line_search_variant="exact_root", collect_iteration_stats=True)
    s1 = model.state()
    solver.step(state, s1, model.control(), contacts, 2.0e-3)
    nc = solver.last_contact_solve_result.contact_count_env
    np.testing.assert_array_equal(nc, np.diff(np.minimum(ref["world_start"], cap)))
    assert np.all(np.isfinite(s1.joint_q.cpu().numpy()))

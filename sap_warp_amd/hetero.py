"""Heterogeneous batches: worlds of several kinds stepped together on one GPU.

The reference's step needs equally shaped worlds (``contact_jacobian.py:1273-1281``),
so a batch of mixed robots has no reference counterpart.  Here a batch is a set
of homogeneous groups — one ``SapModel`` per world kind, each with its own
solver, collision pipeline and states — stepped in one call, in one of two
modes:

* ``mode="batch"`` (default): each group's collide on the caller's stream, then
  one step launch over every world of every group (``sapx_batch_step``: one wave
  per world, each group with its own kernel parameters; a second launch of the
  64-slot contact variant for the worlds above the small capacity).  This mode
  can be captured in a HIP graph (after one eager step).
* ``mode="streams"``: every group's collide + step on its own stream (default
  priority) forked from the caller's stream and joined back.  Eager only: one
  HIP-graph capture of this nested fork ended in a segmentation fault inside
  the runtime's end-of-capture, with no cause shown, so capturing it raises.

World kinds (the heterogeneous workload of the benchmark plan):

* ``arm_gripper`` — a 7-DOF fixed-base arm with a 2-finger parallel gripper
  (9 DOF) above a free box (6 DOF) of side U(0.05, 0.1) m: 15 DOF;
* ``quadruped`` — a floating box torso with 4 legs of 3 revolute joints: 18 DOF;
* ``humanoid`` — ``scenes.build_humanoid_scene``: 35 DOF;
* ``pile`` — 8 free boxes / spheres of size U(0.05, 0.2) m dropped on the
  ground: 48 DOF; with ``build_hetero_models(pile_sizes=(8, 16, 32))`` the pile
  worlds split into groups of 8, 16 and 32 bodies (``pile``, ``pile16``,
  ``pile32``: 48, 96 and 192 DOF, 8 contact slots per body above 8 bodies).

Groups the batch launch cannot hold (above 48 dofs or 64 contact slots: the large
piles, which run the 256-thread kernel with their matrices in device memory) are
stepped by their own solver inside the same call, on the caller's stream, so the
whole batch step still captures into one graph.
"""
from __future__ import annotations

import math

import numpy as np

from .runtime import GEO_BOX, GEO_CAPSULE, GEO_SPHERE, SAP_JOINT_PRISMATIC, SAP_JOINT_REVOLUTE
from .scenes import (DofCfg, SceneBuilder, ShapeCfg, _set_free_joint_state, _with_consistent_body_q,
                     build_humanoid_scene, quat_from_axis_angle, quat_mul, quat_z_to, tf, tf_mul)

WORLD_KINDS = ("arm_gripper", "quadruped", "humanoid", "pile")


def _link_inertia(mass: float, r: float = 0.03) -> np.ndarray:
    return np.eye(3) * (0.4 * mass * r * r)


def build_quadruped_scene(num_worlds: int = 64, seed: int = 11, *, contact_gap: float = 0.01,
                          device: str = "cpu"):
    """Quadrupeds: floating box torso (0.5 x 0.24 x 0.12 m) + 4 legs x (hip abduction about x,
    hip flexion about y, knee about y) = 18 DOF, 13 bodies.

    Thighs and shins are capsules (r 0.035 / 0.03, 0.2 m) with a sphere foot; PD position drives
    (ke 150, kd 4, effort 100) hold the initial crouch; limits +-1.5 rad; armature 0.01.  Half of
    the worlds start standing, half fall from U(0.6, 1.0) m with tilted bases and random joint
    angles and velocities.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    cfg = ShapeCfg(density=600.0, ke=2.0e5, tau=0.005, mu=0.9, gap=contact_gap)
    b.world_count = num_worlds
    hips = [(0.2, 0.1), (0.2, -0.1), (-0.2, 0.1), (-0.2, -0.1)]
    stand = (0.0, 0.6, -1.2)
    for w in range(num_worlds):
        b.current_world = w
        falling = w % 2 == 1
        yaw = rng.uniform(-math.pi, math.pi)
        q = quat_from_axis_angle((0, 0, 1), yaw)
        if falling:
            q = quat_mul(q, quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, 0.8)))
            h = rng.uniform(0.6, 1.0)
        else:
            h = 0.02 + 0.2 * math.cos(0.6) * 2 + 0.035 + 0.005
        root = tf((0.0, 0.0, h), q)
        torso = b.add_body(root)
        torso_shape = b.add_shape(torso, GEO_BOX, tf(), (0.25, 0.12, 0.06), cfg)
        joints = [b.add_free_joint(torso)]
        angles = []
        for hx, hy in hips:
            hip_x = tf((hx, hy, -0.02))
            abd = b.add_body(tf_mul(root, hip_x), mass=0.3, inertia=_link_inertia(0.3))
            dof = lambda axis: DofCfg(axis=axis, target_ke=150.0, target_kd=4.0, actuator_mode=1, limit_lower=-1.5,
                                      limit_upper=1.5, limit_ke=1.0e4, limit_kd=10.0, armature=0.01,
                                      effort_limit=100.0)
            joints.append(b.add_joint(torso, abd, SAP_JOINT_REVOLUTE, angular_axes=[dof((1, 0, 0))],
                                      parent_xform=hip_x))
            thigh = b.add_body(tf_mul(root, hip_x))
            thigh_shape = b.add_shape(thigh, GEO_CAPSULE, tf((0.0, 0.0, -0.1)), (0.035, 0.1, 0.0), cfg)
            b._filter(torso_shape, thigh_shape)  # the hip sits inside the torso box
            joints.append(b.add_joint(abd, thigh, SAP_JOINT_REVOLUTE, angular_axes=[dof((0, 1, 0))]))
            shin = b.add_body(tf_mul(root, tf_mul(hip_x, tf((0.0, 0.0, -0.2)))))
            b.add_shape(shin, GEO_CAPSULE, tf((0.0, 0.0, -0.1)), (0.03, 0.1, 0.0), cfg)
            b.add_shape(shin, GEO_SPHERE, tf((0.0, 0.0, -0.2)), (0.035, 0.0, 0.0), cfg)
            joints.append(b.add_joint(thigh, shin, SAP_JOINT_REVOLUTE, angular_axes=[dof((0, 1, 0))],
                                      parent_xform=tf((0.0, 0.0, -0.2))))
            angles.extend(stand)
        b.add_articulation(joints)
        lin = rng.normal(0.0, 0.3, 3) if falling else np.zeros(3)
        ang = rng.normal(0.0, 0.5, 3) if falling else np.zeros(3)
        _set_free_joint_state(b, joints[0], root, np.concatenate([lin, ang]))
        for j, a in zip(joints[1:], angles):
            qa = a + (rng.uniform(-0.4, 0.4) if falling else rng.uniform(-0.05, 0.05))
            b.joint_q[b.joint_q_start[j]] = qa
            b.dofs[b.joint_qd_start[j]].target_pos = a
            if falling:
                b.joint_qd[b.joint_qd_start[j]] = rng.normal(0.0, 0.5)
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.9, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


# arm: (parent offset along the previous link, joint axis, capsule half length)
_ARM = [((0.0, 0.0, 0.3), (0, 0, 1), 0.10), ((0.0, 0.0, 0.2), (0, 1, 0), 0.15), ((0.0, 0.0, 0.3), (0, 0, 1), 0.10),
        ((0.0, 0.0, 0.2), (0, 1, 0), 0.15), ((0.0, 0.0, 0.3), (0, 0, 1), 0.06), ((0.0, 0.0, 0.12), (0, 1, 0), 0.05),
        ((0.0, 0.0, 0.10), (0, 0, 1), 0.0)]
_ARM_POSE = (0.0, 1.1, 0.0, 1.3, 0.0, 0.75, 0.0)


def _arm_tip(q_arm) -> np.ndarray:
    """Palm frame of the arm at joint angles q_arm (base at the world origin)."""
    x = tf()
    for (off, axis, _), a in zip(_ARM, q_arm):
        x = tf_mul(tf_mul(x, tf(off)), tf((0, 0, 0), quat_from_axis_angle(axis, a)))
    return x


def build_arm_gripper_scene(num_worlds: int = 64, seed: int = 12, *, contact_gap: float = 0.005,
                            device: str = "cpu"):
    """A 7-DOF fixed-base arm (revolute z / y alternating, capsule links r 0.04) ending in a palm box
    with two prismatic fingers (9 DOF), reaching down to a free box (side U(0.05, 0.1) m) on the
    ground under the palm: 15 DOF, 10 bodies.

    The arm holds its pose with PD drives (ke 400, kd 20); the fingers start 1.5 cm wider than the
    box and are driven closed (ke 200, kd 5) so they squeeze it.  Arm links are in distinct
    collision groups (no self-contact); palm, fingers and box share group 1.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = num_worlds
    link_cfgs = [ShapeCfg(density=800.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap, collision_group=2 + i)
                 for i in range(7)]
    grip_cfg = ShapeCfg(density=800.0, ke=2.0e5, tau=0.005, mu=1.0, gap=contact_gap, collision_group=1)
    for w in range(num_worlds):
        b.current_world = w
        half = 0.5 * rng.uniform(0.05, 0.1)
        q_arm = np.array(_ARM_POSE) + rng.uniform(-0.05, 0.05, 7)
        palm = _arm_tip(q_arm)
        fingertip = tf_mul(palm, tf((0.0, 0.0, 0.11)))
        # lower the base so the fingertips sit at the box's mid-height
        base_z = half - fingertip[2]
        base = tf((0.0, 0.0, base_z))
        parent, x = -1, base
        joints = []
        for i, ((off, axis, hl), a) in enumerate(zip(_ARM, q_arm)):
            jx = tf_mul(x, tf(off)) if parent >= 0 else tf_mul(base, tf(off))
            body = b.add_body(tf_mul(jx, tf((0, 0, 0), quat_from_axis_angle(axis, a))))
            if hl > 0.0:
                b.add_shape(body, GEO_CAPSULE, tf((0.0, 0.0, hl)), (0.04, hl, 0.0), link_cfgs[i])
            else:
                b.add_shape(body, GEO_BOX, tf((0.0, 0.0, 0.02)), (0.03, 0.06, 0.02), grip_cfg)
            dof = DofCfg(axis=axis, target_ke=400.0, target_kd=20.0, actuator_mode=1, target_pos=float(a),
                         limit_lower=-3.0, limit_upper=3.0, armature=0.02, effort_limit=500.0)
            joints.append(b.add_joint(parent, body, SAP_JOINT_REVOLUTE, angular_axes=[dof],
                                      parent_xform=tf(off) if parent >= 0 else tf_mul(base, tf(off))))
            b.joint_q[b.joint_q_start[joints[-1]]] = float(a)
            parent, x = body, b.body_q[body]
        palm_body = parent
        open_w = half + 0.0075 + 0.012
        for side in (1.0, -1.0):
            fx = tf((0.0, side * open_w, 0.07))
            finger = b.add_body(tf_mul(b.body_q[palm_body], fx))
            b.add_shape(finger, GEO_BOX, tf((0.0, 0.0, 0.0)), (0.012, 0.012, 0.04), grip_cfg)
            dof = DofCfg(axis=(0.0, side, 0.0), target_ke=200.0, target_kd=5.0, actuator_mode=1,
                         target_pos=-0.03, limit_lower=-0.05, limit_upper=0.02, limit_ke=1.0e4, armature=0.01,
                         effort_limit=40.0)
            joints.append(b.add_joint(palm_body, finger, SAP_JOINT_PRISMATIC, linear_axes=[dof], parent_xform=fx))
        b.add_articulation(joints)
        tip = tf_mul(b.body_q[palm_body], tf((0.0, 0.0, 0.07)))
        box_pose = tf((tip[0], tip[1], half + 0.001), tip[3:])
        box = b.add_body(box_pose)
        b.add_shape(box, GEO_BOX, tf(), (half, half, half), ShapeCfg(density=500.0, ke=2.0e5, tau=0.005, mu=1.0,
                                                                    gap=contact_gap, collision_group=1))
        jb = b.add_free_joint(box)
        b.add_articulation([jb])
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_pile_scene(num_worlds: int = 64, seed: int = 13, *, bodies: int = 8, contact_gap: float = 0.01,
                     device: str = "cpu"):
    """Piles of free boxes and spheres (alternating by slot, sizes U(0.05, 0.2) m, density
    U(500, 2000)) dropped onto the ground from a loose lattice: 2 x 2 per layer, layers 0.3 m
    apart, for up to 8 bodies; 4 x 4 per layer 0.25 m apart for larger piles (16 bodies: one
    layer, 32: two); 6 DOF per body."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = num_worlds
    side, pitch = (2, 0.3) if bodies <= 8 else (4, 0.25)
    for w in range(num_worlds):
        b.current_world = w
        for k in range(bodies):
            i, j, l = k % side, (k // side) % side, k // (side * side)
            size = rng.uniform(0.05, 0.2)
            c = 0.5 * (side - 1)
            p = np.array([(i - c) * pitch, (j - c) * pitch, 0.12 + pitch * l]) + rng.uniform(-0.04, 0.04, 3)
            q = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, math.pi))
            pose = tf(p, q)
            body = b.add_body(pose)
            cfg = ShapeCfg(density=rng.uniform(500.0, 2000.0), ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap)
            if k % 2 == 0:
                b.add_shape(body, GEO_BOX, tf(), (0.5 * size, 0.5 * size * rng.uniform(0.6, 1.0), 0.5 * size), cfg)
            else:
                b.add_shape(body, GEO_SPHERE, tf(), (0.5 * size, 0.0, 0.0), cfg)
            j_ = b.add_free_joint(body)
            b.add_articulation([j_])
            _set_free_joint_state(b, j_, pose, np.concatenate([rng.normal(0.0, 0.2, 3), rng.normal(0.0, 1.0, 3)]))
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_bin_scene(bodies: int = 1024, seed: int = 17, *, columns: int = 6, contact_gap: float = 0.01,
                    device: str = "cpu"):
    """One world: ``bodies`` free boxes, spheres and capsules (by slot, sizes U(0.03, 0.15) m,
    density U(500, 2000)) stacked in columns x columns piles 0.16 m apart inside a bin (the ground
    plane and four static wall boxes, global shapes that do not collide with each other); each body
    rests 2 mm above the one below it, so the columns touch from the first step and collapse into
    a dense pile.  Boxes turn by a small yaw, capsules lie along x.  6 DOF per body."""
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = 1
    b.current_world = 0
    pitch = 0.16
    top = np.zeros(columns * columns)
    for k in range(bodies):
        col = k % (columns * columns)
        i, j = col % columns, col // columns
        size = rng.uniform(0.03, 0.15)
        kind = k % 3
        cfg = ShapeCfg(density=rng.uniform(500.0, 2000.0), ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap)
        if kind == 0:
            half = (0.5 * size, 0.5 * size * rng.uniform(0.6, 1.0), 0.5 * size)
            height = size
            q = quat_from_axis_angle(np.array([0.0, 0.0, 1.0]), rng.uniform(-0.3, 0.3))
        elif kind == 1:
            height = size
            q = np.array([0.0, 0.0, 0.0, 1.0])
        else:
            height = 0.5 * size
            q = quat_from_axis_angle(np.array([0.0, 1.0, 0.0]), 0.5 * math.pi)
        c = 0.5 * (columns - 1)
        z = top[col] + 0.002 + 0.5 * height
        top[col] = z + 0.5 * height
        p = np.array([(i - c) * pitch, (j - c) * pitch, z]) + np.concatenate([rng.uniform(-0.01, 0.01, 2), [0.0]])
        pose = tf(p, q)
        body = b.add_body(pose)
        if kind == 0:
            b.add_shape(body, GEO_BOX, tf(), half, cfg)
        elif kind == 1:
            b.add_shape(body, GEO_SPHERE, tf(), (0.5 * size, 0.0, 0.0), cfg)
        else:
            b.add_shape(body, GEO_CAPSULE, tf(), (0.25 * size, 0.25 * size, 0.0), cfg)
        j_ = b.add_free_joint(body)
        b.add_articulation([j_])
        _set_free_joint_state(b, j_, pose, np.zeros(6))
    b.current_world = -1
    wall = ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap, collision_group=-1)
    b.add_ground_plane(wall)
    half_h = 0.5 * (float(top.max()) + 0.1)
    e = 0.5 * columns * pitch + 0.02
    for x, y, hx, hy in ((e, 0.0, 0.01, e), (-e, 0.0, 0.01, e), (0.0, e, e, 0.01), (0.0, -e, e, 0.01)):
        b.add_shape(-1, GEO_BOX, tf((x, y, half_h)), (hx, hy, half_h), wall)
    return _with_consistent_body_q(b.finalize(device))


_BUILDERS = {"arm_gripper": build_arm_gripper_scene, "quadruped": build_quadruped_scene,
             "humanoid": build_humanoid_scene, "pile": build_pile_scene}


def draw_world_kinds(num_worlds: int, seed: int = 0) -> dict:
    """Seeded draw of each world's kind (uniform over the four kinds); returns {kind: count}."""
    rng = np.random.Generator(np.random.PCG64(seed))
    kinds = rng.integers(0, len(WORLD_KINDS), size=int(num_worlds))
    return {k: int(np.sum(kinds == i)) for i, k in enumerate(WORLD_KINDS) if np.any(kinds == i)}


def build_hetero_models(num_worlds: int = 8192, seed: int = 0, *, device: str = "cpu",
                        pile_sizes: tuple = (8,)) -> dict:
    """One model per world kind, sized by ``draw_world_kinds`` (each group seeded from ``seed``).

    ``pile_sizes``: body counts of the pile worlds; with several sizes each pile world draws its
    size (seeded), and every size is a group of its own (``pile`` for 8 bodies, ``pile<b>``
    otherwise: 16 bodies are 96 dofs, 32 bodies 192)."""
    counts = draw_world_kinds(num_worlds, seed)
    out = {}
    for i, (k, n) in enumerate(counts.items()):
        s = 1000 * seed + 17 * i + 1
        if k != "pile" or tuple(pile_sizes) == (8,):
            out[k] = _BUILDERS[k](n, seed=s, device=device)
            continue
        rng = np.random.Generator(np.random.PCG64(s))
        draw = rng.integers(0, len(pile_sizes), size=n)
        for j, b in enumerate(pile_sizes):
            m = int(np.sum(draw == j))
            if m:
                out["pile" if b == 8 else f"pile{b}"] = build_pile_scene(m, seed=s + 101 * j, bodies=int(b),
                                                                          device=device)
    return out


def contacts_for(name: str, default: int) -> int:
    """Contact capacity per world of a group: 8 per body for piles above 8 bodies."""
    if name.startswith("pile") and name != "pile":
        return 8 * int(name[4:])
    return default


class HeteroBatch:
    """Homogeneous groups stepped together (collide + SAP step per group).

    ``mode="batch"``: one step launch over all groups' worlds (graph-capturable after the first
    step).  ``mode="streams"``: one forked stream per group, joined back (eager only).  Each group
    keeps two states and alternates between them (``state(name)`` is the latest).
    """

    def __init__(self, models: dict, *, dt: float = 2.0e-3, max_contacts_per_world: int = 64,
                 contact_preset_variant: str = "approx32", line_search_variant: str = "exact_root",
                 mode: str = "batch", **solver_kwargs):
        import ctypes

        import torch

        from . import _native as nat
        from .collision import SapCollisionPipeline
        from .solver_sap import SolverSAP

        if mode not in ("batch", "streams"):
            raise ValueError(f"mode must be 'batch' or 'streams', got {mode!r}")
        self.mode = mode
        self.dt = float(dt)
        self.groups = {}
        for name, model in models.items():
            C = contacts_for(name, max_contacts_per_world)
            solver = SolverSAP(model, max_rigid_contact=C, contact_preset_variant=contact_preset_variant,
                               line_search_variant=line_search_variant, **solver_kwargs)
            pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
            self.groups[name] = dict(model=model, solver=solver, pipeline=pipe, contacts=pipe.contacts(),
                                     states=[model.state(), model.state()], control=model.control(),
                                     stream=torch.cuda.Stream())
        self._k = 0
        self._batch = None
        self._bound = [None, None]
        # groups the batch launch holds (one-wave kernel, at most 48 dofs and 64 contact slots);
        # the others (large piles: the 256-thread kernel) step with their own solver in the call
        self.batched = [n for n, g in self.groups.items()
                        if g["solver"].get_schedule()["kernel"] == "wave" and g["solver"].dof_per_env <= 48
                        and g["solver"].max_rigid_contact <= 64]
        self.own = [n for n in self.groups if n not in self.batched]
        if mode == "batch" and self.batched:
            self._lib = nat.load()
            gs = [self.groups[n] for n in self.batched]
            handles = (ctypes.c_void_p * len(gs))(*[g["solver"]._handle.value for g in gs])
            h = ctypes.c_void_p()
            nat.check(self._lib.sapx_batch_create(handles, len(gs), ctypes.byref(h)), "sapx_batch_create")
            self._batch = h

    def close(self) -> None:
        if getattr(self, "_batch", None):
            self._lib.sapx_batch_destroy(self._batch)
            self._batch = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def set_schedule(self, small_capacity: int = 16, fork: bool = False) -> None:
        """Contact variants of the batch launch (``sapx_batch_set_schedule``)."""
        from . import _native as nat

        if self.mode != "batch":
            raise ValueError("set_schedule applies to mode='batch'")
        if self._batch is None:
            return  # no group in the batch launch
        nat.check(self._lib.sapx_batch_set_schedule(self._batch, int(small_capacity), int(bool(fork))),
                  "sapx_batch_set_schedule")

    def _bind(self, slot: int) -> None:
        import ctypes

        from . import _native as nat

        gs = [self.groups[n] for n in self.batched]
        G = len(gs)

        def sview(st):
            return nat.SapxStateView(nat._ptr(st.joint_q), nat._ptr(st.joint_qd), nat._ptr(st.body_q),
                                     nat._ptr(st.body_qd), nat._ptr(st.body_f))

        sin = (nat.SapxStateView * G)(*[sview(g["states"][slot]) for g in gs])
        sout = (nat.SapxStateView * G)(*[sview(g["states"][1 - slot]) for g in gs])
        ctl = (nat.SapxControlView * G)(*[nat.SapxControlView(
            nat._ptr(g["control"].joint_f), nat._ptr(g["control"].joint_target_pos),
            nat._ptr(g["control"].joint_target_vel), nat._ptr(g["control"].joint_act)) for g in gs])
        cv = (nat.SapxContactsView * G)(*[nat.contacts_view(g["contacts"]) for g in gs])
        nat.check(self._lib.sapx_batch_bind(self._batch, slot, sin, sout, ctl, cv, self.dt), "sapx_batch_bind")
        self._bound[slot] = (sin, sout, ctl, cv)

    @property
    def world_count(self) -> int:
        return sum(g["model"].world_count for g in self.groups.values())

    def state(self, name):
        g = self.groups[name]
        return g["states"][self._k % 2] if self.mode == "batch" else g["states"][0]

    def step(self) -> None:
        import torch

        from . import _native as nat

        capturing = torch.cuda.is_current_stream_capturing()
        if self.mode == "batch":
            slot = self._k % 2
            if capturing and any(g["solver"].collect_iteration_stats for g in self.groups.values()):
                raise RuntimeError("construct the batch with collect_iteration_stats=False to capture it")
            if self._batch is not None and self._bound[slot] is None:
                if capturing:
                    raise RuntimeError("step the batch eagerly (three steps) before capturing it: binds both slots")
                self._bind(slot)
            for g in self.groups.values():
                g["pipeline"].collide(g["states"][slot], g["contacts"])
            if self._batch is not None:
                with nat.on_device(self.groups[self.batched[0]]["solver"]._device):
                    nat.check(self._lib.sapx_batch_step(self._batch, slot, nat.stream_ptr()), "sapx_batch_step")
            for n in self.own:
                g = self.groups[n]
                g["solver"].step(g["states"][slot], g["states"][1 - slot], g["control"], g["contacts"], self.dt)
            self._k += 1
            return
        if capturing:
            # A capture of this multi-stream step (capture stream -> per-kind stream -> each solver's
            # second stream) once ended in a crash inside the HIP runtime's end of capture, with no
            # cause shown; it stays switched off.
            raise RuntimeError("HeteroBatch.step in mode='streams' cannot be captured in a HIP graph; use "
                               "mode='batch' or step it eagerly")
        main = torch.cuda.current_stream()
        for g in self.groups.values():
            g["stream"].wait_stream(main)
            with torch.cuda.stream(g["stream"]):
                s0, s1 = g["states"]
                g["pipeline"].collide(s0, g["contacts"])
                g["solver"].step(s0, s1, g["control"], g["contacts"], self.dt)
            g["states"] = [s1, s0]
        for g in self.groups.values():
            main.wait_stream(g["stream"])


__all__ = ["HeteroBatch", "WORLD_KINDS", "build_arm_gripper_scene", "build_hetero_models", "build_pile_scene",
           "build_quadruped_scene", "contacts_for", "draw_world_kinds"]

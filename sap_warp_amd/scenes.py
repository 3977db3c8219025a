"""Synthetic scene builders producing ``SapModel`` instances.

The builder reproduces the model-array conventions of the reference scene
loader (``sim/loader/scene.py``): body/shape/joint bookkeeping
(``_SapSceneBuilder.add_body/add_shape/add_joint``, ``scene.py:282-495``),
density-based inertia accumulation with the parallel-axis theorem
(``_update_body_mass`` ``scene.py:389-411``, inertia formulas
``scene.py:5847-5918``), same-body / parent-child collision filters
(``scene.py:319-373, 477-484``), collision-group pair tests
(``scene.py:5788-5801``) and the explicit candidate pair list
(``_find_shape_contact_pairs`` ``scene.py:1188-1212``).

Scenes:

* :func:`build_ant_scene` — the 14-DOF ant batch (free torso sphere, four
  two-hinge capsule legs) used as the CPU-runnable case.
* :func:`build_humanoid_scene` — the 35-DOF humanoid batch (free pelvis + 29
  revolute joints, 30 bodies, capsule/sphere/box links) used as the GPU
  workload.
* :func:`build_box_pile_scene`, :func:`build_round_shapes_scene`,
  :func:`build_convex_shapes_scene`, :func:`build_ball_chain_scene`,
  :func:`build_d6_ant_scene` — parity scenes for the box pairs, the round-shape
  primitives, the MPR / GJK pairs (cones, cylinders, ellipsoids), ball-joint
  chains and D6 joints.

Worlds are independent (no cross-world pairs), so by default every world sits
at the origin (``spacing=0``): positions stay O(1 m) and float32 geometry keeps
full precision.  A non-zero spacing reproduces the reference's grid layout
(``DEFAULT_REPLICATE_SPACING``, ``scene.py:42``).
"""
from __future__ import annotations

import math
from dataclasses import dataclass, replace

import numpy as np

from .runtime import (
    GEO_BOX,
    GEO_CAPSULE,
    GEO_CONE,
    GEO_CYLINDER,
    GEO_ELLIPSOID,
    GEO_PLANE,
    GEO_SPHERE,
    SAP_BODY_FLAG_DYNAMIC,
    SAP_JOINT_BALL,
    SAP_JOINT_D6,
    SAP_JOINT_DISTANCE,
    SAP_JOINT_FIXED,
    SAP_JOINT_FREE,
    SAP_JOINT_PRISMATIC,
    SAP_JOINT_REVOLUTE,
    SHAPE_FLAG_COLLIDE_PARTICLES,
    SHAPE_FLAG_COLLIDE_SHAPES,
    SHAPE_FLAG_VISIBLE,
    SapModel,
)

IDENTITY_TF = np.array([0, 0, 0, 0, 0, 0, 1], dtype=np.float64)


# ----------------------------------------------------------------------------
# small math helpers (float64, quaternions xyzw)
# ----------------------------------------------------------------------------
def quat_from_axis_angle(axis, angle) -> np.ndarray:
    axis = np.asarray(axis, dtype=np.float64)
    axis = axis / np.linalg.norm(axis)
    s = math.sin(0.5 * angle)
    return np.array([axis[0] * s, axis[1] * s, axis[2] * s, math.cos(0.5 * angle)])


def quat_mul(a, b) -> np.ndarray:
    ax, ay, az, aw = a
    bx, by, bz, bw = b
    return np.array(
        [
            aw * bx + ax * bw + ay * bz - az * by,
            aw * by - ax * bz + ay * bw + az * bx,
            aw * bz + ax * by - ay * bx + az * bw,
            aw * bw - ax * bx - ay * by - az * bz,
        ]
    )


def quat_to_matrix(q) -> np.ndarray:
    x, y, z, w = q
    return np.array(
        [
            [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
            [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
            [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
        ]
    )


def quat_rotate(q, v) -> np.ndarray:
    return quat_to_matrix(q) @ np.asarray(v, dtype=np.float64)


def tf(p=(0.0, 0.0, 0.0), q=(0.0, 0.0, 0.0, 1.0)) -> np.ndarray:
    return np.array([p[0], p[1], p[2], q[0], q[1], q[2], q[3]], dtype=np.float64)


def tf_mul(a, b) -> np.ndarray:
    p = a[:3] + quat_rotate(a[3:], b[:3])
    return np.concatenate([p, quat_mul(a[3:], b[3:])])


def tf_point(a, p) -> np.ndarray:
    return a[:3] + quat_rotate(a[3:], p)


def quat_z_to(direction) -> np.ndarray:
    """Quaternion rotating local +Z onto ``direction`` (capsule axis convention)."""
    d = np.asarray(direction, dtype=np.float64)
    d = d / np.linalg.norm(d)
    z = np.array([0.0, 0.0, 1.0])
    c = float(np.dot(z, d))
    if c > 1.0 - 1e-12:
        return np.array([0.0, 0.0, 0.0, 1.0])
    if c < -1.0 + 1e-12:
        return np.array([1.0, 0.0, 0.0, 0.0])
    axis = np.cross(z, d)
    return quat_from_axis_angle(axis, math.acos(max(-1.0, min(1.0, c))))


# ----------------------------------------------------------------------------
# inertia formulas (scene.py:5847-5918)
# ----------------------------------------------------------------------------
def inertia_sphere(density, r):
    mass = density * 4.0 / 3.0 * math.pi * r ** 3
    i = 2.0 / 5.0 * mass * r * r
    return mass, np.diag([i, i, i])


def inertia_capsule(density, r, half_height):
    h = 2.0 * half_height
    ms = density * (4.0 / 3.0) * math.pi * r ** 3
    mc = density * math.pi * r * r * h
    ixy = mc * (0.25 * r * r + h * h / 12.0) + ms * (0.4 * r * r + 0.375 * r * h + 0.25 * h * h)
    iz = (mc * 0.5 + ms * 0.4) * r * r
    return ms + mc, np.diag([ixy, ixy, iz])


def inertia_box(density, hx, hy, hz):
    mass = density * 8.0 * hx * hy * hz
    return mass, np.diag(
        [mass / 3.0 * (hy * hy + hz * hz), mass / 3.0 * (hx * hx + hz * hz), mass / 3.0 * (hx * hx + hy * hy)]
    )


def inertia_cylinder(density, r, half_height):
    """Solid cylinder about its centre, axis +Z (``_sap_compute_inertia_cylinder`` scene.py:5866-5875)."""
    h = 2.0 * half_height
    mass = density * math.pi * r * r * h
    ixy = mass * (3.0 * r * r + h * h) / 12.0
    return mass, np.diag([ixy, ixy, 0.5 * mass * r * r])


def inertia_ellipsoid(density, rx, ry, rz):
    """Solid ellipsoid about its centre (``_sap_compute_inertia_ellipsoid`` scene.py:5887-5898)."""
    mass = density * 4.0 / 3.0 * math.pi * rx * ry * rz
    return mass, np.diag([0.2 * mass * (ry * ry + rz * rz), 0.2 * mass * (rx * rx + rz * rz),
                          0.2 * mass * (rx * rx + ry * ry)])


def inertia_cone(density, r, half_height):
    """Solid cone, apex at +half_height: mass, centre of mass (0, 0, -h/4) for the height h, and
    the inertia about that centre (``_sap_compute_inertia_cone`` scene.py:5878-5884)."""
    h = 2.0 * half_height
    mass = density * math.pi * r * r * h / 3.0
    ixy = 3.0 / 20.0 * mass * r * r + 3.0 / 80.0 * mass * h * h
    return mass, np.array([0.0, 0.0, -h / 4.0]), np.diag([ixy, ixy, 3.0 / 10.0 * mass * r * r])


def shape_volume(shape_type, scale) -> float:
    if shape_type == GEO_SPHERE:
        return 4.0 / 3.0 * math.pi * scale[0] ** 3
    if shape_type == GEO_CAPSULE:
        r, hh = scale[0], scale[1]
        return 4.0 / 3.0 * math.pi * r ** 3 + math.pi * r * r * 2.0 * hh
    if shape_type == GEO_BOX:
        return 8.0 * scale[0] * scale[1] * scale[2]
    if shape_type == GEO_CYLINDER:
        return math.pi * scale[0] ** 2 * 2.0 * scale[1]
    if shape_type == GEO_ELLIPSOID:
        return 4.0 / 3.0 * math.pi * scale[0] * scale[1] * scale[2]
    if shape_type == GEO_CONE:
        return math.pi * scale[0] ** 2 * 2.0 * scale[1] / 3.0
    return 0.0


def transform_inertia(mass, inertia, offset, q) -> np.ndarray:
    """Parallel-axis + rotation (``_sap_transform_inertia`` scene.py:5913-5918)."""
    rot = quat_to_matrix(q)
    o = np.asarray(offset, dtype=np.float64)
    return rot @ inertia @ rot.T + mass * (np.dot(o, o) * np.eye(3) - np.outer(o, o))


def shape_radius(shape_type, scale) -> float:
    """``_compute_shape_radius`` (scene.py:5804-5820)."""
    if shape_type == GEO_SPHERE:
        return float(scale[0])
    if shape_type == GEO_BOX:
        return float(np.linalg.norm(scale))
    if shape_type in (GEO_CAPSULE, GEO_CYLINDER, GEO_CONE):
        return float(scale[0] + scale[1])
    if shape_type == GEO_ELLIPSOID:
        return float(max(scale[0], scale[1], scale[2]))
    if shape_type == GEO_PLANE:
        if scale[0] > 0.0 and scale[1] > 0.0:
            return float(np.linalg.norm(scale))
        return 1.0e6
    return 10.0


# ----------------------------------------------------------------------------
# builder
# ----------------------------------------------------------------------------
@dataclass
class ShapeCfg:
    """Per-shape material / collision config (defaults: ``_SapShapeConfig`` scene.py:86-110)."""

    density: float = 1000.0
    ke: float = 2.5e3
    tau: float = 0.01
    mu: float = 1.0
    margin: float = 0.0
    gap: float | None = None
    collision_group: int = 1
    collision_filter_parent: bool = True
    has_shape_collision: bool = True
    has_particle_collision: bool = True
    is_visible: bool = True

    def flags(self) -> int:
        f = SHAPE_FLAG_VISIBLE if self.is_visible else 0
        if self.has_particle_collision:
            f |= SHAPE_FLAG_COLLIDE_PARTICLES
        if self.has_shape_collision:
            f |= SHAPE_FLAG_COLLIDE_SHAPES
        return f


@dataclass
class DofCfg:
    """Per-axis joint config (defaults: ``_SapJointDofConfig`` scene.py:169-183)."""

    axis: tuple = (1.0, 0.0, 0.0)
    target_pos: float = 0.0
    target_vel: float = 0.0
    target_ke: float = 0.0
    target_kd: float = 0.0
    limit_lower: float = -1.0e10
    limit_upper: float = 1.0e10
    limit_ke: float = 1.0e4
    limit_kd: float = 10.0
    armature: float = 0.0
    effort_limit: float = 1.0e6
    actuator_mode: int = 0


def _free_axes():
    return [DofCfg(axis=a, limit_ke=0.0, limit_kd=0.0) for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]


def _joint_dof_count(joint_type: int, num_axes: int) -> tuple[int, int]:
    """(dof_count, coord_count) per joint type (scene.py:1215-1222)."""
    if joint_type == SAP_JOINT_BALL:
        return 3, 4
    if joint_type in (SAP_JOINT_FREE, SAP_JOINT_DISTANCE):
        return 6, 7
    if joint_type == SAP_JOINT_FIXED:
        return 0, 0
    return num_axes, num_axes


def _test_group_pair(a: int, b: int) -> bool:
    """scene.py:5788-5795."""
    if a == 0 or b == 0:
        return False
    if a > 0:
        return a == b or b < 0
    if a < 0:
        return a != b
    return False


class SceneBuilder:
    """Host-side model builder (``_SapSceneBuilder`` semantics)."""

    def __init__(self, *, rigid_gap: float = 0.1, gravity: float = -9.81):
        self.rigid_gap = float(rigid_gap)
        self.gravity = float(gravity)
        self.default_body_armature = 0.0
        self.current_world = -1
        self.world_count = 1
        # bodies
        self.body_q: list = []
        self.body_mass: list = []
        self.body_inertia: list = []
        self.body_com: list = []
        self.body_world: list = []
        self.body_lock_inertia: list = []
        self.body_shapes: dict = {-1: []}
        # joints
        self.joint_type: list = []
        self.joint_parent: list = []
        self.joint_child: list = []
        self.joint_parents: dict = {}
        self.joint_children: dict = {}
        self.joint_X_p: list = []
        self.joint_X_c: list = []
        self.joint_q: list = []
        self.joint_qd: list = []
        self.joint_q_start: list = []
        self.joint_qd_start: list = []
        self.joint_dof_dim: list = []
        self.joint_articulation: list = []
        self.dofs: list = []
        self.articulation_start: list = []
        # shapes
        self.shape_transform: list = []
        self.shape_body: list = []
        self.shape_type: list = []
        self.shape_scale: list = []
        self.shape_cfg: list = []
        self.shape_world: list = []
        self.filter_pairs: set = set()

    # ------------------------------------------------------------------ bodies
    def add_body(self, xform, *, mass=0.0, com=None, inertia=None, lock_inertia=False, armature=None) -> int:
        bid = len(self.body_q)
        self.body_q.append(np.asarray(xform, dtype=np.float64).copy())
        self.body_mass.append(float(mass))
        self.body_com.append(np.zeros(3) if com is None else np.asarray(com, dtype=np.float64))
        body_inertia = np.zeros((3, 3)) if inertia is None else np.asarray(inertia, dtype=np.float64).reshape(3, 3)
        arm = self.default_body_armature if armature is None else float(armature)
        if arm != 0.0:  # scene.py:301-303
            body_inertia = body_inertia + np.eye(3) * arm
        self.body_inertia.append(body_inertia)
        self.body_world.append(self.current_world)
        self.body_lock_inertia.append(bool(lock_inertia))
        self.body_shapes[bid] = []
        return bid

    def _update_body_mass(self, body, mass, inertia, p, q):
        """scene.py:389-411."""
        if body == -1:
            return
        new_mass = self.body_mass[body] + mass
        if new_mass == 0.0:
            return
        old_mass = self.body_mass[body]
        new_com = (self.body_com[body] * old_mass + p * mass) / new_mass
        com_offset = new_com - self.body_com[body]
        shape_offset = new_com - p
        new_inertia = transform_inertia(old_mass, self.body_inertia[body], com_offset, (0, 0, 0, 1)) + transform_inertia(
            mass, inertia, shape_offset, q
        )
        self.body_mass[body] = new_mass
        self.body_inertia[body] = new_inertia
        self.body_com[body] = new_com

    # ------------------------------------------------------------------ shapes
    def add_shape(self, body: int, shape_type: int, xform, scale, cfg: ShapeCfg) -> int:
        sid = len(self.shape_type)
        self.body_shapes.setdefault(body, [])
        if cfg.has_shape_collision:
            for other in self.body_shapes[body]:
                self._filter(other, sid)
        self.shape_body.append(body)
        self.body_shapes[body].append(sid)
        self.shape_transform.append(np.asarray(xform, dtype=np.float64).copy())
        self.shape_type.append(int(shape_type))
        self.shape_scale.append(tuple(float(s) for s in scale))
        self.shape_cfg.append(replace(cfg))
        self.shape_world.append(self.current_world)
        if cfg.has_shape_collision and cfg.collision_filter_parent and body > -1:
            for parent_body in self.joint_parents.get(body, []):
                if parent_body > -1:
                    for ps in self.body_shapes[parent_body]:
                        self._filter(ps, sid)
            for child_body in self.joint_children.get(body, []):
                for cs in self.body_shapes[child_body]:
                    self._filter(sid, cs)
        if cfg.density > 0.0 and body >= 0 and not self.body_lock_inertia[body]:
            if shape_type == GEO_SPHERE:
                m, inertia = inertia_sphere(cfg.density, scale[0])
            elif shape_type == GEO_CAPSULE:
                m, inertia = inertia_capsule(cfg.density, scale[0], scale[1])
            elif shape_type == GEO_BOX:
                m, inertia = inertia_box(cfg.density, *scale)
            elif shape_type == GEO_CYLINDER:
                m, inertia = inertia_cylinder(cfg.density, scale[0], scale[1])
            elif shape_type == GEO_ELLIPSOID:
                m, inertia = inertia_ellipsoid(cfg.density, *scale)
            elif shape_type == GEO_CONE:
                m, com, inertia = inertia_cone(cfg.density, scale[0], scale[1])
            else:
                m, inertia = 0.0, np.zeros((3, 3))
            if m > 0.0:
                x = self.shape_transform[sid]
                p = x[:3].copy()
                if shape_type == GEO_CONE:  # the shape's centre of mass in body coordinates
                    p = p + quat_rotate(x[3:], com)
                self._update_body_mass(body, m, inertia, p, x[3:])
        return sid

    def add_ground_plane(self, cfg: ShapeCfg, height: float = 0.0) -> int:
        world = self.current_world
        self.current_world = -1
        sid = self.add_shape(-1, GEO_PLANE, tf((0.0, 0.0, height)), (0.0, 0.0, 0.0), replace(cfg, density=0.0))
        self.current_world = world
        return sid

    def _filter(self, a: int, b: int):
        self.filter_pairs.add((min(a, b), max(a, b)))

    # ------------------------------------------------------------------ joints
    def add_joint(
        self,
        parent: int,
        child: int,
        joint_type: int,
        *,
        linear_axes=(),
        angular_axes=(),
        parent_xform=None,
        child_xform=None,
        collision_filter_parent=True,
    ) -> int:
        jid = len(self.joint_type)
        lin = [a if isinstance(a, DofCfg) else DofCfg(axis=tuple(a)) for a in linear_axes]
        ang = [a if isinstance(a, DofCfg) else DofCfg(axis=tuple(a)) for a in angular_axes]
        self.joint_type.append(int(joint_type))
        self.joint_parent.append(int(parent))
        self.joint_child.append(int(child))
        self.joint_X_p.append(IDENTITY_TF.copy() if parent_xform is None else np.asarray(parent_xform, np.float64))
        self.joint_X_c.append(IDENTITY_TF.copy() if child_xform is None else np.asarray(child_xform, np.float64))
        self.joint_q_start.append(len(self.joint_q))
        self.joint_qd_start.append(len(self.joint_qd))
        self.joint_dof_dim.append((len(lin), len(ang)))
        self.joint_articulation.append(-1)
        self.dofs.extend(lin + ang)
        dof_count, coord_count = _joint_dof_count(joint_type, len(lin) + len(ang))
        self.joint_q.extend([0.0] * coord_count)
        self.joint_qd.extend([0.0] * dof_count)
        if joint_type in (SAP_JOINT_FREE, SAP_JOINT_DISTANCE, SAP_JOINT_BALL) and coord_count > 0:
            self.joint_q[-1] = 1.0
        if joint_type == SAP_JOINT_FREE:
            s = self.joint_q_start[jid]
            self.joint_q[s : s + 7] = list(self.body_q[child])
        self.joint_parents.setdefault(child, []).append(parent)
        ch = self.joint_children.setdefault(parent, [])
        if child not in ch:
            ch.append(child)
        if collision_filter_parent and parent > -1:
            for cs in self.body_shapes.get(child, []):
                if not self.shape_cfg[cs].has_shape_collision:
                    continue
                for ps in self.body_shapes.get(parent, []):
                    if not self.shape_cfg[ps].has_shape_collision:
                        continue
                    self._filter(ps, cs)
        return jid

    def add_free_joint(self, child: int, **kw) -> int:
        return self.add_joint(-1, child, SAP_JOINT_FREE, linear_axes=_free_axes(), angular_axes=_free_axes(), **kw)

    def add_articulation(self, joints) -> None:
        joints = [int(j) for j in joints]
        art = len(self.articulation_start)
        self.articulation_start.append(joints[0])
        for j in joints:
            self.joint_articulation[j] = art

    # -------------------------------------------------------------- replicate
    def replicate(self, world_count: int, *, spacing=(2.0, 2.0, 0.0)) -> None:
        """Copy the built (single) world ``world_count`` times (``_SapSceneBuilder.replicate``,
        scene.py:513-760).

        World w's bodies, free-joint root coordinates, root non-free joint anchors and world-attached
        shapes move by the w-th grid offset (``_compute_sap_world_offsets``, scene.py:5650-5685: a
        grid over the non-zero spacing axes, centred except along z).  Everything already in the
        builder becomes world 0's template; global shapes (the ground) are added afterwards.
        """
        W = max(int(world_count), 1)
        nb, nj, ns = len(self.body_q), len(self.joint_type), len(self.shape_type)
        nq, nd, na = len(self.joint_q), len(self.joint_qd), len(self.articulation_start)
        base = {k: list(getattr(self, k)) for k in (
            "body_q", "body_mass", "body_inertia", "body_com", "body_lock_inertia", "joint_type", "joint_parent",
            "joint_child", "joint_X_p", "joint_X_c", "joint_q", "joint_qd", "joint_q_start", "joint_qd_start",
            "joint_dof_dim", "joint_articulation", "dofs", "articulation_start", "shape_transform", "shape_body",
            "shape_type", "shape_scale", "shape_cfg")}
        base_filters = sorted(self.filter_pairs)
        offsets = world_offsets(W, spacing)
        for k in base:
            setattr(self, k, [])
        self.body_world, self.shape_world = [], []
        self.body_shapes = {-1: []}
        self.joint_parents, self.joint_children = {}, {}
        self.filter_pairs = set()
        for w in range(W):
            off = offsets[w]
            bo, jo, so, qo, do, ao = w * nb, w * nj, w * ns, w * nq, w * nd, w * na
            for x in base["body_q"]:
                y = np.asarray(x, np.float64).copy()
                y[:3] += off
                self.body_q.append(y)
            self.body_mass.extend(base["body_mass"])
            self.body_inertia.extend([np.array(i) for i in base["body_inertia"]])
            self.body_com.extend([np.array(c) for c in base["body_com"]])
            self.body_lock_inertia.extend(base["body_lock_inertia"])
            self.body_world.extend([w] * nb)
            for b in range(nb):
                self.body_shapes[bo + b] = []
            self.joint_type.extend(base["joint_type"])
            self.joint_parent.extend(p + bo if p >= 0 else -1 for p in base["joint_parent"])
            self.joint_child.extend(c + bo for c in base["joint_child"])
            self.joint_X_c.extend([np.array(x) for x in base["joint_X_c"]])
            self.joint_q_start.extend(q + qo for q in base["joint_q_start"])
            self.joint_qd_start.extend(d + do for d in base["joint_qd_start"])
            self.joint_dof_dim.extend(base["joint_dof_dim"])
            self.joint_articulation.extend(a + ao if a >= 0 else -1 for a in base["joint_articulation"])
            self.dofs.extend(replace(d) for d in base["dofs"])
            self.articulation_start.extend(a + jo for a in base["articulation_start"])
            q = list(base["joint_q"])
            for j, jt in enumerate(base["joint_type"]):
                root = base["joint_parent"][j] == -1
                xp = np.asarray(base["joint_X_p"][j], np.float64).copy()
                if root and jt in (SAP_JOINT_FREE, SAP_JOINT_DISTANCE):
                    qs = base["joint_q_start"][j]
                    for k in range(3):
                        q[qs + k] += float(off[k])
                elif root:
                    xp[:3] += off
                self.joint_X_p.append(xp)
            self.joint_q.extend(q)
            self.joint_qd.extend(base["joint_qd"])
            for si in range(ns):
                b = base["shape_body"][si]
                rb = b + bo if b >= 0 else -1
                x = np.asarray(base["shape_transform"][si], np.float64).copy()
                if rb < 0:
                    x[:3] += off
                self.shape_transform.append(x)
                self.shape_body.append(rb)
                self.shape_type.append(base["shape_type"][si])
                self.shape_scale.append(base["shape_scale"][si])
                self.shape_cfg.append(replace(base["shape_cfg"][si]))
                self.shape_world.append(w)
                self.body_shapes.setdefault(rb, []).append(so + si)
            for a, b in base_filters:
                self._filter(a + so, b + so)
            for j in range(nj):
                p, c = self.joint_parent[jo + j], self.joint_child[jo + j]
                self.joint_parents.setdefault(c, []).append(p)
                ch = self.joint_children.setdefault(p, [])
                if c not in ch:
                    ch.append(c)
        self.world_count = W
        self.current_world = -1

    # --------------------------------------------------------------- finalize
    def shape_pairs(self) -> np.ndarray:
        """Explicit candidate pairs (``_find_shape_contact_pairs`` scene.py:1188-1212)."""
        colliding = [i for i, c in enumerate(self.shape_cfg) if c.has_shape_collision]
        order = sorted(colliding, key=lambda i: self.shape_world[i])
        pairs = []
        for i1, s1 in enumerate(order):
            w1 = self.shape_world[s1]
            g1 = self.shape_cfg[s1].collision_group
            for s2 in order[i1 + 1 :]:
                w2 = self.shape_world[s2]
                g2 = self.shape_cfg[s2].collision_group
                if w1 != -1 and w2 != -1 and w1 != w2:
                    break
                if w1 != -1 and w2 != -1 and w1 != w2:
                    continue
                if not _test_group_pair(g1, g2):
                    continue
                a, b = (s2, s1) if s1 > s2 else (s1, s2)
                if (a, b) not in self.filter_pairs:
                    pairs.append((a, b))
        return np.asarray(pairs, dtype=np.int32).reshape(-1, 2)

    def finalize(self, device: str = "cpu") -> SapModel:
        f32 = np.float32
        i32 = np.int32
        joint_count = len(self.joint_type)
        articulation_start = list(self.articulation_start) + [joint_count]
        child_to_joint = {c: j for j, c in enumerate(self.joint_child)}
        joint_ancestor = [child_to_joint.get(p, -1) for p in self.joint_parent]
        world_count = max(self.world_count, 1)
        gravity = np.tile(np.array([0.0, 0.0, self.gravity], dtype=f32), (world_count, 1))

        def dof_field(name):
            return np.asarray([getattr(d, name) for d in self.dofs], dtype=f32)

        shape_gap = [c.gap if c.gap is not None else self.rigid_gap for c in self.shape_cfg]
        # body_qd public convention [v_com, w]; zero at construction
        body_q = np.asarray(self.body_q, dtype=f32).reshape(-1, 7)
        model = SapModel(
            device=device,
            joint_count=joint_count,
            joint_dof_count=len(self.joint_qd),
            joint_coord_count=len(self.joint_q),
            body_count=len(self.body_q),
            articulation_count=len(self.articulation_start),
            world_count=world_count,
            shape_count=len(self.shape_type),
            joint_type=np.asarray(self.joint_type, dtype=i32),
            joint_articulation=np.asarray(self.joint_articulation, dtype=i32),
            joint_parent=np.asarray(self.joint_parent, dtype=i32),
            joint_child=np.asarray(self.joint_child, dtype=i32),
            joint_q_start=np.asarray(self.joint_q_start + [len(self.joint_q)], dtype=i32),
            joint_qd_start=np.asarray(self.joint_qd_start + [len(self.joint_qd)], dtype=i32),
            joint_dof_dim=np.asarray(self.joint_dof_dim, dtype=i32).reshape(-1, 2),
            joint_axis=np.asarray([d.axis for d in self.dofs], dtype=f32).reshape(-1, 3),
            joint_X_p=np.asarray(self.joint_X_p, dtype=f32).reshape(-1, 7),
            joint_X_c=np.asarray(self.joint_X_c, dtype=f32).reshape(-1, 7),
            joint_ancestor=np.asarray(joint_ancestor, dtype=i32),
            articulation_start=np.asarray(articulation_start, dtype=i32),
            joint_q=np.asarray(self.joint_q, dtype=f32),
            joint_qd=np.asarray(self.joint_qd, dtype=f32),
            body_q=body_q,
            body_qd=np.zeros((len(self.body_q), 6), dtype=f32),
            body_mass=np.asarray(self.body_mass, dtype=f32),
            body_inertia=np.asarray(self.body_inertia, dtype=f32).reshape(-1, 3, 3),
            body_com=np.asarray(self.body_com, dtype=f32).reshape(-1, 3),
            body_flags=np.full(len(self.body_q), SAP_BODY_FLAG_DYNAMIC, dtype=i32),
            body_world=np.asarray(self.body_world, dtype=i32),
            gravity=gravity,
            joint_armature=dof_field("armature"),
            shape_body=np.asarray(self.shape_body, dtype=i32),
            shape_material_mu=np.asarray([c.mu for c in self.shape_cfg], dtype=f32),
            shape_material_ke=np.asarray([c.ke for c in self.shape_cfg], dtype=f32),
            shape_material_tau=np.asarray([c.tau for c in self.shape_cfg], dtype=f32),
            joint_target_mode=np.asarray([d.actuator_mode for d in self.dofs], dtype=i32),
            joint_target_ke=dof_field("target_ke"),
            joint_target_kd=dof_field("target_kd"),
            joint_target_pos=dof_field("target_pos"),
            joint_target_vel=dof_field("target_vel"),
            joint_f=np.zeros(len(self.dofs), dtype=f32),
            joint_act=np.zeros(len(self.dofs), dtype=f32),
            joint_effort_limit=dof_field("effort_limit"),
            joint_limit_lower=dof_field("limit_lower"),
            joint_limit_upper=dof_field("limit_upper"),
            joint_limit_ke=dof_field("limit_ke"),
            joint_limit_kd=dof_field("limit_kd"),
            shape_transform=np.asarray(self.shape_transform, dtype=f32).reshape(-1, 7),
            shape_type=np.asarray(self.shape_type, dtype=i32),
            shape_scale=np.asarray(self.shape_scale, dtype=f32).reshape(-1, 3),
            shape_flags=np.asarray([c.flags() for c in self.shape_cfg], dtype=i32),
            shape_margin=np.asarray([c.margin for c in self.shape_cfg], dtype=f32),
            shape_gap=np.asarray(shape_gap, dtype=f32),
            shape_collision_radius=np.asarray(
                [shape_radius(t, s) for t, s in zip(self.shape_type, self.shape_scale)], dtype=f32
            ),
            shape_collision_group=np.asarray([c.collision_group for c in self.shape_cfg], dtype=i32),
            shape_world=np.asarray(self.shape_world, dtype=i32),
            shape_pairs=self.shape_pairs(),
            shape_collision_filter_pairs=np.asarray(sorted(self.filter_pairs), dtype=np.int32).reshape(-1, 2),
        )
        return model


# ----------------------------------------------------------------------------
# scenes
# ----------------------------------------------------------------------------
def world_offsets(world_count: int, spacing, up_axis: int = 2) -> np.ndarray:
    """Per-world origin offsets of a replicated scene (``_compute_sap_world_offsets``,
    scene.py:5650-5685): worlds fill a grid over the axes with non-zero spacing (a line, a square
    or a cube of side ceil(W^(1/dims))), then the grid is centred on the axes other than up."""
    sp = np.asarray(spacing, dtype=np.float32)
    W = int(world_count)
    out = np.zeros((max(W, 0), 3), dtype=np.float32)
    if W <= 0:
        return out
    nz = np.nonzero(sp)[0]
    dims = int(nz.shape[0])
    if dims:
        side = int(np.ceil(W ** (1.0 / dims)))
        for i in range(W):
            if dims == 1:
                out[i] = i * sp
            elif dims == 2:
                out[i, nz[0]] = (i % side) * sp[nz[0]]
                out[i, nz[1]] = (i // side) * sp[nz[1]]
            else:
                out[i, 0] = (i % side) * sp[0]
                out[i, 1] = ((i // side) % side) * sp[1]
                out[i, 2] = (i // (side * side)) * sp[2]
    lo = out.min(axis=0)
    corr = lo + (out.max(axis=0) - lo) / 2.0
    corr[int(up_axis)] = 0.0
    return out - corr


def _grid_offset(world: int, world_count: int, spacing: float) -> np.ndarray:
    if spacing == 0.0:
        return np.zeros(3)
    side = int(math.ceil(math.sqrt(world_count)))
    return np.array([(world % side) * spacing, (world // side) * spacing, 0.0])


def _set_free_joint_state(b: SceneBuilder, joint: int, pose: np.ndarray, vel_public: np.ndarray):
    qs = b.joint_q_start[joint]
    b.joint_q[qs : qs + 7] = list(pose)
    ds = b.joint_qd_start[joint]
    b.joint_qd[ds : ds + 6] = list(vel_public)


def build_ant_scene(
    num_worlds: int = 64,
    seed: int = 1,
    *,
    spacing: float = 0.0,
    device: str = "cpu",
) -> SapModel:
    """64-world ant batch: torso sphere r=0.25 on a free joint + 4 legs x 2 hinges.

    Capsule links r=0.04, segment lengths U(0.2, 0.4); mu=1.0; per-shape
    stiffness 2e5 N/m and dissipation 5 s so each contact pair combines to
    k = 1e5 N/m (series, ``sap_helpers.py:198-207``) and tau = 10 s (sum,
    ``sap_helpers.py:226-237``).  Initial hinge angles U(-0.3, 0.3), base height
    U(0.4, 0.8), all velocities N(0, 0.5).  No drives or limits.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=0.02)
    shape = ShapeCfg(density=1000.0, ke=2.0e5, tau=5.0, mu=1.0, gap=0.02)
    b.world_count = num_worlds
    for w in range(num_worlds):
        b.current_world = w
        off = _grid_offset(w, num_worlds, spacing)
        height = rng.uniform(0.4, 0.8)
        yaw = rng.uniform(-math.pi, math.pi)
        root_q = quat_from_axis_angle((0, 0, 1), yaw)
        root_pose = tf(off + np.array([0.0, 0.0, height]), root_q)
        torso = b.add_body(root_pose)
        b.add_shape(torso, GEO_SPHERE, tf(), (0.25, 0.0, 0.0), shape)
        joints = [b.add_free_joint(torso)]
        for leg in range(4):
            ang = math.pi / 4 + leg * math.pi / 2
            radial = np.array([math.cos(ang), math.sin(ang), 0.0])
            tangent = np.array([-math.sin(ang), math.cos(ang), 0.0])
            l1 = rng.uniform(0.2, 0.4)
            l2 = rng.uniform(0.2, 0.4)
            hip_local = 0.25 * radial
            thigh = b.add_body(tf_mul(root_pose, tf(hip_local)))
            b.add_shape(thigh, GEO_CAPSULE, tf(0.5 * l1 * radial, quat_z_to(radial)), (0.04, 0.5 * l1, 0.0), shape)
            joints.append(
                b.add_joint(torso, thigh, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=(0, 0, 1), limit_ke=0.0)],
                            parent_xform=tf(hip_local))
            )
            down = 0.5 * radial - 0.866 * np.array([0.0, 0.0, 1.0])
            down /= np.linalg.norm(down)
            shin = b.add_body(tf_mul(root_pose, tf(hip_local + l1 * radial)))
            b.add_shape(shin, GEO_CAPSULE, tf(0.5 * l2 * down, quat_z_to(down)), (0.04, 0.5 * l2, 0.0), shape)
            joints.append(
                b.add_joint(thigh, shin, SAP_JOINT_REVOLUTE, angular_axes=[DofCfg(axis=tuple(tangent), limit_ke=0.0)],
                            parent_xform=tf(l1 * radial))
            )
        b.add_articulation(joints)
        # initial state
        lin = rng.normal(0.0, 0.5, size=3)
        angv = rng.normal(0.0, 0.5, size=3)
        _set_free_joint_state(b, joints[0], root_pose, np.concatenate([lin, angv]))
        for j in joints[1:]:
            b.joint_q[b.joint_q_start[j]] = rng.uniform(-0.3, 0.3)
            b.joint_qd[b.joint_qd_start[j]] = rng.normal(0.0, 0.5)
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=5.0, mu=1.0, gap=0.02, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


# Humanoid kinematic template: (name, parent, axis, parent offset, shape)
# shape: None | ("capsule", center, direction, r, half_height) | ("sphere", center, r) | ("box", center, half)
def _humanoid_template():
    t = []
    t.append(("pelvis", None, None, (0.0, 0.0, 0.0), ("capsule", (0, 0, 0.02), (0, 1, 0), 0.07, 0.05)))
    for side, sy in (("l", 1.0), ("r", -1.0)):
        t.append((f"{side}_hip_p", "pelvis", (0, 1, 0), (0.0, 0.09 * sy, -0.05), None))
        t.append((f"{side}_hip_r", f"{side}_hip_p", (1, 0, 0), (0.0, 0.0, 0.0), None))
        t.append((f"{side}_thigh", f"{side}_hip_r", (0, 0, 1), (0.0, 0.0, 0.0),
                  ("capsule", (0, 0, -0.22), (0, 0, 1), 0.06, 0.13)))
        t.append((f"{side}_shin", f"{side}_thigh", (0, 1, 0), (0.0, 0.0, -0.4),
                  ("capsule", (0, 0, -0.2), (0, 0, 1), 0.05, 0.15)))
        t.append((f"{side}_ankle", f"{side}_shin", (0, 1, 0), (0.0, 0.0, -0.4), None))
        t.append((f"{side}_foot", f"{side}_ankle", (1, 0, 0), (0.0, 0.0, 0.0),
                  ("box", (0.04, 0.0, -0.05), (0.10, 0.05, 0.025))))
    t.append(("waist_y", "pelvis", (0, 0, 1), (0.0, 0.0, 0.10), None))
    t.append(("waist_r", "waist_y", (1, 0, 0), (0.0, 0.0, 0.0), None))
    t.append(("torso", "waist_r", (0, 1, 0), (0.0, 0.0, 0.0),
              ("capsule", (0, 0, 0.25), (0, 0, 1), 0.12, 0.10), ("sphere", (0, 0, 0.60), 0.10)))
    for side, sy in (("l", 1.0), ("r", -1.0)):
        t.append((f"{side}_sh_p", "torso", (0, 1, 0), (0.0, 0.2 * sy, 0.35), None))
        t.append((f"{side}_sh_r", f"{side}_sh_p", (1, 0, 0), (0.0, 0.0, 0.0), None))
        t.append((f"{side}_uarm", f"{side}_sh_r", (0, 0, 1), (0.0, 0.0, 0.0),
                  ("capsule", (0, 0, -0.13), (0, 0, 1), 0.045, 0.10)))
        t.append((f"{side}_farm", f"{side}_uarm", (0, 1, 0), (0.0, 0.0, -0.26),
                  ("capsule", (0, 0, -0.12), (0, 0, 1), 0.04, 0.09)))
        t.append((f"{side}_wr_r", f"{side}_farm", (0, 0, 1), (0.0, 0.0, -0.24), None))
        t.append((f"{side}_wr_p", f"{side}_wr_r", (0, 1, 0), (0.0, 0.0, 0.0), None))
        t.append((f"{side}_hand", f"{side}_wr_p", (1, 0, 0), (0.0, 0.0, 0.0), ("sphere", (0, 0, -0.05), 0.045)))
    return t


HUMANOID_STAND_HEIGHT = 0.05 + 0.8 + 0.075


def build_humanoid_scene(
    num_worlds: int = 4096,
    seed: int = 2,
    *,
    spacing: float = 0.0,
    contact_gap: float = 0.01,
    lying_fraction: float = 0.0,
    device: str = "cpu",
) -> SapModel:
    """35-DOF humanoid batch: free pelvis + 29 revolute joints, 30 bodies.

    Tree: legs 2 x 6 (hip pitch/roll/yaw, knee, ankle pitch/roll), waist 3,
    arms 2 x 7 (shoulder pitch/roll/yaw, elbow, wrist roll/pitch/yaw); the
    head is a sphere rigidly attached to the torso.  Every body mass is
    U(0.5, 5) kg (the shapes' density is scaled to it; shapeless joint links
    get a 0.05 m sphere-like inertia).  Per-body friction mu ~ U(0.6, 1.2),
    per-shape stiffness 2e5 N/m and dissipation 5 ms (pair: 1e5 N/m, 10 ms).
    Links are capsules and spheres; the feet are boxes in their own collision
    groups (2 and 3) so they touch only the ground (group -1).  PD position
    drives on every revolute DOF (ke=200, kd=10, effort limit 300, target =
    initial angle), limits +-2 rad, armature 0.01.  Half of the worlds start
    standing on the ground, half start falling from U(1.0, 1.6) m with tilted
    bases and random joint angles and velocities.

    One template world is built with the builder (unit masses) and replicated
    with vectorised per-world parameters; inertia scales linearly with mass
    for a fixed geometry, so this equals building every world separately.

    ``lying_fraction`` (default 0) lays that fraction of the worlds (drawn
    from a separate seeded stream, so the other worlds are unchanged) on
    their back or front with joint angles U(-0.1, 0.1) rad and small
    velocities, lowered until the lowest shape point is U(0, 4) mm into the
    ground: worlds with tens of contacts from the first step.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    W = int(num_worlds)
    tb = SceneBuilder(rigid_gap=contact_gap)
    tb.world_count = 1
    tb.current_world = 0
    template = _humanoid_template()
    names: dict = {}
    joints = []
    world_pose: dict = {}
    root_pose = tf((0.0, 0.0, HUMANOID_STAND_HEIGHT))
    body_group = []
    for entry in template:
        name, parent, axis, p_off = entry[0], entry[1], entry[2], entry[3]
        shapes = [x for x in entry[4:] if x is not None]
        pose = root_pose if parent is None else tf_mul(world_pose[parent], tf(p_off))
        world_pose[name] = pose
        if shapes:
            body = tb.add_body(pose)
            vol = 0.0
            for x in shapes:
                if x[0] == "capsule":
                    vol += shape_volume(GEO_CAPSULE, (x[3], x[4], 0.0))
                elif x[0] == "sphere":
                    vol += shape_volume(GEO_SPHERE, (x[2], 0.0, 0.0))
                else:
                    vol += shape_volume(GEO_BOX, x[2])
            density = 1.0 / vol
        else:
            r = 0.05
            body = tb.add_body(pose, mass=1.0, inertia=np.eye(3) * (0.4 * r * r))
            density = 0.0
        names[name] = body
        if parent is None:
            joints.append(tb.add_free_joint(body))
        else:
            dof = DofCfg(axis=axis, target_ke=200.0, target_kd=10.0, actuator_mode=1, limit_lower=-2.0,
                         limit_upper=2.0, limit_ke=1.0e4, limit_kd=10.0, armature=0.01, effort_limit=300.0)
            joints.append(tb.add_joint(names[parent], body, SAP_JOINT_REVOLUTE, angular_axes=[dof],
                                       parent_xform=tf(p_off)))
        group = 2 if name == "l_foot" else (3 if name == "r_foot" else 1)
        cfg = ShapeCfg(density=density, ke=2.0e5, tau=0.005, mu=1.0, gap=contact_gap, collision_group=group)
        for x in shapes:
            if x[0] == "capsule":
                tb.add_shape(body, GEO_CAPSULE, tf(x[1], quat_z_to(x[2])), (x[3], x[4], 0.0), cfg)
            elif x[0] == "sphere":
                tb.add_shape(body, GEO_SPHERE, tf(x[1]), (x[2], 0.0, 0.0), cfg)
            else:
                tb.add_shape(body, GEO_BOX, tf(x[1]), tuple(x[2]), cfg)
    tb.add_articulation(joints)
    tb.current_world = -1
    tb.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=1.0, gap=contact_gap, collision_group=-1))
    t = tb.finalize(device)
    J, NB, N, NQ = t.joint_count, t.body_count, t.joint_dof_count, t.joint_coord_count
    S = t.shape_count - 1  # world shapes; the ground plane is the last shape

    # ---- per-world random parameters (vectorised, fixed draw order)
    mass = rng.uniform(0.5, 5.0, size=(W, NB))
    mu_body = rng.uniform(0.6, 1.2, size=(W, NB))
    falling = rng.random(W) < 0.5
    height = np.where(falling, rng.uniform(1.0, 1.6, size=W), HUMANOID_STAND_HEIGHT + rng.uniform(0.0, 0.01, size=W))
    yaw = rng.uniform(-math.pi, math.pi, size=W)
    tilt_axis = rng.normal(size=(W, 2))
    tilt = np.where(falling, rng.uniform(0.0, 0.6, size=W), 0.0)
    qscale = np.where(falling, 0.3, 0.03)
    vscale = np.where(falling, 0.5, 0.05)
    root_lin = rng.normal(size=(W, 3)) * vscale[:, None]
    root_ang = rng.normal(size=(W, 3)) * vscale[:, None]
    rev_q = rng.uniform(-1.0, 1.0, size=(W, J - 1)) * qscale[:, None]
    rev_qd = rng.normal(size=(W, J - 1)) * vscale[:, None]

    def tile(a, offset=None):
        a = np.asarray(a)
        reps = np.concatenate([a] * W, axis=0) if W > 1 else a.copy()
        if offset is not None:
            off = np.repeat(np.arange(W) * offset, a.shape[0]).reshape((-1,) + (1,) * (a.ndim - 1))
            reps = np.where(reps >= 0, reps + off, reps)
        return reps

    shape_w = np.asarray(t.shape_transform[:S])
    # root orientation: yaw * tilt
    qr = np.zeros((W, 4))
    for w in range(W):
        ax = np.array([tilt_axis[w, 0], tilt_axis[w, 1], 0.0])
        qr[w] = quat_mul(quat_from_axis_angle((0, 0, 1), yaw[w]), quat_from_axis_angle(ax, tilt[w]))
    off = np.stack([_grid_offset(w, W, spacing) for w in range(W)]) if spacing else np.zeros((W, 3))
    joint_q = np.tile(t.joint_q.astype(np.float64), (W, 1))
    joint_qd = np.tile(t.joint_qd.astype(np.float64), (W, 1))
    joint_q[:, 0:3] = off + np.stack([np.zeros(W), np.zeros(W), height], axis=1)
    joint_q[:, 3:7] = qr
    joint_qd[:, 0:3] = root_lin
    joint_qd[:, 3:6] = root_ang
    joint_q[:, 7:] = rev_q
    joint_qd[:, 6:] = rev_qd
    target_pos = np.tile(t.joint_target_pos.astype(np.float64), (W, 1))
    target_pos[:, 6:] = rev_q
    lying = np.zeros(W, dtype=bool)
    if lying_fraction > 0.0:
        lrng = np.random.Generator(np.random.PCG64(seed + 7919))
        lying = lrng.random(W) < float(lying_fraction)
        for w in np.nonzero(lying)[0]:
            side = math.pi / 2 if lrng.random() < 0.5 else -math.pi / 2
            joint_q[w, 0:3] = off[w] + np.array([0.0, 0.0, 0.5])
            joint_q[w, 3:7] = quat_mul(quat_from_axis_angle((0, 0, 1), yaw[w]), quat_from_axis_angle((1, 0, 0), side))
            joint_qd[w, 0:6] = lrng.normal(size=6) * 0.05
            joint_q[w, 7:] = lrng.uniform(-0.1, 0.1, size=J - 1)
            joint_qd[w, 6:] = lrng.normal(size=J - 1) * 0.05
            target_pos[w, 6:] = joint_q[w, 7:]
    body_mass = mass.reshape(-1)
    body_inertia = (np.tile(t.body_inertia, (W, 1, 1)) * body_mass[:, None, None]).astype(np.float32)
    mu_shape = mu_body[:, t.shape_body[:S]].reshape(-1)
    ground = S
    shape_pairs_t = np.asarray(t.shape_pairs, dtype=np.int64)
    pairs = []
    for w in range(W):
        pw = shape_pairs_t.copy()
        pw = np.where(pw == ground, W * S, pw + w * S)
        pairs.append(pw)
    # the builder lists all ground pairs first (world -1 sorts first), then per-world pairs
    pairs = np.concatenate(pairs, axis=0)
    is_ground = np.any(pairs == W * S, axis=1)
    pairs = np.concatenate([pairs[is_ground], pairs[~is_ground]], axis=0).astype(np.int32)
    filt_t = np.asarray(t.shape_collision_filter_pairs, dtype=np.int64).reshape(-1, 2)
    filt = np.concatenate([np.where(filt_t == ground, W * S, filt_t + w * S) for w in range(W)], axis=0)
    filt = np.unique(filt, axis=0).astype(np.int32)

    def cat_shapes(a_world, a_ground):
        return np.concatenate([np.tile(a_world, (W,) + (1,) * (a_world.ndim - 1)), a_ground], axis=0)

    model = SapModel(
        device=device,
        joint_count=J * W,
        joint_dof_count=N * W,
        joint_coord_count=NQ * W,
        body_count=NB * W,
        articulation_count=W,
        world_count=W,
        shape_count=S * W + 1,
        joint_type=tile(t.joint_type).astype(np.int32),
        joint_articulation=np.repeat(np.arange(W), J).astype(np.int32),
        joint_parent=tile(t.joint_parent, NB).astype(np.int32),
        joint_child=tile(t.joint_child, NB).astype(np.int32),
        joint_q_start=np.concatenate([(np.arange(W)[:, None] * NQ + t.joint_q_start[:-1]).reshape(-1), [NQ * W]]).astype(np.int32),
        joint_qd_start=np.concatenate([(np.arange(W)[:, None] * N + t.joint_qd_start[:-1]).reshape(-1), [N * W]]).astype(np.int32),
        joint_dof_dim=tile(t.joint_dof_dim).astype(np.int32),
        joint_axis=tile(t.joint_axis).astype(np.float32),
        joint_X_p=tile(t.joint_X_p).astype(np.float32),
        joint_X_c=tile(t.joint_X_c).astype(np.float32),
        joint_ancestor=tile(t.joint_ancestor, J).astype(np.int32),
        articulation_start=(np.arange(W + 1) * J).astype(np.int32),
        joint_q=joint_q.reshape(-1).astype(np.float32),
        joint_qd=joint_qd.reshape(-1).astype(np.float32),
        body_q=np.zeros((NB * W, 7), np.float32),
        body_qd=np.zeros((NB * W, 6), np.float32),
        body_mass=body_mass.astype(np.float32),
        body_inertia=body_inertia,
        body_com=tile(t.body_com).astype(np.float32),
        body_flags=tile(t.body_flags).astype(np.int32),
        body_world=np.repeat(np.arange(W), NB).astype(np.int32),
        gravity=np.tile(t.gravity[:1], (W, 1)).astype(np.float32),
        joint_armature=tile(t.joint_armature).astype(np.float32),
        shape_body=np.concatenate([tile(t.shape_body[:S], NB), [-1]]).astype(np.int32),
        shape_material_mu=np.concatenate([mu_shape, t.shape_material_mu[S:]]).astype(np.float32),
        shape_material_ke=cat_shapes(t.shape_material_ke[:S], t.shape_material_ke[S:]).astype(np.float32),
        shape_material_tau=cat_shapes(t.shape_material_tau[:S], t.shape_material_tau[S:]).astype(np.float32),
        joint_target_mode=tile(t.joint_target_mode).astype(np.int32),
        joint_target_ke=tile(t.joint_target_ke).astype(np.float32),
        joint_target_kd=tile(t.joint_target_kd).astype(np.float32),
        joint_target_pos=target_pos.reshape(-1).astype(np.float32),
        joint_target_vel=tile(t.joint_target_vel).astype(np.float32),
        joint_f=np.zeros(N * W, np.float32),
        joint_act=np.zeros(N * W, np.float32),
        joint_effort_limit=tile(t.joint_effort_limit).astype(np.float32),
        joint_limit_lower=tile(t.joint_limit_lower).astype(np.float32),
        joint_limit_upper=tile(t.joint_limit_upper).astype(np.float32),
        joint_limit_ke=tile(t.joint_limit_ke).astype(np.float32),
        joint_limit_kd=tile(t.joint_limit_kd).astype(np.float32),
        shape_transform=cat_shapes(shape_w, t.shape_transform[S:]).astype(np.float32),
        shape_type=cat_shapes(t.shape_type[:S], t.shape_type[S:]).astype(np.int32),
        shape_scale=cat_shapes(t.shape_scale[:S], t.shape_scale[S:]).astype(np.float32),
        shape_flags=cat_shapes(t.shape_flags[:S], t.shape_flags[S:]).astype(np.int32),
        shape_margin=cat_shapes(t.shape_margin[:S], t.shape_margin[S:]).astype(np.float32),
        shape_gap=cat_shapes(t.shape_gap[:S], t.shape_gap[S:]).astype(np.float32),
        shape_collision_radius=cat_shapes(t.shape_collision_radius[:S], t.shape_collision_radius[S:]).astype(np.float32),
        shape_collision_group=cat_shapes(t.shape_collision_group[:S], t.shape_collision_group[S:]).astype(np.int32),
        shape_world=np.concatenate([np.repeat(np.arange(W), S), [-1]]).astype(np.int32),
        shape_pairs=pairs,
        shape_collision_filter_pairs=filt,
    )
    model = _with_consistent_body_q(model)
    if np.any(lying):
        pen = np.random.Generator(np.random.PCG64(seed + 7920)).uniform(0.0, 0.004, size=W)
        low = _lowest_shape_points(model)
        jq = model.joint_q.astype(np.float64).reshape(W, NQ)
        jq[lying, 2] += -low[lying] - pen[lying]
        model = _with_consistent_body_q(replace(model, joint_q=jq.reshape(-1).astype(np.float32)))
    return model


def _lowest_shape_points(model: SapModel) -> np.ndarray:
    """Per world, the lowest world-z of its shapes' surfaces (spheres, capsules, boxes)."""
    low = np.full(model.world_count, np.inf)
    for s in range(model.shape_count):
        w, b = int(model.shape_world[s]), int(model.shape_body[s])
        if w < 0 or b < 0:
            continue
        x = tf_mul(tuple(model.body_q[b]), tuple(model.shape_transform[s]))
        R = quat_to_matrix(np.asarray(x[3:7], dtype=np.float64))
        sc = model.shape_scale[s]
        t = int(model.shape_type[s])
        if t == GEO_SPHERE:
            z = x[2] - sc[0]
        elif t == GEO_CAPSULE:
            z = x[2] - abs(R[2, 2]) * sc[1] - sc[0]
        elif t == GEO_BOX:
            z = x[2] - (abs(R[2, 0]) * sc[0] + abs(R[2, 1]) * sc[1] + abs(R[2, 2]) * sc[2])
        else:
            continue
        low[w] = min(low[w], z)
    return low


def build_box_pile_scene(
    num_worlds: int = 64,
    seed: int = 5,
    *,
    spacing: float = 0.0,
    contact_gap: float = 0.01,
    device: str = "cpu",
) -> SapModel:
    """Free boxes and capsules piled on the ground: the convex (MPR / GJK) contact path.

    Six free bodies per world, each on its own free joint and articulation:
    a slab box (half 0.25 x 0.2 x 0.1) on the ground; a box (half 0.1 x 0.1 x
    0.08) on the slab, yawed and slightly tilted; a capsule (r 0.04, half
    0.12) lying on the slab; a capsule (r 0.03, half 0.06) standing on the
    box; a small tilted box (half 0.06 x 0.05 x 0.04) above the box; a cube
    (half 0.1) beside the slab, facing its side.  Every resting offset is a
    penetration drawn from U(-0.005, 0.01) m (negative: a gap), positions
    jitter by U(-0.02, 0.02) m and yaws are uniform, so the batch holds
    face-face, edge-face, separated-within-gap and overhanging manifolds of
    box-box and capsule-box pairs next to the plane-box / plane-capsule
    primitives.  Density 500 kg/m^3, mu 0.8, per-shape stiffness 2e5 N/m and
    dissipation 5 ms; bodies start at rest.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    cfg = ShapeCfg(density=500.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap, collision_group=1)
    b.world_count = num_worlds

    def pen():
        return rng.uniform(-0.005, 0.01)

    def jitter():
        return np.array([rng.uniform(-0.02, 0.02), rng.uniform(-0.02, 0.02), 0.0])

    def yawq(tilt_axis=None, tilt=0.0):
        q = quat_from_axis_angle((0, 0, 1), rng.uniform(-math.pi, math.pi))
        if tilt_axis is not None:
            q = quat_mul(q, quat_from_axis_angle(tilt_axis, tilt))
        return q

    for w in range(num_worlds):
        b.current_world = w
        off = _grid_offset(w, num_worlds, spacing)
        bodies = []
        # slab on the ground
        slab_top = 0.2 - pen()
        bodies.append((tf(off + np.array([0.0, 0.0, slab_top - 0.1]) + jitter(),
                          quat_from_axis_angle((0, 0, 1), rng.uniform(-0.3, 0.3))),
                       [(GEO_BOX, tf(), (0.25, 0.2, 0.1))]))
        # box on the slab (yawed, tilted by up to 4 degrees)
        box_top = slab_top + 0.16 - pen()
        bodies.append((tf(off + np.array([0.05, 0.0, box_top - 0.08]) + jitter(),
                          yawq((1, 0, 0), rng.uniform(-0.07, 0.07))),
                       [(GEO_BOX, tf(), (0.1, 0.1, 0.08))]))
        # capsule lying on the slab
        bodies.append((tf(off + np.array([-0.12, 0.0, slab_top + 0.04 - pen()]) + jitter(), yawq()),
                       [(GEO_CAPSULE, tf((0.0, 0.0, 0.0), quat_z_to((1.0, 0.0, 0.0))), (0.04, 0.12, 0.0))]))
        # capsule standing on the box
        bodies.append((tf(off + np.array([0.05, 0.0, box_top + 0.09 - pen()]) + jitter(),
                          quat_from_axis_angle((1, 0, 0), rng.uniform(-0.05, 0.05))),
                       [(GEO_CAPSULE, tf(), (0.03, 0.06, 0.0))]))
        # small tilted box above the box
        bodies.append((tf(off + np.array([0.12, 0.05, box_top + 0.045 - pen()]) + jitter(),
                          yawq((0, 1, 0), rng.uniform(-0.6, 0.6))),
                       [(GEO_BOX, tf(), (0.06, 0.05, 0.04))]))
        # cube beside the slab, against its +x side face
        bodies.append((tf(off + np.array([0.35 - pen(), 0.0, 0.1 - pen()]) + jitter(),
                          quat_from_axis_angle((0, 0, 1), rng.uniform(-0.1, 0.1))),
                       [(GEO_BOX, tf(), (0.1, 0.1, 0.1))]))
        for pose, shapes in bodies:
            body = b.add_body(pose)
            for st, x, scale in shapes:
                b.add_shape(body, st, x, scale, cfg)
            j = b.add_free_joint(body)
            b.add_articulation([j])
            _set_free_joint_state(b, j, pose, np.zeros(6))
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_round_shapes_scene(
    num_worlds: int = 64,
    seed: int = 9,
    *,
    spacing: float = 0.0,
    contact_gap: float = 0.01,
    device: str = "cpu",
) -> SapModel:
    """Free cylinders, ellipsoids and spheres on the ground: the plane-cylinder, plane-ellipsoid and
    sphere-cylinder primitives (``primitive.py:322-351, 431-657``).

    Six free bodies per world: an upright cylinder (tilt up to 15 degrees: the near-cap tripod
    mode) with a sphere resting on its top cap; a cylinder lying on its side (rolling mode, axis up
    to 20 degrees off horizontal) with a sphere against its curved side; two ellipsoids of random
    semi-axes U(0.04, 0.12) m at random orientations.  Cylinder radii U(0.05, 0.1) m, half heights
    U(0.05, 0.12) m, sphere radii U(0.03, 0.06) m.  Each cylinder shares a collision group with its
    sphere and each ellipsoid has a group of its own, so every candidate pair is one of the three
    primitives or a ground pair of a sphere.  Every resting offset is a penetration U(0, 4) mm;
    density U(400, 1000) kg/m^3, mu 0.7, per-shape stiffness 2e5 N/m and dissipation 5 ms; bodies
    start at rest.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = num_worlds

    def cfg(group):
        return ShapeCfg(density=rng.uniform(400.0, 1000.0), ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap,
                        collision_group=group)

    def pen():
        return rng.uniform(0.0, 0.004)

    for w in range(num_worlds):
        b.current_world = w
        off = _grid_offset(w, num_worlds, spacing)
        bodies = []
        # upright cylinder + sphere on its top cap
        r, hh = rng.uniform(0.05, 0.1), rng.uniform(0.05, 0.12)
        q = quat_mul(quat_from_axis_angle((0, 0, 1), rng.uniform(-math.pi, math.pi)),
                     quat_from_axis_angle((1, 0, 0), rng.uniform(0.0, 0.26)))
        a = quat_rotate(q, (0.0, 0.0, 1.0))
        low = r * math.sqrt(max(1.0 - a[2] * a[2], 0.0)) + hh * abs(a[2])
        c = off + np.array([-0.3, 0.0, low - pen()])
        bodies.append((tf(c, q), GEO_CYLINDER, (r, hh, 0.0), 1))
        rs = rng.uniform(0.03, 0.06)
        bodies.append((tf(c + a * (hh + rs - pen())), GEO_SPHERE, (rs, 0.0, 0.0), 1))
        # cylinder on its side + sphere against its curved side
        r, hh = rng.uniform(0.05, 0.1), rng.uniform(0.05, 0.12)
        yaw = rng.uniform(-math.pi, math.pi)
        q = quat_mul(quat_from_axis_angle((0, 0, 1), yaw),
                     quat_from_axis_angle((1, 0, 0), math.pi / 2 - rng.uniform(0.0, 0.35)))
        a = quat_rotate(q, (0.0, 0.0, 1.0))
        low = r * math.sqrt(max(1.0 - a[2] * a[2], 0.0)) + hh * abs(a[2])
        c = off + np.array([0.3, 0.0, low - pen()])
        bodies.append((tf(c, q), GEO_CYLINDER, (r, hh, 0.0), 2))
        side = np.cross(a, (0.0, 0.0, 1.0))
        side = side / np.linalg.norm(side)
        rs = rng.uniform(0.03, 0.06)
        bodies.append((tf(c + side * (r + rs - pen())), GEO_SPHERE, (rs, 0.0, 0.0), 2))
        # two ellipsoids at random orientations
        for k in range(2):
            semi = tuple(rng.uniform(0.04, 0.12, size=3))
            q = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, math.pi))
            R = quat_to_matrix(q)
            low = math.sqrt(sum((R[2, i] * semi[i]) ** 2 for i in range(3)))
            bodies.append((tf(off + np.array([0.0, -0.3 + 0.6 * k, low - pen()]), q), GEO_ELLIPSOID, semi, 3 + k))
        for pose, st, scale, group in bodies:
            body = b.add_body(pose)
            b.add_shape(body, st, tf(), scale, cfg(group))
            j = b.add_free_joint(body)
            b.add_articulation([j])
            _set_free_joint_state(b, j, pose, np.zeros(6))
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.7, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_convex_shapes_scene(
    num_worlds: int = 64,
    seed: int = 11,
    *,
    spacing: float = 0.0,
    contact_gap: float = 0.01,
    device: str = "cpu",
) -> SapModel:
    """Cones, cylinders, ellipsoids and a capsule resting on each other: the pairs the analytic
    routines leave to MPR / GJK (``narrow_phase.py:570-747``), including an infinite plane against a
    cone (box proxy, ``core.py:507-566``), the single-contact rule of sphere and ellipsoid pairs
    (``core.py:360-378``) and the rolling projection of cylinders and cones on boxes and planes
    (``core.py:172-275``).

    Nine free bodies per world, in four collision groups:

    * group 1: a box slab (half extents 0.2 x 0.2 x 0.05 m) on the ground, a cylinder lying on it
      (axis within 1.5 degrees of horizontal: the rolling mode) and an upright cone standing on its
      base on it (box-cylinder, box-cone, and cone-cylinder out of reach);
    * group 2: a cone lying on its slant on the ground (plane-cone in the rolling mode);
    * group 3: two ellipsoids of equal vertical semi-axis side by side on the ground, touching at
      their equators (ellipsoid-ellipsoid), and a sphere resting on the first (sphere-ellipsoid);
    * group 4: a cylinder lying on the ground and a capsule lying across it (capsule-cylinder).

    Radii and half heights U(0.04, 0.08) m, ellipsoid semi-axes U(0.05, 0.1) m; every resting offset
    is a penetration U(0, 3) mm; density U(400, 1000) kg/m^3, mu 0.6, stiffness 2e5 N/m, dissipation
    5 ms; bodies start at rest.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = num_worlds

    def cfg(group):
        return ShapeCfg(density=rng.uniform(400.0, 1000.0), ke=2.0e5, tau=0.005, mu=0.6, gap=contact_gap,
                        collision_group=group)

    def pen():
        return rng.uniform(0.0, 0.003)

    def yaw_q(extra=None):
        q = quat_from_axis_angle((0, 0, 1), rng.uniform(-math.pi, math.pi))
        return q if extra is None else quat_mul(q, extra)

    for w in range(num_worlds):
        b.current_world = w
        off = _grid_offset(w, num_worlds, spacing)
        bodies = []
        # group 1: slab, rolling cylinder and upright cone on it
        c0 = off + np.array([0.0, 0.0, 0.05 - pen()])
        bodies.append((tf(c0, yaw_q()), GEO_BOX, (0.2, 0.2, 0.05), 1))
        top = c0[2] + 0.05
        r, hh = rng.uniform(0.04, 0.08), rng.uniform(0.04, 0.08)
        tilt = math.radians(rng.uniform(0.0, 1.5))
        q = quat_mul(quat_from_axis_angle((0, 0, 1), rng.uniform(-0.3, 0.3)),
                     quat_from_axis_angle((1, 0, 0), math.pi / 2 - tilt))
        a = quat_rotate(q, (0.0, 0.0, 1.0))
        low = r * math.sqrt(max(1.0 - a[2] * a[2], 0.0)) + hh * abs(a[2])
        bodies.append((tf(c0 + np.array([0.0, -0.1, top - c0[2] + low - pen()]), q), GEO_CYLINDER, (r, hh, 0.0), 1))
        r, hh = rng.uniform(0.04, 0.08), rng.uniform(0.04, 0.08)
        bodies.append((tf(c0 + np.array([0.0, 0.1, top - c0[2] + hh - pen()]), yaw_q()), GEO_CONE, (r, hh, 0.0), 1))
        # group 2: a cone lying on its slant on the ground: the axis points down by the half angle
        r, hh = rng.uniform(0.04, 0.08), rng.uniform(0.04, 0.08)
        alpha = math.atan2(r, 2.0 * hh)
        q = yaw_q(quat_from_axis_angle((0, 1, 0), math.pi / 2 + alpha))
        bodies.append((tf(off + np.array([0.45, 0.0, hh * math.sin(alpha) - pen()]), q), GEO_CONE, (r, hh, 0.0), 2))
        # group 3: two ellipsoids side by side, a sphere against the first
        e1 = tuple(rng.uniform(0.05, 0.1, size=3))
        e2 = (rng.uniform(0.05, 0.1), rng.uniform(0.05, 0.1), e1[2])  # same height: equators touch
        z = e1[2] - pen()
        p1 = off + np.array([-0.45, 0.0, z])
        p2 = off + np.array([-0.45, e1[1] + e2[1] - pen(), z])
        bodies.append((tf(p1), GEO_ELLIPSOID, e1, 3))
        bodies.append((tf(p2), GEO_ELLIPSOID, e2, 3))
        rs = rng.uniform(0.03, 0.05)
        bodies.append((tf(p1 + np.array([0.0, 0.0, e1[2] + rs - pen()])), GEO_SPHERE, (rs, 0.0, 0.0), 3))
        # group 4: a cylinder on its side on the ground, a capsule across it
        r, hh = rng.uniform(0.04, 0.08), rng.uniform(0.04, 0.08)
        q = quat_from_axis_angle((1, 0, 0), math.pi / 2)
        pc = off + np.array([0.0, 0.45, r - pen()])
        bodies.append((tf(pc, q), GEO_CYLINDER, (r, hh, 0.0), 4))
        rc, hc = rng.uniform(0.02, 0.04), rng.uniform(0.06, 0.1)
        qc = quat_from_axis_angle((0, 1, 0), math.pi / 2 + rng.uniform(-0.2, 0.2))
        bodies.append((tf(pc + np.array([0.0, 0.0, r + rc - pen()]), qc), GEO_CAPSULE, (rc, hc, 0.0), 4))
        for pose, st, scale, group in bodies:
            body = b.add_body(pose)
            b.add_shape(body, st, tf(), scale, cfg(group))
            j = b.add_free_joint(body)
            b.add_articulation([j])
            _set_free_joint_state(b, j, pose, np.zeros(6))
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.6, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_ball_chain_scene(
    num_worlds: int = 64,
    seed: int = 7,
    *,
    links: int = 4,
    anchor_height: float = 1.19,
    max_tilt: float = 0.35,
    contact_gap: float = 0.01,
    device: str = "cpu",
) -> SapModel:
    """Chains of capsules on ball joints swinging onto a free box.

    Each world: ``links`` capsules (r 0.04, 0.24 m long) joined end to end by ball joints, the
    first hung from a world anchor ``anchor_height`` up; a free box (half 0.12 m) on the ground
    under the chain.  Every ball joint starts at a random rotation (angle U(0, ``max_tilt``) rad)
    with a random angular velocity N(0, 1) rad/s and a velocity-damping drive (kd 0.05); so
    3 * links + 6 DOF, capsule-box and box-plane contacts.  With the defaults a
    hanging chain's tip starts at most 1 cm into the box top (a chain hung lower starts several
    cm inside the box, and its first step is an ejection at 100+ rad/s rather than a contact).
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=contact_gap)
    b.world_count = num_worlds
    cfg = ShapeCfg(density=700.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap)
    half_len = 0.12
    for w in range(num_worlds):
        b.current_world = w
        parent, x = -1, tf((0.0, 0.0, anchor_height))
        joints = []
        for k in range(links):
            body = b.add_body(x)
            b.add_shape(body, GEO_CAPSULE, tf((0.0, 0.0, -half_len)), (0.04, half_len - 0.04, 0.0), cfg)
            axes = [DofCfg(axis=a, target_kd=0.05, actuator_mode=2, limit_ke=0.0, limit_kd=0.0)
                    for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
            px = tf((0.0, 0.0, anchor_height)) if parent < 0 else tf((0.0, 0.0, -2.0 * half_len))
            j = b.add_joint(parent, body, SAP_JOINT_BALL, angular_axes=axes, parent_xform=px)
            qs, ds = b.joint_q_start[j], b.joint_qd_start[j]
            b.joint_q[qs:qs + 4] = list(quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, max_tilt)))
            b.joint_qd[ds:ds + 3] = list(rng.normal(0.0, 1.0, 3))
            joints.append(j)
            parent, x = body, tf_mul(x, tf((0.0, 0.0, -2.0 * half_len)))
        b.add_articulation(joints)
        box_pose = tf((rng.uniform(-0.15, 0.15), rng.uniform(-0.15, 0.15), 0.121),
                      quat_from_axis_angle((0, 0, 1), rng.uniform(-math.pi, math.pi)))
        box = b.add_body(box_pose)
        b.add_shape(box, GEO_BOX, tf(), (0.12, 0.12, 0.12), ShapeCfg(density=300.0, ke=2.0e5, tau=0.005, mu=0.8,
                                                                    gap=contact_gap))
        jb = b.add_free_joint(box)
        b.add_articulation([jb])
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.005, mu=0.8, gap=contact_gap, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def build_d6_ant_scene(
    num_worlds: int = 64,
    seed: int = 3,
    *,
    device: str = "cpu",
) -> SapModel:
    """Ant batch whose legs use D6 joints (the joints the reference's MJCF / USD / URDF loaders
    make from bodies with several hinges or slides, ``loader/scene.py:2239-2251, 2856-2894,
    4930-4937``).

    Torso sphere r=0.25 on a free joint; four two-capsule legs (r=0.04, segments U(0.2, 0.35)):

    * leg 0: hip D6 with two angular axes (z, tangent), revolute knee;
    * leg 1: hip D6 with three angular axes (z, tangent, radial), one-axis D6 knee with limits
      (+-0.4 rad) and a PD drive (ke 20, kd 1);
    * leg 2: hip D6 with one linear (radial) and one angular (z) axis, revolute knee;
    * leg 3: hip D6 with two skew angular axes (orthogonalised by the joint), knee D6 with one
      linear axis along the shin and one angular axis, PD drives on both (ke 40, kd 2).

    6 + 3 + 4 + 3 + 4 = 20 DOF.  Hinge angles U(-0.3, 0.3), slides U(-0.02, 0.02), base height
    U(0.36, 0.6) (the lowest worlds start with a foot on the ground), velocities N(0, 0.5); mu 1, per-shape stiffness 2e5 N/m, tau 5 s, as the ant.
    """
    rng = np.random.Generator(np.random.PCG64(seed))
    b = SceneBuilder(rigid_gap=0.02)
    shape = ShapeCfg(density=1000.0, ke=2.0e5, tau=5.0, mu=1.0, gap=0.02)
    b.world_count = num_worlds
    z = np.array([0.0, 0.0, 1.0])

    def dof(axis, **kw):
        a = np.asarray(axis, dtype=np.float64)
        return DofCfg(axis=tuple(a / np.linalg.norm(a)), **({"limit_ke": 0.0} | kw))

    for w in range(num_worlds):
        b.current_world = w
        height = rng.uniform(0.36, 0.6)
        root_pose = tf(np.array([0.0, 0.0, height]), quat_from_axis_angle((0, 0, 1), rng.uniform(-math.pi, math.pi)))
        torso = b.add_body(root_pose)
        b.add_shape(torso, GEO_SPHERE, tf(), (0.25, 0.0, 0.0), shape)
        joints = [b.add_free_joint(torso)]
        for leg in range(4):
            ang = math.pi / 4 + leg * math.pi / 2
            radial = np.array([math.cos(ang), math.sin(ang), 0.0])
            tangent = np.array([-math.sin(ang), math.cos(ang), 0.0])
            l1, l2 = rng.uniform(0.2, 0.35), rng.uniform(0.2, 0.35)
            hip = 0.25 * radial
            down = 0.5 * radial - 0.866 * z
            down /= np.linalg.norm(down)
            thigh = b.add_body(tf_mul(root_pose, tf(hip)))
            b.add_shape(thigh, GEO_CAPSULE, tf(0.5 * l1 * radial, quat_z_to(radial)), (0.04, 0.5 * l1, 0.0), shape)
            shin = b.add_body(tf_mul(root_pose, tf(hip + l1 * radial)))
            b.add_shape(shin, GEO_CAPSULE, tf(0.5 * l2 * down, quat_z_to(down)), (0.04, 0.5 * l2, 0.0), shape)
            hip_x = dict(parent_xform=tf(hip))
            knee_x = dict(parent_xform=tf(l1 * radial))
            if leg == 0:
                jh = b.add_joint(torso, thigh, SAP_JOINT_D6, angular_axes=[dof(z), dof(tangent)], **hip_x)
                jk = b.add_joint(thigh, shin, SAP_JOINT_REVOLUTE, angular_axes=[dof(tangent)], **knee_x)
            elif leg == 1:
                jh = b.add_joint(torso, thigh, SAP_JOINT_D6, angular_axes=[dof(z), dof(tangent), dof(radial)], **hip_x)
                jk = b.add_joint(thigh, shin, SAP_JOINT_D6, angular_axes=[dof(tangent, limit_lower=-0.4, limit_upper=0.4,
                                                                              limit_ke=1.0e4, target_ke=20.0, target_kd=1.0,
                                                                              actuator_mode=1)],
                                 **knee_x)
            elif leg == 2:
                jh = b.add_joint(torso, thigh, SAP_JOINT_D6, linear_axes=[dof(radial)], angular_axes=[dof(z)], **hip_x)
                jk = b.add_joint(thigh, shin, SAP_JOINT_REVOLUTE, angular_axes=[dof(tangent)], **knee_x)
            else:
                jh = b.add_joint(torso, thigh, SAP_JOINT_D6, angular_axes=[dof(z + 0.3 * radial), dof(tangent + 0.2 * z)],
                                 **hip_x)
                pd = dict(target_ke=40.0, target_kd=2.0, actuator_mode=1)
                jk = b.add_joint(thigh, shin, SAP_JOINT_D6, linear_axes=[dof(down, **pd)],
                                 angular_axes=[dof(tangent, **pd)], **knee_x)
            joints += [jh, jk]
        b.add_articulation(joints)
        _set_free_joint_state(b, joints[0], root_pose, rng.normal(0.0, 0.5, size=6))
        for j in joints[1:]:
            qs, ds = b.joint_q_start[j], b.joint_qd_start[j]
            nl, na = b.joint_dof_dim[j]
            for k in range(nl + na):
                b.joint_q[qs + k] = rng.uniform(-0.02, 0.02) if k < nl else rng.uniform(-0.3, 0.3)
                b.joint_qd[ds + k] = rng.normal(0.0, 0.5)
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=5.0, mu=1.0, gap=0.02, collision_group=-1))
    return _with_consistent_body_q(b.finalize(device))


def _with_consistent_body_q(model: SapModel) -> SapModel:
    """Recompute body_q/body_qd from joint_q/joint_qd with the public FK."""
    from .kinematics_host import eval_fk_public

    body_q, body_qd = eval_fk_public(model, model.joint_q, model.joint_qd)
    return replace(model, body_q=body_q.astype(np.float32), body_qd=body_qd.astype(np.float32))


__all__ = ["SceneBuilder", "ShapeCfg", "DofCfg", "build_ant_scene", "build_ball_chain_scene", "build_box_pile_scene",
           "build_humanoid_scene", "build_round_shapes_scene", "build_convex_shapes_scene", "build_d6_ant_scene"]

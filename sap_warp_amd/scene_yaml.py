"""Inline YAML / JSON scene files -> ``SapModel`` (``load_sap_scene`` for inline scenes).

Reads the reference's scene schema (version 1) for scenes described inline —
``simulation``, ``builder`` (gravity, rigid gap, default body armature, joint
and shape defaults), ``bodies`` with box / sphere / capsule / cylinder / cone /
ellipsoid shapes, ``joints``
(fixed, free, revolute, prismatic), ``articulations``, ``initial_joint_q``,
``post_build`` operations, ``replicate`` and ``ground`` — and builds the model
with ``SceneBuilder`` (``sim/loader/scene.py:1372-1509, 5100-5780``).  Asset
entries (URDF / USD / MJCF files), meshes and the other shape types are
outside this build's path: they are reported as unsupported (an error when
``strict``).

Numbers may be written as arithmetic over ``pi`` (``"pi * 0.5"``), evaluated by
a restricted parser (no names other than ``pi``, no calls), as in
``scene.py:5752-5780``.
"""
from __future__ import annotations

import ast
import json
import math
import os
from dataclasses import dataclass, field, fields

import numpy as np

from .runtime import (GEO_BOX, GEO_CAPSULE, GEO_CONE, GEO_CYLINDER, GEO_ELLIPSOID, GEO_SPHERE, SAP_JOINT_FIXED,
                      SAP_JOINT_FREE, SAP_JOINT_PRISMATIC, SAP_JOINT_REVOLUTE, SHAPE_FLAG_COLLIDE_PARTICLES,
                      SHAPE_FLAG_COLLIDE_SHAPES, SHAPE_FLAG_VISIBLE)
from .scenes import DofCfg, SceneBuilder, ShapeCfg, _with_consistent_body_q, quat_from_axis_angle

SCHEMA_VERSION = 1
DEFAULT_SPACING = (2.0, 2.0, 0.0)


class SceneLoaderError(ValueError):
    """Malformed scene file (``SapSceneLoaderError``)."""


class UnsupportedSceneFeature(NotImplementedError):
    """A strict load met features outside this build (``SapUnsupportedSceneFeature``)."""

    def __init__(self, features):
        self.features = list(features)
        super().__init__("unsupported scene features: " + "; ".join(self.features))


@dataclass
class LoadedScene:
    """What a scene file defines: the model plus the simulation settings to step it with."""

    model: object
    name: str
    dt: float
    max_rigid_contact: int
    solver: dict
    body_ids: dict
    joint_ids: dict
    unsupported: list = field(default_factory=list)


# ---------------------------------------------------------------- numbers
def number(value) -> float:
    """int / float, or a string of arithmetic over ``pi`` (``_float``, scene.py:5752-5780)."""
    if isinstance(value, bool):
        raise SceneLoaderError(f"Expected numeric value, got {value!r}.")
    if isinstance(value, (int, float)):
        return float(value)
    if isinstance(value, str):
        return _eval_expr(value)
    raise SceneLoaderError(f"Expected numeric value, got {value!r}.")


def _eval_expr(expr: str) -> float:
    def visit(node):
        if isinstance(node, ast.Expression):
            return visit(node.body)
        if isinstance(node, ast.Constant) and isinstance(node.value, (int, float)) and not isinstance(node.value, bool):
            return float(node.value)
        if isinstance(node, ast.Name) and node.id == "pi":
            return math.pi
        if isinstance(node, ast.UnaryOp) and isinstance(node.op, (ast.UAdd, ast.USub)):
            v = visit(node.operand)
            return v if isinstance(node.op, ast.UAdd) else -v
        if isinstance(node, ast.BinOp) and isinstance(node.op, (ast.Add, ast.Sub, ast.Mult, ast.Div)):
            a, b = visit(node.left), visit(node.right)
            if isinstance(node.op, ast.Add):
                return a + b
            if isinstance(node.op, ast.Sub):
                return a - b
            if isinstance(node.op, ast.Mult):
                return a * b
            return a / b
        raise ValueError(f"Unsupported numeric expression: {expr!r}")

    return visit(ast.parse(expr, mode="eval"))


def _coerce(value):
    """Config value: numbers and numeric strings to float, anything else unchanged (scene.py:5737)."""
    if isinstance(value, bool):
        return value
    if isinstance(value, int):
        return int(value)
    if isinstance(value, float):
        return float(value)
    if isinstance(value, str):
        try:
            return number(value)
        except (ValueError, SyntaxError):
            return value
    return value


def _vec3(value):
    if not isinstance(value, (list, tuple)) or len(value) != 3:
        raise SceneLoaderError(f"Expected a 3-vector, got {value!r}.")
    return tuple(number(v) for v in value)


def _axis(value):
    """Joint axis: a 3-vector, an axis index or a token x / y / z (scene.py:1362-1370, 3525-3547)."""
    if isinstance(value, str):
        t = value.strip().upper()
        k = 0 if t in ("X", "AXIS.X", "0") else (1 if t in ("Y", "AXIS.Y", "1") else 2)
        return tuple(1.0 if i == k else 0.0 for i in range(3))
    if isinstance(value, (int, np.integer)) and not isinstance(value, bool):
        k = int(value)
        return tuple(1.0 if i == (k if k in (0, 1) else 2) else 0.0 for i in range(3))
    return _vec3(value)


def _quat(spec):
    if spec is None or spec == "identity":
        return (0.0, 0.0, 0.0, 1.0)
    if isinstance(spec, (list, tuple)):
        if len(spec) != 4:
            raise SceneLoaderError("Quaternion lists must be [x, y, z, w].")
        return tuple(number(v) for v in spec)
    if isinstance(spec, dict) and "axis_angle" in spec:
        aa = spec["axis_angle"]
        return tuple(float(v) for v in quat_from_axis_angle(_vec3(aa["axis"]), number(aa["angle"])))
    raise SceneLoaderError(f"Unsupported quaternion spec: {spec}")


def _transform(spec):
    """``{p, q}`` / ``identity`` / None -> 7-vector (scene.py:5697-5728)."""
    if spec is None or spec == "identity":
        return np.array([0, 0, 0, 0, 0, 0, 1], dtype=np.float64)
    if not isinstance(spec, dict):
        raise SceneLoaderError(f"Unsupported transform spec: {spec!r}")
    return np.array(_vec3(spec.get("p", (0.0, 0.0, 0.0))) + _quat(spec.get("q", "identity")), dtype=np.float64)


# ---------------------------------------------------------- shape configs
_SHAPE_ALIASES = {"contact_margin": "margin", "shape_margin": "margin", "margin": "margin", "shape_gap": "gap",
                  "gap": "gap", "mu": "mu", "shape_mu": "mu", "ke": "ke", "shape_ke": "ke", "tau": "tau",
                  "shape_tau": "tau", "relaxation_time": "tau"}
_SHAPE_OVERRIDE_KEYS = ("margin", "gap", "mu", "ke", "tau")
# accepted and stored nowhere: the SAP contact path does not read them
_SHAPE_IGNORED = {"kd", "kf", "ka", "shape_kd", "shape_kf", "shape_ka", "contact_ka", "shape_material_ka",
                  "restitution", "mu_torsional", "mu_rolling", "is_solid", "is_hydroelastic", "kh",
                  "sdf_narrow_band_range", "sdf_target_voxel_size", "sdf_max_resolution", "sdf_texture_format"}
_SHAPE_FIELDS = {f.name for f in fields(ShapeCfg)}


def _normalize_shape_cfg(cfg) -> dict:
    if not isinstance(cfg, dict):
        return {}
    return {_SHAPE_ALIASES.get(str(k), str(k)): v for k, v in cfg.items() if v is not None}


def _object_overrides(cfg) -> dict:
    """Body- or shape-level material shorthands (``physics`` / ``shape`` blocks, direct keys)."""
    if not isinstance(cfg, dict):
        return {}
    out = {}
    for src in ("physics", "shape"):
        if isinstance(cfg.get(src), dict):
            out.update({k: v for k, v in _normalize_shape_cfg(cfg[src]).items() if k in _SHAPE_OVERRIDE_KEYS})
    out.update({k: v for k, v in _normalize_shape_cfg(cfg).items() if k in _SHAPE_OVERRIDE_KEYS})
    return out


def _apply_shape_cfg(cfg: ShapeCfg, values: dict) -> ShapeCfg:
    """``_SapShapeConfig.apply`` (scene.py:127-168) onto this build's ``ShapeCfg``."""
    for name, value in values.items():
        value = _coerce(value)
        if name == "is_site":
            if bool(value):  # a site never collides and carries no mass (scene.py:150-155)
                cfg.has_shape_collision = cfg.has_particle_collision = False
                cfg.density, cfg.collision_group = 0.0, 0
            continue
        if name == "flags":
            f = int(value)
            cfg.is_visible = bool(f & SHAPE_FLAG_VISIBLE)
            cfg.has_shape_collision = bool(f & SHAPE_FLAG_COLLIDE_SHAPES)
            cfg.has_particle_collision = bool(f & SHAPE_FLAG_COLLIDE_PARTICLES)
        elif name in _SHAPE_FIELDS:
            if name == "collision_group":
                value = int(value)
            elif name in ("collision_filter_parent", "has_shape_collision", "has_particle_collision", "is_visible"):
                value = bool(value)
            setattr(cfg, name, value)
    return cfg


# ----------------------------------------------------------------- loader
class _Build:
    def __init__(self, config: dict, strict: bool):
        self.config = config
        self.strict = strict
        self.unsupported: list = []
        bcfg = config.get("builder", {}) or {}
        self.b = SceneBuilder(rigid_gap=number(bcfg.get("rigid_gap", 0.1)))
        if "gravity" in bcfg:
            self.b.gravity = number(bcfg["gravity"])
        if "default_body_armature" in bcfg:
            self.b.default_body_armature = number(bcfg["default_body_armature"])
        defaults = bcfg.get("defaults", {}) or {}
        self.joint_defaults = DofCfg()
        for k, v in (defaults.get("joint", {}) or {}).items():
            if v is None or not hasattr(self.joint_defaults, str(k)):
                continue
            setattr(self.joint_defaults, str(k), _axis(v) if k == "axis" else
                    (int(v) if k == "actuator_mode" else _coerce(v)))
        self.shape_defaults = _apply_shape_cfg(ShapeCfg(), _normalize_shape_cfg(defaults.get("shape", {}) or {}))
        self.body_ids: dict = {}
        self.joint_ids: dict = {}
        self.joint_labels: list = []
        self.b.world_count = 1
        self.b.current_world = 0

    # ------------------------------------------------------------ shapes
    def shape_cfg(self, raw) -> ShapeCfg:
        return _apply_shape_cfg(ShapeCfg(**{f: getattr(self.shape_defaults, f) for f in _SHAPE_FIELDS}),
                                _normalize_shape_cfg(raw))

    def add_shape(self, body: int, cfg: dict, body_defaults=None) -> int:
        kind = str(cfg.get("type", "box")).lower()
        eff = dict(body_defaults or {})
        eff.update(_object_overrides(cfg))
        if isinstance(cfg.get("cfg"), dict):
            eff.update(_normalize_shape_cfg(cfg["cfg"]))
        for k in list(eff):
            if k in _SHAPE_IGNORED:
                eff.pop(k)
        sc = self.shape_cfg(eff)
        x = _transform(cfg.get("transform"))
        if kind == "box":
            scale = (number(cfg.get("hx", 0.5)), number(cfg.get("hy", 0.5)), number(cfg.get("hz", 0.5)))
            return self.b.add_shape(body, GEO_BOX, x, scale, sc)
        if kind == "sphere":
            return self.b.add_shape(body, GEO_SPHERE, x, (number(cfg.get("radius", 1.0)), 0.0, 0.0), sc)
        if kind in ("capsule", "cylinder", "cone"):
            scale = (number(cfg.get("radius", 1.0)), number(cfg.get("half_height", 0.5)), 0.0)
            typ = {"capsule": GEO_CAPSULE, "cylinder": GEO_CYLINDER, "cone": GEO_CONE}[kind]
            return self.b.add_shape(body, typ, x, scale, sc)
        if kind == "ellipsoid":
            scale = (number(cfg.get("a", 1.0)), number(cfg.get("b", 0.75)), number(cfg.get("c", 0.5)))
            return self.b.add_shape(body, GEO_ELLIPSOID, x, scale, sc)
        raise SceneLoaderError(f"Unsupported inline shape type: {kind}")

    def body_ref(self, value) -> int:
        if value in (-1, "world", None):
            return -1
        if isinstance(value, int) and not isinstance(value, bool):
            return int(value)
        if value not in self.body_ids:
            raise SceneLoaderError(f"Unknown body reference: {value}")
        return self.body_ids[value]

    # ------------------------------------------------------------ joints
    def dof(self, cfg: dict) -> DofCfg:
        d = DofCfg(**{f.name: getattr(self.joint_defaults, f.name) for f in fields(DofCfg)})
        d.axis = _axis(cfg.get("axis", d.axis))
        for name in ("target_pos", "target_vel", "target_ke", "target_kd", "limit_lower", "limit_upper",
                     "limit_ke", "limit_kd", "armature", "effort_limit"):
            if name in cfg and cfg[name] is not None:
                setattr(d, name, number(cfg[name]))
        if cfg.get("actuator_mode") is not None:
            d.actuator_mode = int(cfg["actuator_mode"])
        return d

    def add_joint(self, cfg: dict, label: str) -> int:
        kind = str(cfg.get("type", "")).lower()
        common = dict(parent_xform=_transform(cfg.get("parent_xform")), child_xform=_transform(cfg.get("child_xform")),
                      collision_filter_parent=bool(cfg.get("collision_filter_parent", True)))
        if "child" not in cfg:
            raise SceneLoaderError(f"joint {label!r} has no child")
        parent, child = self.body_ref(cfg.get("parent", -1)), self.body_ref(cfg["child"])
        if kind == "fixed":
            j = self.b.add_joint(parent, child, SAP_JOINT_FIXED, **common)
        elif kind == "free":
            axes = lambda: [DofCfg(axis=a, limit_ke=0.0, limit_kd=0.0) for a in ((1, 0, 0), (0, 1, 0), (0, 0, 1))]
            j = self.b.add_joint(parent, child, SAP_JOINT_FREE, linear_axes=axes(), angular_axes=axes(), **common)
        elif kind == "revolute":
            j = self.b.add_joint(parent, child, SAP_JOINT_REVOLUTE, angular_axes=[self.dof(cfg)], **common)
        elif kind == "prismatic":
            j = self.b.add_joint(parent, child, SAP_JOINT_PRISMATIC, linear_axes=[self.dof(cfg)], **common)
        else:
            raise SceneLoaderError(f"joint type={kind or '<missing>'}")
        self.joint_labels.append(str(cfg.get("label", label)))
        return j

    def joint_index(self, value) -> int:
        """Joint by id, by label, or by label suffix after '/' (scene.py:5305-5320)."""
        if isinstance(value, int) and not isinstance(value, bool):
            return int(value)
        if value in self.joint_ids:
            return self.joint_ids[value]
        text = str(value)
        exact = [i for i, lab in enumerate(self.joint_labels) if lab == text]
        if len(exact) == 1:
            return exact[0]
        suffix = [i for i, lab in enumerate(self.joint_labels) if lab.endswith(f"/{text}")]
        if len(suffix) == 1:
            return suffix[0]
        if exact or suffix:
            raise SceneLoaderError(f"Ambiguous joint reference {value!r}")
        raise SceneLoaderError(f"Unknown joint reference: {value!r}")

    def joint_indices(self, spec: dict) -> list:
        if "joint" in spec or "joint_label" in spec:
            return [self.joint_index(spec.get("joint", spec.get("joint_label")))]
        if "joints" in spec or "joint_labels" in spec:
            return [self.joint_index(v) for v in (spec.get("joints", spec.get("joint_labels")) or ())]
        if "label_prefix" in spec:
            return [i for i, lab in enumerate(self.joint_labels) if lab.startswith(str(spec["label_prefix"]))]
        if "match" in spec:
            return [i for i, lab in enumerate(self.joint_labels) if str(spec["match"]) in lab]
        raise SceneLoaderError(f"Expected a joint selector in op: {spec!r}")

    def q_range(self, j):
        b = self.b
        s = b.joint_q_start[j]
        e = b.joint_q_start[j + 1] if j + 1 < len(b.joint_q_start) else len(b.joint_q)
        return s, e

    def qd_range(self, j):
        b = self.b
        s = b.joint_qd_start[j]
        e = b.joint_qd_start[j + 1] if j + 1 < len(b.joint_qd_start) else len(b.joint_qd)
        return s, e

    def set_joint_q(self, entry: dict):
        if "index" in entry:
            index = int(entry["index"])
        else:
            j = self.joint_index(entry.get("joint", entry.get("joint_label")))
            index = self.b.joint_q_start[j] + int(entry.get("offset", 0))
        self.b.joint_q[index] = number(entry["value"])

    # ------------------------------------------------------------ arrays
    def array(self, name: str):
        """(length, get, set) of a post-build array (``_POST_BUILD_ARRAYS``, scene.py:5445-5466)."""
        b = self.b
        shape_fields = {"shape_margin": "margin", "shape_gap": "gap", "shape_material_ke": "ke",
                        "shape_material_tau": "tau", "shape_material_mu": "mu",
                        "shape_collision_group": "collision_group"}
        dof_fields = {"joint_target_pos": "target_pos", "joint_target_vel": "target_vel",
                      "joint_target_ke": "target_ke", "joint_target_kd": "target_kd",
                      "joint_target_mode": "actuator_mode", "joint_armature": "armature"}
        if name in shape_fields:
            f = shape_fields[name]
            conv = int if f == "collision_group" else float
            return (len(b.shape_cfg), lambda i: getattr(b.shape_cfg[i], f),
                    lambda i, v: setattr(b.shape_cfg[i], f, conv(v)))
        if name == "shape_flags":
            return (len(b.shape_cfg), lambda i: b.shape_cfg[i].flags(),
                    lambda i, v: _apply_shape_cfg(b.shape_cfg[i], {"flags": int(v)}))
        if name in ("shape_material_restitution", "shape_material_mu_torsional", "shape_material_mu_rolling",
                    "shape_material_kh"):
            return (len(b.shape_cfg), lambda i: 0.0, lambda i, v: None)  # not read by the SAP contact path
        if name in dof_fields:
            f = dof_fields[name]
            conv = int if f == "actuator_mode" else float
            return (len(b.dofs), lambda i: getattr(b.dofs[i], f), lambda i, v: setattr(b.dofs[i], f, conv(v)))
        if name in ("joint_q", "joint_qd"):
            arr = getattr(b, name)
            return (len(arr), lambda i: arr[i], lambda i, v: arr.__setitem__(i, float(v)))
        return None

    @staticmethod
    def indices(n: int, spec: dict) -> list:
        """``_indices_for`` (scene.py:5552-5576): index, or range all / head:k / tail:k / from:i / [a, b]."""
        norm = lambda i: n + i if i < 0 else i
        if "index" in spec:
            return [norm(int(spec["index"]))]
        r = spec.get("range", "all")
        if r == "all":
            return list(range(n))
        if isinstance(r, str):
            if r.startswith("head:"):
                return list(range(min(int(r.split(":", 1)[1]), n)))
            if r.startswith("tail:"):
                k = min(int(r.split(":", 1)[1]), n)
                return list(range(n - k, n))
            if r.startswith("from:"):
                return list(range(norm(int(r.split(":", 1)[1])), n))
        if isinstance(r, (list, tuple)) and len(r) == 2:
            s = norm(int(r[0]))
            e = norm(int(r[1])) if r[1] is not None else n
            return list(range(s, min(e, n)))
        raise SceneLoaderError(f"Unsupported range spec: {r!r}")

    def post_build(self, op: dict):
        """``_apply_post_build_op`` (scene.py:5380-5442)."""
        kind = str(op.get("op", "")).lower()
        b = self.b
        if kind == "add_shape":
            self.add_shape(self.body_ref(op.get("body", -1)), op)
        elif kind == "set_array":
            arr = self.array(str(op.get("array", "")))
            if arr is None:
                self.unsupported.append(f"post_build set_array {op.get('array') or '<missing>'}")
                return
            n, get, put = arr
            idx = self.indices(n, op)
            if "from_array" in op:
                src = self.array(str(op["from_array"]))
                if src is None:
                    self.unsupported.append(f"post_build set_array from_array {op['from_array']}")
                    return
                sidx = self.indices(src[0], op.get("from_range", op))
                if len(sidx) != len(idx):
                    raise SceneLoaderError(f"set_array from_array length mismatch: {len(sidx)} != {len(idx)}")
                values = [src[1](i) for i in sidx]
            elif "values" in op:
                values = [_coerce(v) for v in op["values"]]
                if len(values) != len(idx):
                    raise SceneLoaderError(f"set_array expected {len(idx)} values, got {len(values)}")
            else:
                values = [_coerce(op["value"])] * len(idx)
            for i, v in zip(idx, values):
                put(i, v)
        elif kind == "copy_array":
            src, dst = self.array(str(op.get("src", ""))), self.array(str(op.get("dst", "")))
            if src is None or dst is None:
                self.unsupported.append(f"post_build copy_array {op.get('src')}->{op.get('dst')}")
                return
            si, di = self.indices(src[0], op.get("src_range", op)), self.indices(dst[0], op.get("dst_range", op))
            if len(si) != len(di):
                raise SceneLoaderError(f"copy_array length mismatch: {len(si)} != {len(di)}")
            for a, c in zip(si, di):
                dst[2](c, src[1](a))
        elif kind == "scale_sphere_shapes":
            f = number(op.get("factor", 1.0))
            for i, t in enumerate(b.shape_type):
                if t == GEO_SPHERE:
                    b.shape_scale[i] = (float(b.shape_scale[i][0]) * f, 0.0, 0.0)
        elif kind == "set_attr":
            name = str(op.get("name", ""))
            if name in ("gravity", "rigid_gap", "default_body_armature"):
                setattr(b, name, number(op["value"]))
            else:
                self.unsupported.append(f"post_build set_attr {name or '<missing>'}")
        elif kind == "set_joint_targets":
            if any(k in op for k in ("joint", "joints", "joint_label", "joint_labels", "match", "label_prefix")):
                idx = [d for j in self.joint_indices(op) for d in range(*self.qd_range(j))]
            else:
                idx = self.indices(len(b.dofs), op)
            mode = op.get("mode", "position")
            modes = {"none": 0, "position": 1, "velocity": 2, "position_velocity": 3, "effort": 4}
            if not isinstance(mode, int):
                if str(mode).lower() not in modes:
                    raise SceneLoaderError(f"Unknown joint target mode: {mode!r}")
                mode = modes[str(mode).lower()]
            ke, kd = number(op.get("ke", 0.0)), number(op.get("kd", 0.0))
            for d in idx:
                b.dofs[d].target_ke, b.dofs[d].target_kd, b.dofs[d].actuator_mode = ke, kd, int(mode)
        elif kind == "set_joint_q":
            self.set_joint_q(op)
        elif kind == "copy_joint_q_to_joint_targets":
            for j in self.joint_indices(op):
                qs, qe = self.q_range(j)
                ds, de = self.qd_range(j)
                for k in range(min(qe - qs, de - ds)):
                    b.dofs[ds + k].target_pos = b.joint_q[qs + k]
        elif kind == "set_joint_armature":
            value = number(op["value"])
            for d in self.indices(len(b.dofs), op):
                b.dofs[d].armature = value
        else:
            self.unsupported.append(f"post_build op={kind or '<missing>'}")


def _writes_joint_state(op) -> bool:
    kind = str(op.get("op", "")).lower()
    if kind in ("set_joint_q",):
        return True
    if kind == "set_array":
        return str(op.get("array", "")) in ("joint_q", "joint_qd")
    if kind == "copy_array":
        return str(op.get("dst", "")) in ("joint_q", "joint_qd")
    return False


def load_scene_config(path) -> dict:
    """Read a YAML or JSON scene file and check its schema version (scene.py:1372-1384)."""
    import yaml

    with open(path, "r", encoding="utf-8") as f:
        data = json.load(f) if str(path).lower().endswith(".json") else yaml.safe_load(f)
    if not isinstance(data, dict):
        raise SceneLoaderError(f"Scene file {path} must contain a mapping.")
    version = int(data.get("schema_version", SCHEMA_VERSION))
    if version != SCHEMA_VERSION:
        raise SceneLoaderError(f"Unsupported scene schema_version={version}; expected {SCHEMA_VERSION}.")
    return data


def build_scene(config: dict, *, num_worlds=None, spacing=None, strict: bool = True, device: str = "cpu"):
    """Scene mapping -> ``LoadedScene`` (``_build_sap_scene_builder`` + model build, scene.py:1387-1509).

    Order as in the reference: bodies and their shapes, joints, articulations,
    ``initial_joint_q``, ``post_build`` on the single template world, then
    replication over ``num_worlds`` worlds, then the ground plane.  Body poses
    follow from the joint coordinates by forward kinematics when the scene sets
    joint coordinates (``initial_joint_q`` or a post-build write to
    ``joint_q`` / ``joint_qd``), as the reference evaluates its initial FK only
    then (scene.py:1525-1528).
    """
    if not isinstance(config, dict):
        raise SceneLoaderError("scene config must be a mapping")
    sim = config.get("simulation", {}) or {}
    rep = config.get("replicate", {}) or {}
    W = int(num_worlds if num_worlds is not None else sim.get("num_worlds", rep.get("num_worlds", 1)))
    sp = spacing if spacing is not None else _vec3(rep.get("spacing", DEFAULT_SPACING))
    st = _Build(config, strict)
    b = st.b
    for asset in config.get("assets", []) or []:
        st.unsupported.append(f"asset[{asset.get('id', '?')!r}] {asset.get('type', '?')}: file assets are not loaded")
    for body_cfg in config.get("bodies", []) or []:
        bid = body_cfg.get("id")
        if not isinstance(bid, str) or not bid:
            raise SceneLoaderError("Expected non-empty string field 'id'.")
        body = b.add_body(_transform(body_cfg.get("transform")), mass=number(body_cfg.get("mass", 0.0)))
        st.body_ids[bid] = body
        overrides = _object_overrides(body_cfg)
        for shape_cfg in body_cfg.get("shapes", []) or []:
            try:
                st.add_shape(body, shape_cfg, overrides)
            except SceneLoaderError as exc:
                st.unsupported.append(str(exc))
    for joint_cfg in config.get("joints", []) or []:
        jid = joint_cfg.get("id")
        if not isinstance(jid, str) or not jid:
            raise SceneLoaderError("Expected non-empty string field 'id'.")
        try:
            st.joint_ids[jid] = st.add_joint(joint_cfg, jid)
        except SceneLoaderError as exc:
            st.unsupported.append(str(exc))
    for art in config.get("articulations", []) or []:
        names = art.get("joints", []) or []
        for n in names:
            if n not in st.joint_ids:
                raise SceneLoaderError(f"Unknown joint reference: {n!r}")
        if names:
            b.add_articulation([st.joint_ids[n] for n in names])
    for entry in config.get("initial_joint_q", []) or []:
        st.set_joint_q(entry)
    for op in config.get("post_build", []) or []:
        try:
            st.post_build(op)
        except SceneLoaderError as exc:
            st.unsupported.append(f"post_build {op.get('op')}: {exc}")
    if strict and st.unsupported:
        raise UnsupportedSceneFeature(st.unsupported)
    b.replicate(W, spacing=sp)
    ground = config.get("ground", {"enabled": True})
    if ground is True or ground is None:
        ground = {"enabled": True}
    if ground is not False and bool(ground.get("enabled", True)):
        b.add_ground_plane(st.shape_cfg((ground.get("cfg") or {})), height=number(ground.get("height", 0.0)))
    model = b.finalize(device)
    needs_fk = bool(config.get("initial_joint_q")) or any(_writes_joint_state(op) for op in config.get("post_build", []) or [])
    if needs_fk:
        model = _with_consistent_body_q(model)
    solver = dict(sim.get("solver", {}) or {})
    return LoadedScene(model=model, name=str(config.get("name", "")), dt=number(sim.get("dt", 1.0e-3)),
                       max_rigid_contact=int(sim.get("max_rigid_contact", 0) or 0), solver=solver,
                       body_ids=dict(st.body_ids), joint_ids=dict(st.joint_ids), unsupported=list(st.unsupported))


def load_scene(path, *, num_worlds=None, spacing=None, strict: bool = True, device: str = "cpu") -> LoadedScene:
    """``load_sap_scene`` for inline scene files: the model plus dt, contact capacity and solver settings."""
    return build_scene(load_scene_config(os.fspath(path)), num_worlds=num_worlds, spacing=spacing, strict=strict,
                       device=device)


__all__ = ["LoadedScene", "SceneLoaderError", "UnsupportedSceneFeature", "build_scene", "load_scene",
           "load_scene_config", "number"]

"""Python side of the float32 collision oracle (TEST INFRASTRUCTURE ONLY).

Loads ``oracle/_build/libsapx_oracle_np.so`` (built from ``oracle/narrowphase.c``
by ``oracle/Makefile``) and runs the reference's primitive contact generation
on the CPU.  See the C file header for the reference lines it restates.
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsapx_oracle_np.so")
_lib = None


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.sapx_oracle_collide.restype = ctypes.c_int
    return _lib


def pair_worlds(model) -> np.ndarray:
    """World of each explicit pair: the world of its non-global shape."""
    pairs = np.asarray(model.shape_pairs, dtype=np.int64).reshape(-1, 2)
    sw = np.asarray(model.shape_world)
    w0 = sw[pairs[:, 0]]
    w1 = sw[pairs[:, 1]]
    return np.where(w0 >= 0, w0, w1).astype(np.int32)


def world_major_pairs(model):
    """Canonical candidate order: stable sort of the explicit pair list by world."""
    pairs = np.asarray(model.shape_pairs, dtype=np.int32).reshape(-1, 2)
    pw = pair_worlds(model)
    order = np.argsort(pw, kind="stable")
    return np.ascontiguousarray(pairs[order]), np.ascontiguousarray(pw[order])


def _ptr(a, ct):
    return a.ctypes.data_as(ctypes.POINTER(ct))


def collide(model, body_q, max_contacts_per_world: int):
    """Returns a dict of compact contact arrays in world-major / pair order."""
    lib = _load()
    pairs, pw = world_major_pairs(model)
    W = model.world_count
    S = model.shape_count
    cap = max(1, W * max_contacts_per_world)
    f32 = np.float32
    i32 = np.int32
    bq = np.ascontiguousarray(np.asarray(body_q, dtype=f32).reshape(-1, 7))
    out = dict(
        count=np.zeros(1, i32), world_start=np.zeros(W + 1, i32), shape0=np.full(cap, -1, i32),
        shape1=np.full(cap, -1, i32), point0=np.zeros((cap, 3), f32), point1=np.zeros((cap, 3), f32),
        normal=np.zeros((cap, 3), f32), margin0=np.zeros(cap, f32), margin1=np.zeros(cap, f32),
        pair=np.full(cap, -1, i32),
    )
    scratch_tf = np.zeros((S, 7), f32)
    scratch_aabb = np.zeros((S, 6), f32)
    args = [
        ctypes.c_int(W), ctypes.c_int(S), _ptr(bq, ctypes.c_float),
        _ptr(np.ascontiguousarray(model.shape_transform, f32), ctypes.c_float),
        _ptr(np.ascontiguousarray(model.shape_body, i32), ctypes.c_int),
        _ptr(np.ascontiguousarray(model.shape_type, i32), ctypes.c_int),
        _ptr(np.ascontiguousarray(model.shape_scale, f32), ctypes.c_float),
        _ptr(np.ascontiguousarray(model.shape_margin, f32), ctypes.c_float),
        _ptr(np.ascontiguousarray(model.shape_gap, f32), ctypes.c_float),
        _ptr(np.ascontiguousarray(model.shape_collision_radius, f32), ctypes.c_float),
        _ptr(pairs, ctypes.c_int), _ptr(pw, ctypes.c_int), ctypes.c_int(len(pw)),
        ctypes.c_int(max_contacts_per_world), ctypes.c_int(cap),
        _ptr(out["count"], ctypes.c_int), _ptr(out["world_start"], ctypes.c_int),
        _ptr(out["shape0"], ctypes.c_int), _ptr(out["shape1"], ctypes.c_int),
        _ptr(out["point0"], ctypes.c_float), _ptr(out["point1"], ctypes.c_float),
        _ptr(out["normal"], ctypes.c_float), _ptr(out["margin0"], ctypes.c_float),
        _ptr(out["margin1"], ctypes.c_float), _ptr(out["pair"], ctypes.c_int),
        scratch_tf.ctypes.data_as(ctypes.c_void_p), _ptr(scratch_aabb, ctypes.c_float),
    ]
    unsupported = lib.sapx_oracle_collide(*args)
    out["unsupported"] = int(unsupported)
    out["count"] = int(out["count"][0])
    return out


def convex_pair(type_a, scale_a, xf_a, type_b, scale_b, xf_b, gap_sum=0.0, margin_a=0.0, margin_b=0.0):
    """Contacts of one convex pair (box-box / capsule-box), float32.

    ``xf`` = (px, py, pz, qx, qy, qz, qw).  Returns a dict with ``pos`` (n x 3
    contact centres), ``dist`` (n signed distances), ``normal`` (world, A to B)
    and ``radii`` (effective radii), as the per-contact writer receives them.
    """
    lib = _load()
    f32 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    sa, xa, sb, xb = f32(scale_a), f32(xf_a), f32(scale_b), f32(xf_b)
    pos = np.zeros((5, 3), np.float32)
    dist = np.zeros(5, np.float32)
    normal = np.zeros(3, np.float32)
    radii = np.zeros(2, np.float32)
    fp = ctypes.POINTER(ctypes.c_float)
    n = lib.sapx_oracle_convex_pair(
        ctypes.c_int(type_a), _ptr(sa, ctypes.c_float), _ptr(xa, ctypes.c_float), ctypes.c_int(type_b),
        _ptr(sb, ctypes.c_float), _ptr(xb, ctypes.c_float), ctypes.c_float(gap_sum), ctypes.c_float(margin_a),
        ctypes.c_float(margin_b), _ptr(pos, ctypes.c_float), _ptr(dist, ctypes.c_float),
        _ptr(normal, ctypes.c_float), _ptr(radii, ctypes.c_float))
    del fp
    return {"count": int(n), "pos": pos[:n].copy(), "dist": dist[:n].copy(), "normal": normal, "radii": radii}

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


def _group_pair_ok(ga: int, gb: int) -> bool:
    """Collision-group rule (sim/collision/broad_phase_common.py:130-150)."""
    if ga == 0 or gb == 0:
        return False
    if ga > 0:
        return ga == gb or gb < 0
    return ga != gb


def all_pairs(model):
    """Candidate pairs of the all-pairs broadphase, world by world, in lower-triangular order.

    Restates SapBroadPhaseAllPairs (sim/collision/broad_phase_nxn.py:60-207,
    215-325) and its world map (broad_phase_common.py:183-300): a world's slice
    is its shapes with the COLLIDE_SHAPES flag followed by the colliding global
    shapes; pair (r, c > r) of the slice is canonicalised to (min, max) and
    kept when not both global (the dedicated global segment is not collided
    here), the worlds match or one is global, the collision groups may
    collide (broad_phase_common.py:153-180) and the pair is not excluded
    (pipeline.py:533-544).  The AABB test is left to the narrowphase driver.
    Returns (pairs, pair_world) like ``world_major_pairs``.
    """
    sw = np.asarray(model.shape_world, dtype=np.int64).reshape(-1)
    grp = np.asarray(model.shape_collision_group, dtype=np.int64).reshape(-1)
    flags = np.asarray(model.shape_flags, dtype=np.int64).reshape(-1)
    fp = model.shape_collision_filter_pairs
    excluded = set() if fp is None else {(min(a, b), max(a, b)) for a, b in np.asarray(fp).reshape(-1, 2).tolist()}
    colliding = (flags & (1 << 1)) != 0
    shared = [s for s in range(sw.shape[0]) if sw[s] == -1 and colliding[s]]
    pairs, pw = [], []
    for w in range(int(model.world_count)):
        sl = [s for s in np.nonzero((sw == w) & colliding)[0].tolist()] + shared
        for r in range(len(sl)):
            for c in range(r + 1, len(sl)):
                a, b = min(sl[r], sl[c]), max(sl[r], sl[c])
                if sw[a] == -1 and sw[b] == -1:
                    continue
                if sw[a] != -1 and sw[b] != -1 and sw[a] != sw[b]:
                    continue
                if not _group_pair_ok(int(grp[a]), int(grp[b])) or (a, b) in excluded:
                    continue
                pairs.append((a, b))
                pw.append(w)
    return (np.ascontiguousarray(np.asarray(pairs, dtype=np.int32).reshape(-1, 2)),
            np.ascontiguousarray(np.asarray(pw, dtype=np.int32)))


def collide(model, body_q, max_contacts_per_world: int, pairs=None, threads: int = 1):
    """Returns a dict of compact contact arrays in world-major / pair order.

    ``pairs``: ``(pairs, pair_world)`` grouped by world (``all_pairs``); default:
    the model's explicit pairs in world-major order.  ``threads`` > 1 runs the
    worlds on that many OpenMP threads (same output; the CPU baseline's collide).
    """
    lib = _load()
    pairs, pw = world_major_pairs(model) if pairs is None else pairs
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
    if threads > 1:
        C = max_contacts_per_world
        wps = np.ascontiguousarray(np.searchsorted(pw, np.arange(W + 1)).astype(i32))
        stage_i = np.zeros(W * C * 3, i32)
        stage_f = np.zeros(W * C * 11, f32)
        wcount = np.zeros(W, i32)
        args += [_ptr(wps, ctypes.c_int), _ptr(stage_i, ctypes.c_int), _ptr(stage_f, ctypes.c_float),
                 _ptr(wcount, ctypes.c_int), ctypes.c_int(int(threads))]
        unsupported = lib.sapx_oracle_collide_mt(*args)
    else:
        unsupported = lib.sapx_oracle_collide(*args)
    out["unsupported"] = int(unsupported)
    out["count"] = int(out["count"][0])
    return out


def support(shape_type, scale, direction):
    """Support point of one shape in its own frame, float32 (the MPR / GJK support maps)."""
    lib = _load()
    f32 = lambda x: np.ascontiguousarray(np.asarray(x, dtype=np.float32))  # noqa: E731
    sc, d = f32(scale), f32(direction)
    out = np.zeros(3, np.float32)
    lib.sapx_oracle_support(ctypes.c_int(shape_type), _ptr(sc, ctypes.c_float), _ptr(d, ctypes.c_float),
                            _ptr(out, ctypes.c_float))
    return out


def convex_pair(type_a, scale_a, xf_a, type_b, scale_b, xf_b, gap_sum=0.0, margin_a=0.0, margin_b=0.0):
    """Contacts of one pair on the MPR / GJK path (sorted types, type_a <= type_b), float32.

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

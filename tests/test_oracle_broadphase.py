"""Pins of the oracle's all-pairs broadphase enumeration (``oracle.collide.all_pairs``).

* The kept set equals the scene builder's explicit candidate list, an
  independent restatement of ``_find_shape_contact_pairs`` (scene.py:1188-1212):
  both apply the world, collision-group and filter-pair rules, one per step over
  all pairs, the other once at build time.
* The order inside a world is the reference kernel's thread order: thread t of a
  world handles pair ``_sap_get_lower_triangular_indices(t, n)``
  (broad_phase_nxn.py:60-78), restated below from its binary search; the
  oracle's pairs must come out with strictly increasing t.
* Group semantics on a hand-built scene: positive groups collide only with
  themselves and negative groups, negative groups with everything but
  themselves, group 0 with nothing.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle.collide import all_pairs, world_major_pairs
from sap_warp_amd.runtime import GEO_BOX, GEO_SPHERE
from sap_warp_amd.scenes import (SceneBuilder, ShapeCfg, build_ant_scene, build_box_pile_scene,
                                 build_humanoid_scene, tf)


def _lower_tri(index, n):
    """Thread index -> (row, col) of the strict upper triangle, by the reference's binary search."""
    total = (n * (n - 1)) >> 1
    if index >= total:
        return -1, -1
    low, high = 0, n - 1
    while low < high:
        mid = (low + high) >> 1
        if (mid * (2 * n - mid - 1)) >> 1 <= index:
            low = mid + 1
        else:
            high = mid
    r = low - 1
    f = (r * (2 * n - r - 1)) >> 1
    return r, (index - f) + r + 1


@pytest.mark.parametrize("n", [2, 3, 7, 32, 41])
def test_lower_tri_map_is_row_major(n):
    seen = [_lower_tri(t, n) for t in range(n * (n - 1) // 2)]
    expected = [(r, c) for r in range(n) for c in range(r + 1, n)]
    assert seen == expected


@pytest.mark.parametrize("build", [lambda: build_humanoid_scene(3, seed=1), lambda: build_ant_scene(3),
                                   lambda: build_box_pile_scene(3)])
def test_all_pairs_set_equals_explicit(build):
    m = build()
    a, aw = all_pairs(m)
    e, ew = world_major_pairs(m)
    assert sorted(map(tuple, a.tolist())) == sorted(map(tuple, e.tolist()))
    assert np.all(np.diff(aw) >= 0)


def test_all_pairs_thread_order():
    m = build_humanoid_scene(2, seed=4)
    a, aw = all_pairs(m)
    sw = np.asarray(m.shape_world)
    flags = np.asarray(m.shape_flags)
    shared = [s for s in range(len(sw)) if sw[s] == -1 and flags[s] & 2]
    for w in range(m.world_count):
        sl = [s for s in range(len(sw)) if sw[s] == w and flags[s] & 2] + shared
        pos = {s: i for i, s in enumerate(sl)}
        n = len(sl)
        tids = []
        for x, y in a[aw == w].tolist():
            r, c = sorted((pos[x], pos[y]))
            tids.append(next(t for t in range(n * (n - 1) // 2) if _lower_tri(t, n) == (r, c)))
        assert tids == sorted(tids) and len(set(tids)) == len(tids)


def test_all_pairs_group_rules():
    b = SceneBuilder()
    b.world_count = 1
    b.current_world = 0
    groups = [1, 1, 2, -1, -2, 0]
    for g in groups:
        body = b.add_body(tf())
        b.add_shape(body, GEO_SPHERE, tf(), (0.1, 0.0, 0.0), ShapeCfg(collision_group=g))
    b.current_world = -1
    b.add_shape(-1, GEO_BOX, tf(), (1.0, 1.0, 1.0), ShapeCfg(density=0.0, collision_group=-1))
    m = b.finalize()
    a, _ = all_pairs(m)
    got = set(map(tuple, a.tolist()))
    allg = groups + [-1]
    want = set()
    for i in range(len(allg)):
        for j in range(i + 1, len(allg)):
            gi, gj = allg[i], allg[j]
            ok = gi != 0 and gj != 0 and ((gi > 0 and (gi == gj or gj < 0)) or (gi < 0 and gi != gj))
            if ok:
                want.add((i, j))
    assert got == want
    assert (0, 1) in got and (0, 2) not in got and (3, 4) in got and (3, 6) not in got

"""Pins of the oracle's convex-pair path (box-box, capsule-box: MPR, GJK, manifold).

Every expected value here comes from geometry, not from the oracle's formulas:
face-contact corners and depths of axis-aligned boxes, the vertices of the exact
intersection of two overlapping face rectangles (corner containment + edge
crossings, not polygon clipping), distances between separated boxes / capsules
from a bounded least-squares solve over the shapes' parameterisations, the
lowest point of a tilted box, and the a/b swap symmetry.  The reference
behaviours kept by the oracle and checked here: the MPR witness points sit on
shapes inflated by 1e-4 (so manifold depths read 1e-4 deeper than the true
overlap while the single deepest contact of a separated pair is exact), the
capsule radius is carried as an effective radius, an edge resting on a face
gives its two end points, and a corner gives only the deepest point.
"""
from __future__ import annotations

import itertools
import math

import numpy as np
import pytest
from scipy.optimize import lsq_linear, minimize

from oracle.collide import convex_pair

BOX, CAP = 7, 4
ID = (0.0, 0.0, 0.0, 1.0)
INFLATE = 1e-4


def qaxis(axis, ang):
    a = np.asarray(axis, float)
    a = a / np.linalg.norm(a)
    s = math.sin(0.5 * ang)
    return (a[0] * s, a[1] * s, a[2] * s, math.cos(0.5 * ang))


def qmat(q):
    x, y, z, w = q
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def xf(p, q=ID):
    return tuple(p) + tuple(q)


def box_box(ha, pa, qa, hb, pb, qb, gap=0.02):
    return convex_pair(BOX, ha, xf(pa, qa), BOX, hb, xf(pb, qb), gap)


def sorted_rows(a):
    a = np.asarray(a, float)
    return a[np.lexsort(a.T[::-1])]


# ------------------------------------------------------------------ face contacts
@pytest.mark.parametrize("pen", [0.01, 0.002])
def test_box_resting_inside_face(pen):
    """A small box resting on a larger box: four contacts at the small box's bottom corners."""
    c = (0.1, 0.2)
    r = box_box((1, 1, 0.5), (0, 0, 0), ID, (0.5, 0.5, 0.5), (c[0], c[1], 1.0 - pen), ID)
    assert r["count"] == 4
    np.testing.assert_allclose(r["normal"], [0, 0, 1], atol=1e-5)
    corners = [(c[0] + sx * 0.5, c[1] + sy * 0.5) for sx in (-1, 1) for sy in (-1, 1)]
    np.testing.assert_allclose(sorted_rows(r["pos"][:, :2]), sorted_rows(corners), atol=2e-5)
    # contact centre midway through the overlap; depth 1e-4 deeper than the true overlap
    np.testing.assert_allclose(r["pos"][:, 2], 0.5 - 0.5 * pen, atol=2e-6)
    np.testing.assert_allclose(r["dist"], -(pen + INFLATE), rtol=1e-4, atol=2e-7)


def test_box_separated_within_gap():
    """A box hovering 5 mm above another (inside the contact threshold): GJK, exact distance."""
    r = box_box((1, 1, 0.5), (0, 0, 0), ID, (0.5, 0.5, 0.5), (0.1, 0.2, 1.005), ID)
    assert r["count"] == 4
    np.testing.assert_allclose(r["dist"], 0.005, atol=2e-7)
    np.testing.assert_allclose(r["normal"], [0, 0, 1], atol=1e-4)
    np.testing.assert_allclose(r["pos"][:, 2], 0.5025, atol=1e-6)


def _rect(c, h, yaw):
    R = np.array([[math.cos(yaw), -math.sin(yaw)], [math.sin(yaw), math.cos(yaw)]])
    return np.array([np.asarray(c) + R @ (sx * h[0], sy * h[1]) for sx, sy in ((-1, -1), (1, -1), (1, 1), (-1, 1))])


def _inside(p, poly, eps=1e-7):
    n = len(poly)
    cr = lambda u, v: u[0] * v[1] - u[1] * v[0]
    sgn = [cr(poly[(i + 1) % n] - poly[i], p - poly[i]) for i in range(n)]
    return all(s >= -eps for s in sgn) or all(s <= eps for s in sgn)


def _intersection_vertices(P, Q):
    """Vertices of the intersection of two convex polygons: contained corners + edge crossings."""
    pts = [p for p in P if _inside(p, Q)] + [q for q in Q if _inside(q, P)]
    for i, j in itertools.product(range(len(P)), range(len(Q))):
        a, b = P[i], P[(i + 1) % len(P)]
        c, d = Q[j], Q[(j + 1) % len(Q)]
        M = np.array([b - a, c - d]).T
        if abs(np.linalg.det(M)) < 1e-12:
            continue
        t, u = np.linalg.solve(M, c - a)
        if -1e-9 <= t <= 1 + 1e-9 and -1e-9 <= u <= 1 + 1e-9:
            pts.append(a + t * (b - a))
    return np.array(pts)


def _quad_area(pts):
    c = pts.mean(axis=0)
    order = np.argsort(np.arctan2(pts[:, 1] - c[1], pts[:, 0] - c[0]))
    x, y = pts[order, 0], pts[order, 1]
    return 0.5 * abs(np.dot(x, np.roll(y, -1)) - np.dot(y, np.roll(x, -1)))


@pytest.mark.parametrize("center,yaw,big", [((0.8, 0.1), 0.0, (1.0, 1.0, 0.5)), ((0.75, -0.7), 0.0, (1.0, 1.0, 0.5)),
                                            ((0.9, 0.2), 0.6, (1.0, 1.0, 0.5)),
                                            ((0.0, 0.0), math.pi / 4, (1.0, 1.0, 0.5)),
                                            ((0.0, 0.0), math.pi / 4, (0.5, 0.5, 0.5)),
                                            ((0.05, -0.03), 0.5, (0.55, 0.45, 0.5))])
def test_overhanging_face_contacts_are_intersection_vertices(center, yaw, big):
    """Contacts of an overhanging / rotated face pair lie on vertices of the exact face intersection.

    With more than four intersection vertices (two equal squares at 45 degrees: an octagon) four are
    kept; the rotating-calipers choice must keep the polygon's diameter and come within 10 % of the
    largest quadrilateral on its vertices (brute force over all 4-subsets).
    """
    hb = (0.5, 0.5, 0.5)
    pen = 0.004
    r = box_box(big, (0, 0, 0), ID, hb, (center[0], center[1], 1.0 - pen), qaxis((0, 0, 1), yaw))
    poly = _intersection_vertices(_rect((0, 0), big, 0.0), _rect(center, hb, yaw))
    assert 2 <= r["count"] <= 5
    xy = r["pos"][:, :2]
    for p in xy:
        d = np.min(np.linalg.norm(poly - p, axis=1))
        assert d < 5e-5, (p, poly)
    np.testing.assert_allclose(r["dist"], -(pen + INFLATE), rtol=1e-3)
    uniq = np.unique(np.round(poly, 6), axis=0)
    if len(uniq) > 4:
        assert r["count"] >= 4
        diam = max(np.linalg.norm(a - b) for a, b in itertools.combinations(uniq, 2))
        kept = max(np.linalg.norm(a - b) for a, b in itertools.combinations(xy, 2))
        assert kept == pytest.approx(diam, rel=1e-3)
        best = max(_quad_area(uniq[list(c)]) for c in itertools.combinations(range(len(uniq)), 4))
        assert _quad_area(np.unique(np.round(xy[:4], 5), axis=0)) >= 0.9 * best
    else:
        assert r["count"] == len(uniq)


def test_swap_symmetry_box_box():
    """Swapping the shapes flips the normal and keeps the contact set."""
    qa, qb = qaxis((0, 0, 1), 0.3), qaxis((0, 0, 1), -0.2)
    pa, pb = (0.02, -0.03, 0.0), (0.15, 0.1, 0.896)
    ra = box_box((1, 0.8, 0.5), pa, qa, (0.3, 0.4, 0.4), pb, qb)
    rb = box_box((0.3, 0.4, 0.4), pb, qb, (1, 0.8, 0.5), pa, qa)
    np.testing.assert_allclose(ra["normal"], -rb["normal"], atol=1e-5)
    assert ra["count"] == rb["count"]
    np.testing.assert_allclose(sorted_rows(ra["pos"]), sorted_rows(rb["pos"]), atol=5e-5)
    np.testing.assert_allclose(np.sort(ra["dist"]), np.sort(rb["dist"]), atol=1e-6)


@pytest.mark.parametrize("tilt", [0.06, 0.15, 0.4])
def test_tilted_box_edge_on_face(tilt):
    """A box rolled onto one bottom edge: two contacts at that edge's ends.

    The face's plane passes through the (inflated) MPR witness point and the edge's line through
    the edge itself, so the depth reads pen + half the inflation.
    """
    hb = np.array([0.3, 0.2, 0.15])
    q = qaxis((1, 0, 0), tilt)
    R = qmat(q)
    corners = [R @ (hb * np.array(s)) for s in itertools.product((-1, 1), repeat=3)]
    low = min(c[2] for c in corners)
    edge = [c for c in corners if abs(c[2] - low) < 1e-9]
    assert len(edge) == 2
    pen, p = 0.003, np.array([0.05, 0.0, 0.0])
    p[2] = 0.5 - pen - low
    r = box_box((1, 1, 0.5), (0, 0, 0), ID, tuple(hb), tuple(p), q)
    assert r["count"] == 2
    np.testing.assert_allclose(r["normal"], [0, 0, 1], atol=1e-4)
    np.testing.assert_allclose(r["dist"], -(pen + 0.5 * INFLATE), atol=2e-6)
    ends = np.array([p + c for c in edge])[:, :2]
    np.testing.assert_allclose(sorted_rows(r["pos"][:, :2]), sorted_rows(ends), atol=2e-5)


@pytest.mark.parametrize("tilt", [0.1, 0.3])
def test_tilted_box_corner_on_face(tilt):
    """A box tipped onto one corner: only the deepest point, at that corner, exact depth."""
    hb = np.array([0.3, 0.2, 0.15])
    q = qaxis((1, -1.3, 0), tilt)
    R = qmat(q)
    corners = sorted((R @ (hb * np.array(s)) for s in itertools.product((-1, 1), repeat=3)), key=lambda c: c[2])
    assert corners[1][2] - corners[0][2] > 1e-3
    pen, p = 0.003, np.array([-0.1, 0.05, 0.0])
    p[2] = 0.5 - pen - corners[0][2]
    r = box_box((1, 1, 0.5), (0, 0, 0), ID, tuple(hb), tuple(p), q)
    assert r["count"] == 1
    np.testing.assert_allclose(r["normal"], [0, 0, 1], atol=1e-4)
    np.testing.assert_allclose(r["dist"][0], -(pen + INFLATE), atol=2e-6)
    np.testing.assert_allclose(r["pos"][0, :2], (p + corners[0])[:2], atol=2e-5)
    assert r["pos"][0, 2] == pytest.approx(0.5 - 0.5 * pen, abs=1e-4)


# -------------------------------------------------------------- separated distance
def _box_points(h, p, q, u):
    return np.asarray(p) + qmat(q) @ (np.asarray(h) * u)


def _box_distance(ha, pa, qa, hb, pb, qb):
    """min |x_a - x_b| over both boxes: bounded least squares in the 6 box coordinates."""
    Ra, Rb = qmat(qa) * np.asarray(ha), qmat(qb) * np.asarray(hb)
    A = np.hstack([Ra, -Rb])
    res = lsq_linear(A, np.asarray(pb) - np.asarray(pa), bounds=(-1, 1), tol=1e-14, lsmr_tol="auto",
                     max_iter=2000)
    xa = _box_points(ha, pa, qa, res.x[:3])
    xb = _box_points(hb, pb, qb, res.x[3:])
    return np.linalg.norm(xb - xa), xa, xb


@pytest.mark.parametrize("seed", range(6))
def test_separated_boxes_distance(seed):
    """Random rotated boxes farther apart than the contact threshold: one contact at the exact distance."""
    rng = np.random.default_rng(seed)
    ha = rng.uniform(0.1, 0.5, 3)
    hb = rng.uniform(0.1, 0.5, 3)
    qa = qaxis(rng.normal(size=3), rng.uniform(0, math.pi))
    qb = qaxis(rng.normal(size=3), rng.uniform(0, math.pi))
    pa = rng.uniform(-0.2, 0.2, 3)
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    # push B out along d until the true distance is about 0.05 - 0.3
    for s in np.linspace(0.1, 3.0, 300):
        dist, _, _ = _box_distance(ha, pa, qa, hb, pa + s * d, qb)
        if dist > 0.05:
            break
    pb = pa + (s + rng.uniform(0.0, 0.25)) * d
    dist, xa, xb = _box_distance(ha, pa, qa, hb, pb, qb)
    r = box_box(tuple(ha), tuple(pa), qa, tuple(hb), tuple(pb), qb, gap=0.01)
    assert r["count"] == 1
    assert r["dist"][0] == pytest.approx(dist, abs=2e-5)
    np.testing.assert_allclose(r["normal"], (xb - xa) / dist, atol=2e-3)


def _segment_box_distance(c, axis, hh, hb):
    """distance from the segment c + t axis, |t| <= hh, to an axis-aligned box at the origin."""
    f = lambda t: np.linalg.norm(np.maximum(np.abs(c + t[0] * axis) - hb, 0.0))
    best = min((minimize(f, [t0], bounds=[(-hh, hh)]) for t0 in np.linspace(-hh, hh, 9)), key=lambda r: r.fun)
    return best.fun


# ------------------------------------------------------------------- capsule-box
def test_capsule_lying_on_box():
    """Horizontal capsule on a box top: two contacts under the segment ends, effective radius = r.

    The pair runs on the capsule's core (radius 1e-4), so the depth reads pen + 1e-4.
    """
    r_c, hh, pen = 0.1, 0.3, 0.004
    q = qaxis((0, 1, 0), math.pi / 2)  # capsule axis (local z) along world x
    r = convex_pair(CAP, (r_c, hh, 0), xf((0.2, 0.05, 0.5 + r_c - pen), q), BOX, (1, 1, 0.5), xf((0, 0, 0)), 0.02)
    assert r["count"] == 2
    np.testing.assert_allclose(r["radii"], [r_c, 0.0])
    np.testing.assert_allclose(r["normal"], [0, 0, -1], atol=1e-4)
    np.testing.assert_allclose(np.sort(r["pos"][:, 0]), [0.2 - hh, 0.2 + hh], atol=2e-5)
    np.testing.assert_allclose(r["pos"][:, 1], 0.05, atol=2e-5)
    np.testing.assert_allclose(r["dist"], -(pen + INFLATE), atol=2e-6)
    # writer witnesses: centre - n (d/2 + r) lies on the capsule's segment (the radius travels as
    # margin), centre + n d/2 on the box top
    xa = r["pos"] - r["normal"] * (0.5 * r["dist"][:, None] + r_c)
    xb = r["pos"] + r["normal"] * (0.5 * r["dist"][:, None])
    np.testing.assert_allclose(xa[:, 2], 0.5 + r_c - pen, atol=1e-4)
    np.testing.assert_allclose(xb[:, 2], 0.5, atol=1e-4)


def test_capsule_standing_on_box():
    """Vertical capsule: one contact under the lower cap."""
    r_c, hh, pen = 0.05, 0.2, 0.003
    r = convex_pair(CAP, (r_c, hh, 0), xf((0.1, -0.2, 0.5 + hh + r_c - pen)), BOX, (1, 1, 0.5), xf((0, 0, 0)),
                    0.02)
    assert r["count"] == 1
    np.testing.assert_allclose(r["pos"][0, :2], [0.1, -0.2], atol=2e-5)
    # the cap's core is a 1e-4 sphere: MPR stops within its 1e-5 tolerance on the curved tip
    np.testing.assert_allclose(r["dist"][0], -(pen + INFLATE), atol=1e-5)


@pytest.mark.parametrize("seed", range(4))
def test_separated_capsule_box_distance(seed):
    """Capsule away from a box: one contact at the segment-to-box distance - r - 1e-4.

    The GJK runs on the capsule's core (its segment with a 1e-4 radius) and the true radius is
    subtracted afterwards, so the distance reads 1e-4 short.
    """
    rng = np.random.default_rng(10 + seed)
    hb = rng.uniform(0.2, 0.5, 3)
    r_c, hh = rng.uniform(0.03, 0.1), rng.uniform(0.05, 0.3)
    axis_q = qaxis(rng.normal(size=3), rng.uniform(0, math.pi))
    axis = qmat(axis_q)[:, 2]
    d = rng.normal(size=3)
    d /= np.linalg.norm(d)
    c = d * (np.linalg.norm(hb) + hh + r_c + rng.uniform(0.05, 0.3))
    expected = _segment_box_distance(c, axis, hh, hb) - r_c - 1e-4
    r = convex_pair(CAP, (r_c, hh, 0), xf(tuple(c), axis_q), BOX, tuple(hb), xf((0, 0, 0)), 0.01)
    assert r["count"] == 1
    assert r["dist"][0] == pytest.approx(expected, abs=5e-5)


# --------------------------------------------------------------------- scene level
def _tf_point(x, p):
    return np.asarray(x[:3], float) + qmat(x[3:]) @ np.asarray(p, float)


def test_box_pile_scene_contacts():
    """The pile's contacts: every pair type supported, unit normals, and the writer's gap rule
    (witness separation minus margins <= gap) holds for every written contact."""
    from oracle.collide import collide
    from sap_warp_amd.scenes import build_box_pile_scene

    model = build_box_pile_scene(8, seed=5)
    r = collide(model, model.body_q, 64)
    assert r["unsupported"] == 0
    n = r["count"]
    st = np.asarray(model.shape_type)
    s0, s1 = r["shape0"][:n], r["shape1"][:n]
    kinds = set(zip(st[s0].tolist(), st[s1].tolist()))
    assert {(7, 7), (4, 7), (1, 7)} <= kinds
    np.testing.assert_allclose(np.linalg.norm(r["normal"][:n], axis=1), 1.0, atol=1e-6)
    sb = np.asarray(model.shape_body)
    bq = np.asarray(model.body_q, dtype=np.float64)
    gap = np.asarray(model.shape_gap)
    for i in range(n):
        xa = _tf_point(bq[sb[s0[i]]], r["point0"][i]) if sb[s0[i]] >= 0 else r["point0"][i]
        xb = _tf_point(bq[sb[s1[i]]], r["point1"][i]) if sb[s1[i]] >= 0 else r["point1"][i]
        d = np.dot(xb - xa, r["normal"][i]) - (r["margin0"][i] + r["margin1"][i])
        assert d <= gap[s0[i]] + gap[s1[i]] + 1e-6


@pytest.mark.parametrize("C", [2, 3, 5, 9])
def test_box_pile_capacity_prefix(C):
    """At a small capacity each world's list is a prefix of its full list (claims only grow, so once
    a batch or contact misses, everything after it misses too)."""
    from oracle.collide import collide
    from sap_warp_amd.scenes import build_box_pile_scene

    model = build_box_pile_scene(6, seed=8)
    full = collide(model, model.body_q, 64)
    small = collide(model, model.body_q, C)
    for w in range(model.world_count):
        a0, a1 = full["world_start"][w], full["world_start"][w + 1]
        b0, b1 = small["world_start"][w], small["world_start"][w + 1]
        assert b1 - b0 <= C
        for key in ("shape0", "shape1", "point0", "point1", "normal"):
            np.testing.assert_array_equal(small[key][b0:b1], full[key][a0:a0 + (b1 - b0)])

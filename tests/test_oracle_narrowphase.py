"""Pin the float32 collision oracle with analytic geometry.

For every primitive pair the reference handles (plane-sphere/capsule/box,
sphere-sphere/capsule/box, capsule-capsule; ``sim/collision/primitive.py``),
the signed distance recovered from a written contact,
phi = (x1 - x0) . n - margin0 - margin1 with x0, x1 the world witness points,
must equal the analytic surface distance; the normal must point from shape0 to
shape1; a pair is reported iff phi <= gap0 + gap1 (``contact_data.py:36-56``).
Distances to segments and boxes are checked against brute force on dense
samples, which shares nothing with the closed-form routines under test.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import collide as oc
from sap_warp_amd.runtime import GEO_BOX, GEO_CAPSULE, GEO_SPHERE
from sap_warp_amd.scenes import SceneBuilder, ShapeCfg, quat_from_axis_angle, quat_to_matrix, tf

GAP = 0.05


def _scene(shapes, ground=False, margin=0.0):
    """One world, one free body per shape (body pose = shape pose, shape at the body origin)."""
    b = SceneBuilder(rigid_gap=GAP)
    b.current_world = 0
    joints = []
    for (typ, scale, pose) in shapes:
        body = b.add_body(pose)
        b.add_shape(body, typ, tf(), scale, ShapeCfg(density=500.0, gap=GAP, margin=margin))
        joints.append(b.add_free_joint(body))
    b.add_articulation(joints[:1])
    for j in joints[1:]:
        b.add_articulation([j])
    b.current_world = -1
    if ground:
        b.add_ground_plane(ShapeCfg(density=0.0, gap=GAP, collision_group=-1, margin=margin))
    model = b.finalize()
    body_q = np.stack([np.asarray(s[2], dtype=np.float32) for s in shapes])
    return model, body_q


def _world_point(model, body_q, shape, p):
    body = int(model.shape_body[shape])
    p = np.asarray(p, dtype=np.float64)
    if body < 0:
        return p
    x = body_q[body].astype(np.float64)
    return x[:3] + quat_to_matrix(x[3:]) @ p


def _contacts(model, body_q, C=16):
    c = oc.collide(model, body_q, C)
    assert c["unsupported"] == 0
    out = []
    for i in range(c["count"]):
        s0, s1 = int(c["shape0"][i]), int(c["shape1"][i])
        x0 = _world_point(model, body_q, s0, c["point0"][i])
        x1 = _world_point(model, body_q, s1, c["point1"][i])
        n = c["normal"][i].astype(np.float64)
        phi = float((x1 - x0) @ n) - float(c["margin0"][i]) - float(c["margin1"][i])
        out.append(dict(s0=s0, s1=s1, x0=x0, x1=x1, n=n, phi=phi))
    return out


def _seg(center, axis, hh):
    return center - hh * axis, center + hh * axis


def _seg_seg_distance_brute(a0, a1, b0, b1, k=801):
    """Dense sampling of both segments, then a bounded local polish of the best pair."""
    from scipy.optimize import minimize

    t = np.linspace(0.0, 1.0, k)
    pa = a0[None] + t[:, None] * (a1 - a0)[None]
    pb = b0[None] + t[:, None] * (b1 - b0)[None]
    d = np.linalg.norm(pa[:, None, :] - pb[None, :, :], axis=2)
    i, j = np.unravel_index(np.argmin(d), d.shape)
    f = lambda x: float(np.linalg.norm(a0 + x[0] * (a1 - a0) - b0 - x[1] * (b1 - b0)))  # noqa: E731
    res = minimize(f, x0=[t[i], t[j]], bounds=[(0.0, 1.0), (0.0, 1.0)], method="L-BFGS-B",
                   options=dict(ftol=1e-15, gtol=1e-12))
    return min(float(d[i, j]), float(res.fun))


def _point_box_signed_distance(p_local, half):
    q = np.abs(p_local) - half
    outside = np.linalg.norm(np.maximum(q, 0.0))
    inside = min(float(np.max(q)), 0.0)
    return outside + inside


# ----------------------------------------------------------------------------
@pytest.mark.parametrize("z", [0.3, 0.249, 0.21, 0.2999])
def test_plane_sphere(z):
    r = 0.25
    model, bq = _scene([(GEO_SPHERE, (r, 0, 0), tf((0.1, -0.3, z)))], ground=True)
    cs = _contacts(model, bq)
    expect = z - r
    if expect > 2 * GAP:
        assert cs == []
        return
    assert len(cs) == 1
    c = cs[0]
    assert abs(c["phi"] - expect) < 2e-6
    assert abs(abs(c["n"][2]) - 1.0) < 1e-6
    # normal points from shape0 to shape1: sphere center lies on shape1's side
    centre = np.array([0.1, -0.3, z])
    side = (centre - c["x0"]) @ c["n"] if int(model.shape_type[c["s1"]]) == GEO_SPHERE else (c["x1"] - centre) @ c["n"]
    assert side > 0.0


def test_gap_threshold():
    r = 0.2
    for d, expect in ((2 * GAP - 1e-4, 1), (2 * GAP + 1e-4, 0)):
        model, bq = _scene([(GEO_SPHERE, (r, 0, 0), tf((0.0, 0.0, r + d)))], ground=True)
        assert len(_contacts(model, bq)) == expect, d


@pytest.mark.parametrize("seed", range(6))
def test_sphere_sphere(seed):
    rng = np.random.default_rng(seed)
    r1, r2 = rng.uniform(0.05, 0.3, size=2)
    c1 = rng.uniform(-0.5, 0.5, size=3)
    direction = rng.normal(size=3)
    direction /= np.linalg.norm(direction)
    d = rng.uniform(-0.05, 0.09)
    c2 = c1 + direction * (r1 + r2 + d)
    model, bq = _scene([(GEO_SPHERE, (r1, 0, 0), tf(c1)), (GEO_SPHERE, (r2, 0, 0), tf(c2))])
    cs = _contacts(model, bq)
    if d > 2 * GAP:
        assert cs == []
        return
    assert len(cs) == 1
    assert abs(cs[0]["phi"] - d) < 3e-6
    np.testing.assert_allclose(cs[0]["n"], direction, atol=2e-6)


@pytest.mark.parametrize("tilt", [0.0, 0.3, 1.2])
def test_plane_capsule(tilt):
    r, hh, z = 0.05, 0.2, 0.1
    q = quat_from_axis_angle((1.0, 0.3, 0.0), math.pi / 2 + tilt)
    model, bq = _scene([(GEO_CAPSULE, (r, hh, 0), tf((0.0, 0.0, z), q))], ground=True)
    cs = _contacts(model, bq)
    axis = quat_to_matrix(q) @ np.array([0.0, 0.0, 1.0])
    ends = [np.array([0.0, 0.0, z]) + s * hh * axis for s in (-1.0, 1.0)]
    expect = sorted(e[2] - r for e in ends if e[2] - r <= 2 * GAP)
    got = sorted(c["phi"] for c in cs)
    assert len(got) == len(expect)
    np.testing.assert_allclose(got, expect, atol=3e-6)


@pytest.mark.parametrize("seed", range(4))
def test_plane_box_corners(seed):
    rng = np.random.default_rng(seed)
    half = rng.uniform(0.05, 0.2, size=3)
    q = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, 0.4))
    R = quat_to_matrix(q)
    corners = np.array([[sx, sy, sz] for sx in (-1, 1) for sy in (-1, 1) for sz in (-1, 1)]) * half
    zmin = float(np.min((corners @ R.T)[:, 2]))
    z = -zmin + rng.uniform(-0.01, 0.03)
    model, bq = _scene([(GEO_BOX, tuple(half), tf((0.2, 0.1, z), q))], ground=True)
    cs = _contacts(model, bq)
    heights = np.sort(z + (corners @ R.T)[:, 2])
    expect = heights[heights <= 2 * GAP][:4]
    got = np.sort([c["phi"] for c in cs])
    assert len(got) == len(expect)
    np.testing.assert_allclose(got, expect, atol=3e-6)


@pytest.mark.parametrize("seed", range(6))
def test_capsule_capsule(seed):
    rng = np.random.default_rng(100 + seed)
    r1, r2 = rng.uniform(0.03, 0.08, size=2)
    h1, h2 = rng.uniform(0.1, 0.3, size=2)
    q1 = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    q2 = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    a1 = quat_to_matrix(q1) @ np.array([0, 0, 1.0])
    a2 = quat_to_matrix(q2) @ np.array([0, 0, 1.0])
    c1 = np.zeros(3)
    # place capsule 2 so that its axis passes near capsule 1
    offset = rng.normal(size=3)
    offset /= np.linalg.norm(offset)
    c2 = c1 + offset * (r1 + r2 + rng.uniform(0.0, 0.06)) + rng.uniform(-0.5, 0.5) * h2 * a2
    model, bq = _scene([(GEO_CAPSULE, (r1, h1, 0), tf(c1, q1)), (GEO_CAPSULE, (r2, h2, 0), tf(c2, q2))])
    cs = _contacts(model, bq)
    s1 = _seg(c1, a1, h1)
    s2 = _seg(c2, a2, h2)
    dist = _seg_seg_distance_brute(*s1, *s2) - r1 - r2
    if dist > 2 * GAP + 1e-3:
        assert cs == []
        return
    assert len(cs) >= 1
    # the closest-feature contact reproduces the segment distance
    assert abs(min(c["phi"] for c in cs) - dist) < 5e-6


@pytest.mark.parametrize("seed", range(6))
def test_sphere_capsule(seed):
    rng = np.random.default_rng(200 + seed)
    rs, rc, hh = rng.uniform(0.05, 0.15), rng.uniform(0.03, 0.08), rng.uniform(0.1, 0.3)
    qc = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    ac = quat_to_matrix(qc) @ np.array([0, 0, 1.0])
    cc = np.array([0.1, 0.2, 0.3])
    foot = cc + rng.uniform(-1.2, 1.2) * hh * ac
    perp = np.cross(ac, rng.normal(size=3))
    perp /= np.linalg.norm(perp)
    sc = foot + perp * (rs + rc + rng.uniform(-0.02, 0.08))
    model, bq = _scene([(GEO_SPHERE, (rs, 0, 0), tf(sc)), (GEO_CAPSULE, (rc, hh, 0), tf(cc, qc))])
    cs = _contacts(model, bq)
    t = np.clip((sc - cc) @ ac, -hh, hh)
    dist = float(np.linalg.norm(sc - (cc + t * ac))) - rs - rc
    if dist > 2 * GAP:
        assert cs == []
        return
    assert len(cs) == 1
    assert abs(cs[0]["phi"] - dist) < 3e-6


@pytest.mark.parametrize("seed", range(8))
def test_sphere_box(seed):
    rng = np.random.default_rng(300 + seed)
    half = rng.uniform(0.05, 0.2, size=3)
    qb = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0, math.pi))
    Rb = quat_to_matrix(qb)
    cb = np.array([-0.1, 0.05, 0.2])
    r = rng.uniform(0.03, 0.1)
    # a centre outside the box at surface distance about r (or slightly inside the box)
    p_local = rng.normal(size=3) * half * 2.0
    closest = np.clip(p_local, -half, half)
    if np.all(closest == p_local):
        p_local = p_local * 0.5  # inside: keep it inside, the sphere overlaps deeply
    else:
        out = (p_local - closest) / np.linalg.norm(p_local - closest)
        p_local = closest + out * (r + rng.uniform(-0.01, 0.07))
    sc = cb + Rb @ p_local
    model, bq = _scene([(GEO_SPHERE, (r, 0, 0), tf(sc)), (GEO_BOX, tuple(half), tf(cb, qb))])
    cs = _contacts(model, bq)
    dist = _point_box_signed_distance(p_local, half) - r
    if dist > 2 * GAP:
        assert cs == []
        return
    assert len(cs) == 1
    assert abs(cs[0]["phi"] - dist) < 3e-6


def test_normal_flips_with_shape_order():
    """Swapping which shape comes first flips the normal and keeps phi."""
    r1, r2 = 0.1, 0.2
    c1, c2 = np.array([0.0, 0.0, 0.0]), np.array([0.25, 0.1, -0.05])
    m_a, bq_a = _scene([(GEO_SPHERE, (r1, 0, 0), tf(c1)), (GEO_SPHERE, (r2, 0, 0), tf(c2))])
    m_b, bq_b = _scene([(GEO_SPHERE, (r2, 0, 0), tf(c2)), (GEO_SPHERE, (r1, 0, 0), tf(c1))])
    ca, cb = _contacts(m_a, bq_a), _contacts(m_b, bq_b)
    assert len(ca) == len(cb) == 1
    np.testing.assert_allclose(ca[0]["n"], -cb[0]["n"], atol=1e-7)
    assert abs(ca[0]["phi"] - cb[0]["phi"]) < 1e-6


def test_shape_margin_is_subtracted():
    r, z, m = 0.2, 0.23, 0.01
    model, bq = _scene([(GEO_SPHERE, (r, 0, 0), tf((0.0, 0.0, z)))], ground=True, margin=m)
    cs = _contacts(model, bq)
    assert len(cs) == 1
    assert abs(cs[0]["phi"] - (z - r - 2 * m)) < 2e-6

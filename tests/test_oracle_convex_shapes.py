"""Pins of the oracle's MPR / GJK path for round support-mapped shapes (ellipsoid, cylinder,
cone) and of the infinite-plane box proxy.

``sim/collision/support.py:190-258`` (ellipsoid, cylinder, cone support maps),
``core.py:172-275`` (rolling projection), ``core.py:360-378`` (single contact for sphere and
ellipsoid pairs) and ``core.py:507-566`` / ``narrow_phase.py:640-720`` (plane as a box proxy
behind a bounding-sphere half-space test), restated in ``oracle/convex.h`` and
``oracle/narrowphase.c``.  Every expected value comes from geometry sampled here, not from the
oracle's formulas:

* support maps: the returned point lies on the shape's surface (an implicit surface test) and
  maximises the dot product with the direction over a dense surface sample;
* separated pairs: the reported distance equals the brute-force minimum distance between dense
  surface samples of the two shapes, and the normal points along the brute-force closest-point
  difference; sphere and ellipsoid pairs give exactly one contact;
* resting pairs with a known penetration: an upright cone on a box face gives rim points at the
  known depth, a cylinder rolling on a box and a cone lying on its slant on the ground give
  contacts in the plane through the shape's axis and the normal, at the known depth;
* plane proxy: a cone whose bounding sphere clears the plane gives nothing, a cone hovering
  within the gap gives its true clearance, and a finite plane is reported unsupported.

Planting a mistake (a cone apex test with the wrong slope, an ellipsoid support without the
squared semi-axes, a missing rolling projection, a proxy box on the wrong side of the plane)
fails these tests.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.spatial import cKDTree

from oracle import collide as oc
from sap_warp_amd.runtime import GEO_BOX, GEO_CAPSULE, GEO_CONE, GEO_CYLINDER, GEO_ELLIPSOID, GEO_SPHERE
from sap_warp_amd.scenes import (SceneBuilder, ShapeCfg, quat_from_axis_angle, quat_mul, quat_to_matrix,
                                 tf)

INFLATE = 1e-4  # MPR runs on shapes inflated by this much (manifold depths read this much deeper)
CORE = 1e-4  # a sphere / capsule enters MPR / GJK with this radius and its full radius is subtracted
# afterwards, so its distances read this much deeper


# ------------------------------------------------------------------ shape geometry (float64)
def _grid(n_u, n_v):
    return np.meshgrid(np.linspace(0.0, 2.0 * np.pi, n_u, endpoint=False), np.linspace(0.0, 1.0, n_v))


def surface(t, s, n=160):
    """Dense sample of a shape's surface in its own frame."""
    u, v = _grid(2 * n, n)
    u, v = u.ravel(), v.ravel()
    if t in (GEO_SPHERE, GEO_ELLIPSOID):
        ax = (s[0], s[0], s[0]) if t == GEO_SPHERE else s
        th = v * np.pi
        return np.stack([ax[0] * np.sin(th) * np.cos(u), ax[1] * np.sin(th) * np.sin(u), ax[2] * np.cos(th)], -1)
    if t == GEO_CAPSULE:
        r, hh = s[0], s[1]
        side = np.stack([r * np.cos(u), r * np.sin(u), hh * (2.0 * v - 1.0)], -1)
        th = v * np.pi
        ball = np.stack([r * np.sin(th) * np.cos(u), r * np.sin(th) * np.sin(u), r * np.cos(th)], -1)
        return np.concatenate([side, ball + [0, 0, hh], ball - [0, 0, hh]])
    if t == GEO_CYLINDER:
        r, hh = s[0], s[1]
        side = np.stack([r * np.cos(u), r * np.sin(u), hh * (2.0 * v - 1.0)], -1)
        disk = np.stack([r * np.sqrt(v) * np.cos(u), r * np.sqrt(v) * np.sin(u), np.zeros_like(u)], -1)
        return np.concatenate([side, disk + [0, 0, hh], disk - [0, 0, hh]])
    if t == GEO_CONE:
        r, hh = s[0], s[1]
        rad = r * v  # v = 0 apex, 1 base rim
        side = np.stack([rad * np.cos(u), rad * np.sin(u), hh - 2.0 * hh * v], -1)
        disk = np.stack([r * np.sqrt(v) * np.cos(u), r * np.sqrt(v) * np.sin(u), np.full_like(u, -hh)], -1)
        return np.concatenate([side, disk])
    if t == GEO_BOX:
        a, b = np.meshgrid(np.linspace(-1, 1, n), np.linspace(-1, 1, n))
        a, b = a.ravel(), b.ravel()
        faces = []
        for k in range(3):
            i, j = [m for m in range(3) if m != k]
            for sg in (-1.0, 1.0):
                f = np.zeros((a.size, 3))
                f[:, i], f[:, j], f[:, k] = a * s[i], b * s[j], sg * s[k]
                faces.append(f)
        return np.concatenate(faces)
    raise ValueError(t)


def residual(t, s, p):
    """Signed surface test: 0 on the surface, negative inside (lengths, first order)."""
    x, y, z = p
    if t == GEO_SPHERE:
        return math.sqrt(x * x + y * y + z * z) - s[0]
    if t == GEO_ELLIPSOID:
        return (math.sqrt((x / s[0]) ** 2 + (y / s[1]) ** 2 + (z / s[2]) ** 2) - 1.0) * min(s)
    if t == GEO_CAPSULE:
        zc = min(max(z, -s[1]), s[1])
        return math.sqrt(x * x + y * y + (z - zc) ** 2) - s[0]
    if t == GEO_CYLINDER:
        return max(math.hypot(x, y) - s[0], abs(z) - s[1])
    if t == GEO_CONE:
        r, hh = s[0], s[1]
        rho = r * (hh - z) / (2.0 * hh)  # radius of the cross-section at height z
        cos_a = 2.0 * hh / math.hypot(r, 2.0 * hh)
        return max((math.hypot(x, y) - rho) * cos_a, -hh - z, z - hh)
    if t == GEO_BOX:
        return max(abs(x) - s[0], abs(y) - s[1], abs(z) - s[2])
    raise ValueError(t)


SCALES = {
    GEO_SPHERE: lambda g: (g.uniform(0.03, 0.08), 0.0, 0.0),
    GEO_CAPSULE: lambda g: (g.uniform(0.02, 0.05), g.uniform(0.04, 0.1), 0.0),
    GEO_ELLIPSOID: lambda g: tuple(g.uniform(0.04, 0.12, size=3)),
    GEO_CYLINDER: lambda g: (g.uniform(0.04, 0.1), g.uniform(0.03, 0.1), 0.0),
    GEO_CONE: lambda g: (g.uniform(0.04, 0.1), g.uniform(0.03, 0.1), 0.0),
    GEO_BOX: lambda g: tuple(g.uniform(0.04, 0.1, size=3)),
}


def _rand_q(g):
    return quat_from_axis_angle(g.normal(size=3), g.uniform(0.0, math.pi))


# ------------------------------------------------------------------ support maps
@pytest.mark.parametrize("t", [GEO_ELLIPSOID, GEO_CYLINDER, GEO_CONE, GEO_SPHERE, GEO_CAPSULE, GEO_BOX])
def test_support_maximises_over_surface(t):
    g = np.random.default_rng(t)
    for _ in range(4):
        s = SCALES[t](g)
        pts = surface(t, s, n=240)
        dirs = list(g.normal(size=(60, 3)))
        dirs += [np.array(v, float) for v in ((0, 0, 1), (0, 0, -1), (1, 0, 0), (0, 1, 0), (0.3, -0.2, 0.0))]
        for d in dirs:
            p = oc.support(t, s, d).astype(np.float64)
            assert abs(residual(t, s, p)) < 2e-6 * max(s)
            best = float(np.max(pts @ d))
            assert float(p @ d) >= best - 1e-6 * np.linalg.norm(d), (t, s, d)


def test_cone_support_apex_rim_boundary():
    """The apex wins exactly when the direction's elevation exceeds the slant's: dz >= r / (2 hh)
    |d_xy| (a direction just below the slant picks the rim, just above picks the apex)."""
    r, hh = 0.06, 0.08
    slope = r / (2.0 * hh)
    for az in np.linspace(0.0, 2.0 * np.pi, 7):
        for sgn, want_apex in ((1.01, True), (0.99, False)):
            d = np.array([math.cos(az), math.sin(az), sgn * slope])
            p = oc.support(GEO_CONE, (r, hh, 0.0), d)
            if want_apex:
                np.testing.assert_allclose(p, [0.0, 0.0, hh], atol=1e-7)
            else:
                np.testing.assert_allclose(p, [r * math.cos(az), r * math.sin(az), -hh], atol=1e-6)


# ------------------------------------------------------------------ separated pairs vs brute force
def _world(t, s, pose, n):
    R = quat_to_matrix(pose[3:])
    return surface(t, s, n=n) @ R.T + np.asarray(pose[:3])


def _brute(ta, sa, pa, tb, sb, pb, n=120):
    A, B = _world(ta, sa, pa, n), _world(tb, sb, pb, n)
    d, j = cKDTree(B).query(A)
    i = int(np.argmin(d))
    return float(d[i]), A[i], B[j[i]]


PAIRS = [(GEO_SPHERE, GEO_ELLIPSOID), (GEO_ELLIPSOID, GEO_ELLIPSOID), (GEO_ELLIPSOID, GEO_CYLINDER),
         (GEO_CAPSULE, GEO_CYLINDER), (GEO_CYLINDER, GEO_CYLINDER), (GEO_CYLINDER, GEO_BOX),
         (GEO_BOX, GEO_CONE), (GEO_CYLINDER, GEO_CONE), (GEO_CONE, GEO_CONE), (GEO_SPHERE, GEO_CONE),
         (GEO_CAPSULE, GEO_ELLIPSOID), (GEO_ELLIPSOID, GEO_CONE), (GEO_ELLIPSOID, GEO_BOX)]


@pytest.mark.parametrize("ta,tb", PAIRS)
@pytest.mark.parametrize("seed", range(2))
def test_separated_pair_distance(ta, tb, seed):
    g = np.random.default_rng(1000 * ta + 10 * tb + seed)
    sa, sb = SCALES[ta](g), SCALES[tb](g)
    pa = tf((0.0, 0.0, 0.0), _rand_q(g))
    qb = _rand_q(g)
    u = g.normal(size=3)
    u /= np.linalg.norm(u)
    target = g.uniform(0.006, 0.03)
    t = max(sa) + max(sb) + target
    for _ in range(6):  # walk B along u until the sampled clearance is the target
        d, _, _ = _brute(ta, sa, pa, tb, sb, tf(u * t, qb), n=40)
        t += target - d
    pb = tf(u * t, qb)
    d_bf, xa, xb = _brute(ta, sa, pa, tb, sb, pb)
    res = oc.convex_pair(ta, sa, pa, tb, sb, pb, gap_sum=0.05)
    assert res["count"] >= 1
    if GEO_SPHERE in (ta, tb) or GEO_ELLIPSOID in (ta, tb):
        assert res["count"] == 1
    dmin = float(np.min(res["dist"]))
    expect = d_bf - CORE * sum(t in (GEO_SPHERE, GEO_CAPSULE) for t in (ta, tb))
    # the sampled clearance is an upper bound of the true one, within the sample spacing; GJK stops
    # once the support along -v gains less than 1e-4 on |v| (or repeats a simplex vertex), so it
    # overshoots the true distance by about 1e-4
    assert dmin == pytest.approx(expect, abs=1.5e-3)
    assert dmin <= expect + 2e-4
    k = (xb - xa) / np.linalg.norm(xb - xa)
    assert float(k @ res["normal"]) > 0.97


# ------------------------------------------------------------------ resting pairs with a known depth
def _qz(yaw):
    return quat_from_axis_angle((0, 0, 1), yaw)


@pytest.mark.parametrize("seed", range(4))
def test_upright_cone_on_box_face(seed):
    g = np.random.default_rng(seed)
    hb = (0.3, 0.3, 0.05)
    r, hh = g.uniform(0.04, 0.08), g.uniform(0.04, 0.08)
    pen = g.uniform(0.0005, 0.003)
    yaw = g.uniform(-np.pi, np.pi)
    c = np.array([g.uniform(-0.1, 0.1), g.uniform(-0.1, 0.1), 0.05 + hh - pen])
    res = oc.convex_pair(GEO_BOX, hb, tf((0, 0, 0)), GEO_CONE, (r, hh, 0.0), tf(c, _qz(yaw)), gap_sum=0.02)
    assert res["count"] >= 3
    np.testing.assert_allclose(res["normal"], [0.0, 0.0, 1.0], atol=1e-5)
    for p, d in zip(res["pos"], res["dist"]):
        assert d == pytest.approx(-pen - INFLATE, abs=2e-5)
        assert p[2] == pytest.approx(0.05 - 0.5 * pen, abs=2e-5)  # midway between the surfaces
        assert math.hypot(p[0] - c[0], p[1] - c[1]) == pytest.approx(r, abs=1e-4)  # on the base rim


@pytest.mark.parametrize("seed", range(4))
def test_cylinder_rolling_on_box(seed):
    g = np.random.default_rng(10 + seed)
    hb = (0.3, 0.3, 0.05)
    r, hh = g.uniform(0.04, 0.08), g.uniform(0.04, 0.1)
    pen = g.uniform(0.0005, 0.003)
    tilt = math.radians(g.uniform(0.0, 1.5))  # within the 2 degree rolling band
    q = quat_mul(_qz(g.uniform(-np.pi, np.pi)), quat_from_axis_angle((1, 0, 0), math.pi / 2 - tilt))
    a = quat_to_matrix(q)[:, 2]
    low = r * math.sqrt(1.0 - a[2] ** 2) + hh * abs(a[2])
    c = np.array([0.02, -0.03, 0.05 + low - pen])
    res = oc.convex_pair(GEO_CYLINDER, (r, hh, 0.0), tf(c, q), GEO_BOX, hb, tf((0, 0, 0)), gap_sum=0.02)
    assert res["count"] >= 1
    n = res["normal"]  # cylinder (A) to box (B): down
    np.testing.assert_allclose(n, [0.0, 0.0, -1.0], atol=1e-4)
    pn = np.cross(a, -n)
    pn /= np.linalg.norm(pn)
    # the manifold depth of a curved side comes from planes fitted through support points probed 2
    # degrees off the normal, which sit up to r (1 - cos 4 deg) ~ 1e-4 above the lowest generator
    assert min(res["dist"]) == pytest.approx(-pen - INFLATE, abs=1.5e-4)
    for p in res["pos"]:
        assert abs(float((p - c) @ pn)) < 2e-6  # projected onto the plane of axis and normal
        assert abs(float((p - c) @ a)) <= hh + 1e-4


def _ground_scene(t, s, pose, plane_scale=(0.0, 0.0, 0.0), gap=0.02):
    b = SceneBuilder(rigid_gap=gap)
    b.current_world = 0
    body = b.add_body(pose)
    b.add_shape(body, t, tf(), s, ShapeCfg(density=500.0, gap=gap))
    b.add_articulation([b.add_free_joint(body)])
    b.current_world = -1
    from sap_warp_amd.runtime import GEO_PLANE

    b.add_shape(-1, GEO_PLANE, tf(), plane_scale, ShapeCfg(density=0.0, gap=gap, collision_group=-1))
    model = b.finalize()
    return model, np.asarray(pose, np.float32)[None, :]


def _contacts(model, bq):
    c = oc.collide(model, bq, 16)
    out = []
    x = bq[0].astype(np.float64)
    R = quat_to_matrix(x[3:])
    for i in range(c["count"]):
        p0 = np.asarray(c["point0"][i], np.float64)  # on the plane (a global shape: world frame)
        p1 = x[:3] + R @ np.asarray(c["point1"][i], np.float64)
        n = c["normal"][i].astype(np.float64)
        phi = float((p1 - p0) @ n) - float(c["margin0"][i]) - float(c["margin1"][i])
        out.append(dict(x0=p0, x1=p1, n=n, phi=phi))
    return out, c["unsupported"]


@pytest.mark.parametrize("seed", range(4))
def test_cone_lying_on_ground(seed):
    g = np.random.default_rng(20 + seed)
    r, hh = g.uniform(0.04, 0.08), g.uniform(0.04, 0.08)
    pen = g.uniform(0.0005, 0.003)
    alpha = math.atan2(r, 2.0 * hh)
    q = quat_mul(_qz(g.uniform(-np.pi, np.pi)), quat_from_axis_angle((0, 1, 0), math.pi / 2 + alpha))
    a = quat_to_matrix(q)[:, 2]
    c = np.array([0.1, 0.2, hh * math.sin(alpha) - pen])
    model, bq = _ground_scene(GEO_CONE, (r, hh, 0.0), tf(c, q))
    cs, unsup = _contacts(model, bq)
    assert unsup == 0 and len(cs) >= 1
    pn = np.cross(a, [0.0, 0.0, -1.0])  # the cone is shape B: its normal is -n
    pn /= np.linalg.norm(pn)
    apex = c + hh * a
    for k in cs:
        np.testing.assert_allclose(k["n"], [0.0, 0.0, 1.0], atol=1e-5)
        assert k["phi"] == pytest.approx(-pen - INFLATE, abs=6e-5)
        mid = 0.5 * (k["x0"] + k["x1"])
        assert abs(float((mid - c) @ pn)) < 2e-6
        along = float((mid - apex) @ -a)  # distance from the apex along the axis
        assert -pen - 1e-4 <= along <= 2.0 * hh + pen + 1e-4  # midway points sit pen / 2 off the slant


@pytest.mark.parametrize("seed", range(3))
def test_cone_plane_proxy_reach(seed):
    g = np.random.default_rng(30 + seed)
    r, hh = g.uniform(0.04, 0.08), g.uniform(0.04, 0.08)
    q = _rand_q(g)
    R = quat_to_matrix(q)
    low = float(np.min((surface(GEO_CONE, (r, hh, 0.0)) @ R.T)[:, 2]))
    # hovering within the gap: one or more contacts, the nearest at the true clearance
    clear = g.uniform(0.002, 0.015)
    model, bq = _ground_scene(GEO_CONE, (r, hh, 0.0), tf((0.0, 0.0, clear - low), q))
    cs, unsup = _contacts(model, bq)
    assert unsup == 0 and cs
    assert min(k["phi"] for k in cs) == pytest.approx(clear, abs=5e-4)
    # far above: the bounding sphere clears the plane's half-space, no contacts
    model, bq = _ground_scene(GEO_CONE, (r, hh, 0.0), tf((0.0, 0.0, 1.0), q))
    cs, unsup = _contacts(model, bq)
    assert unsup == 0 and not cs
    # a finite plane against a cone is not on this path
    model, bq = _ground_scene(GEO_CONE, (r, hh, 0.0), tf((0.0, 0.0, clear - low), q), plane_scale=(1.0, 1.0, 0.0))
    cs, unsup = _contacts(model, bq)
    assert unsup == 1 and not cs


def test_sphere_on_ellipsoid_axis_depth():
    """A sphere on an ellipsoid's semi-axis: one contact at the known depth, normal along the axis."""
    g = np.random.default_rng(5)
    for k in range(3):
        semi = tuple(g.uniform(0.05, 0.12, size=3))
        rs = g.uniform(0.02, 0.05)
        pen = g.uniform(0.0005, 0.004)
        e = np.eye(3)[k]
        res = oc.convex_pair(GEO_SPHERE, (rs, 0, 0), tf(e * (semi[k] + rs - pen)), GEO_ELLIPSOID, semi,
                             tf((0, 0, 0)), gap_sum=0.02)
        assert res["count"] == 1
        assert res["dist"][0] == pytest.approx(-pen - CORE, abs=5e-5)
        np.testing.assert_allclose(res["normal"], -e, atol=1e-3)

"""Pin the oracle's plane-ellipsoid, sphere-cylinder and plane-cylinder contacts with geometry.

``sim/collision/primitive.py:322-351, 431-657`` restated in ``oracle/narrowphase.c``.
Checks that share nothing with the closed-form routines under test:

* plane-ellipsoid: phi equals the minimum height over a dense parametric sample of the
  ellipsoid surface; the ellipsoid witness lies on the surface with its outward normal
  along -n;
* sphere-cylinder: phi equals the capped-cylinder signed distance field (the
  max / length form, inside and outside) at the sphere centre minus the sphere radius,
  in the side, cap and rim regions, and the normal is the field's negative gradient;
* plane-cylinder: the first contact is the deepest point of a dense rim sample; every
  contact's cylinder witness lies on a rim circle (distance r from a cap centre, in the
  cap plane) or a side generator end; near upright the others are the fixed tripod on
  the near cap, 120 degrees apart; rolling, two side-generator ends and one near-cap rim
  point.  Each reported phi equals the witness height above the plane.

Planting a mistake (no half-depth shift, a cap normal not flipped, a dropped tilt test)
fails these tests.
"""
from __future__ import annotations

import math

import numpy as np
import pytest

from oracle import collide as oc
from sap_warp_amd.runtime import GEO_CYLINDER, GEO_ELLIPSOID, GEO_SPHERE
from sap_warp_amd.scenes import SceneBuilder, ShapeCfg, quat_from_axis_angle, quat_to_matrix, tf

GAP = 0.05


def _scene(shapes, ground=False):
    b = SceneBuilder(rigid_gap=GAP)
    b.current_world = 0
    joints = []
    for k, (typ, scale, pose) in enumerate(shapes):
        body = b.add_body(pose)
        b.add_shape(body, typ, tf(), scale, ShapeCfg(density=500.0, gap=GAP))
        joints.append(b.add_free_joint(body))
    for j in joints:
        b.add_articulation([j])
    b.current_world = -1
    if ground:
        b.add_ground_plane(ShapeCfg(density=0.0, gap=GAP, collision_group=-1))
    model = b.finalize()
    body_q = np.stack([np.asarray(s[2], dtype=np.float32) for s in shapes])
    return model, body_q


def _contacts(model, body_q):
    c = oc.collide(model, body_q, 16)
    assert c["unsupported"] == 0
    out = []
    for i in range(c["count"]):
        s0, s1 = int(c["shape0"][i]), int(c["shape1"][i])
        pts = []
        for s, p in ((s0, c["point0"][i]), (s1, c["point1"][i])):
            b = int(model.shape_body[s])
            p = np.asarray(p, np.float64)
            if b >= 0:
                x = body_q[b].astype(np.float64)
                p = x[:3] + quat_to_matrix(x[3:]) @ p
            pts.append(p)
        n = c["normal"][i].astype(np.float64)
        phi = float((pts[1] - pts[0]) @ n) - float(c["margin0"][i]) - float(c["margin1"][i])
        out.append(dict(s0=s0, s1=s1, x0=pts[0], x1=pts[1], n=n, phi=phi))
    return out


def _pose(rng, center):
    q = quat_from_axis_angle(rng.normal(size=3), rng.uniform(0.0, math.pi))
    return tf(center, q), quat_to_matrix(q)


@pytest.mark.parametrize("seed", range(6))
def test_plane_ellipsoid(seed):
    rng = np.random.default_rng(seed)
    semi = rng.uniform(0.03, 0.15, size=3)
    pose, R = _pose(rng, (0.1, -0.2, 0.0))
    # dense parametric sample of the surface: lowest height
    u, v = np.meshgrid(np.linspace(0, 2 * np.pi, 721), np.linspace(0, np.pi, 361))
    local = np.stack([semi[0] * np.cos(u) * np.sin(v), semi[1] * np.sin(u) * np.sin(v), semi[2] * np.cos(v)], -1)
    low = float(np.min((local.reshape(-1, 3) @ R.T)[:, 2]))
    height = -low + rng.uniform(-0.01, 0.03)
    pose[2] = height
    model, bq = _scene([(GEO_ELLIPSOID, tuple(semi), pose)], ground=True)
    cs = _contacts(model, bq)
    assert len(cs) == 1
    c = cs[0]
    assert c["phi"] == pytest.approx(height + low, abs=2e-5)
    np.testing.assert_allclose(c["n"], [0.0, 0.0, 1.0], atol=1e-7)
    # ellipsoid witness: on the surface, outward normal along -n
    loc = R.T @ (c["x1"] - np.array(pose[:3], np.float64))
    assert float(np.sum((loc / semi) ** 2)) == pytest.approx(1.0, abs=1e-5)
    g = R @ (loc / semi ** 2)
    np.testing.assert_allclose(g / np.linalg.norm(g), [0.0, 0.0, -1.0], atol=1e-4)


def _cyl_sdf(p_local, r, hh):
    d = np.array([math.hypot(p_local[0], p_local[1]) - r, abs(p_local[2]) - hh])
    return min(max(d[0], d[1]), 0.0) + float(np.linalg.norm(np.maximum(d, 0.0)))


@pytest.mark.parametrize("region", ["side", "cap", "rim", "inside_side", "inside_cap"])
@pytest.mark.parametrize("seed", range(4))
def test_sphere_cylinder(region, seed):
    rng = np.random.default_rng(10 * seed + len(region))
    r, hh, rs = rng.uniform(0.05, 0.1), rng.uniform(0.05, 0.12), rng.uniform(0.02, 0.05)
    cpose, R = _pose(rng, (0.0, 0.0, 0.5))
    phi_t = rng.uniform(-0.01, 0.03)
    ang = rng.uniform(0, 2 * np.pi)
    radial = np.array([math.cos(ang), math.sin(ang), 0.0])
    sgn = 1.0 if rng.random() < 0.5 else -1.0
    if region == "side":
        loc = radial * (r + rs + phi_t) + np.array([0, 0, sgn * rng.uniform(0, 0.8) * hh])
    elif region == "cap":
        loc = radial * rng.uniform(0, 0.8) * r + np.array([0, 0, sgn * (hh + rs + phi_t)])
    elif region == "rim":
        e = np.array([math.cos(0.7), 0.0, math.sin(0.7)]) * (rs + abs(phi_t) + 0.005)
        loc = radial * (r + e[0]) + np.array([0, 0, sgn * (hh + e[2])])
    elif region == "inside_side":
        loc = radial * (r - 0.2 * min(r, hh) * rng.uniform(0.1, 0.9)) + np.array([0, 0, sgn * 0.2 * hh])
    else:
        loc = radial * 0.2 * r + np.array([0, 0, sgn * (hh - 0.2 * min(r, hh) * rng.uniform(0.1, 0.9))])
    sp = np.array(cpose[:3]) + R @ loc
    model, bq = _scene([(GEO_SPHERE, (rs, 0.0, 0.0), tf(sp)), (GEO_CYLINDER, (r, hh, 0.0), cpose)])
    cs = _contacts(model, bq)
    sdf = _cyl_sdf(loc, r, hh)
    expect = sdf - rs
    if expect > GAP:
        assert not cs
        return
    assert len(cs) == 1
    c = cs[0]
    assert (c["s0"], c["s1"]) == (0, 1)  # sphere first (type order)
    assert c["phi"] == pytest.approx(expect, abs=2e-5)
    # normal from sphere to cylinder: minus the field's gradient at the sphere centre
    h = 1e-6
    grad = np.array([(_cyl_sdf(loc + h * e, r, hh) - _cyl_sdf(loc - h * e, r, hh)) / (2 * h) for e in np.eye(3)])
    np.testing.assert_allclose(c["n"], -(R @ grad) / np.linalg.norm(grad), atol=2e-3)


@pytest.mark.parametrize("mode", ["flat", "rolling"])
@pytest.mark.parametrize("seed", range(5))
def test_plane_cylinder(mode, seed):
    rng = np.random.default_rng(100 * seed + len(mode))
    r, hh = rng.uniform(0.05, 0.1), rng.uniform(0.05, 0.12)
    tilt = rng.uniform(0.02, 0.3) if mode == "flat" else rng.uniform(0.6, math.pi / 2 - 0.05)
    q = quat_from_axis_angle((0, 0, 1), rng.uniform(-np.pi, np.pi))
    from sap_warp_amd.scenes import quat_mul

    q = quat_mul(q, quat_from_axis_angle((1, 0, 0), tilt))
    R = quat_to_matrix(q)
    a = R[:, 2]
    # dense rim sample: lowest point of the solid cylinder
    t = np.linspace(0, 2 * np.pi, 20001)
    ring = np.stack([np.cos(t), np.sin(t), np.zeros_like(t)], -1) * r
    rims = [ring + np.array([0, 0, hh]), ring - np.array([0, 0, hh])]
    low = min(float(np.min((x @ R.T)[:, 2])) for x in rims)
    height = -low + rng.uniform(-0.004, 0.002)
    pose = tf((0.2, 0.1, height), q)
    model, bq = _scene([(GEO_CYLINDER, (r, hh, 0.0), pose)], ground=True)
    cs = _contacts(model, bq)
    c0 = np.array(pose[:3])
    assert len(cs) >= 2
    assert cs[0]["phi"] == pytest.approx(height + low, abs=2e-5)
    assert cs[0]["phi"] == pytest.approx(min(c["phi"] for c in cs), abs=1e-7)
    near = c0 + a * hh if a[2] < 0 else c0 - a * hh  # the cap closer to the plane
    far = 2 * c0 - near
    for c in cs:
        np.testing.assert_allclose(c["n"], [0.0, 0.0, 1.0], atol=1e-7)
        assert c["phi"] == pytest.approx(c["x1"][2], abs=2e-6)
        on_near = abs(np.linalg.norm(c["x1"] - near) - r) < 2e-5 and abs((c["x1"] - near) @ a) < 2e-5
        on_far = abs(np.linalg.norm(c["x1"] - far) - r) < 2e-5 and abs((c["x1"] - far) @ a) < 2e-5
        assert on_near or on_far
    if mode == "flat":
        assert tilt < math.radians(22.5)
        rest = [c for c in cs[1:]]
        assert all(abs(np.linalg.norm(c["x1"] - near) - r) < 2e-5 for c in cs)
        # tripod points 120 degrees apart on the near rim
        dirs = [(c["x1"] - near) / r for c in rest]
        for i in range(len(dirs)):
            for j in range(i + 1, len(dirs)):
                assert float(dirs[i] @ dirs[j]) == pytest.approx(-0.5, abs=1e-4)
    else:
        assert len(cs) <= 4
        # at least one contact on the far cap: a side generator's other end
        assert any(abs(np.linalg.norm(c["x1"] - far) - r) < 2e-5 for c in cs)

"""Host side of the large-world path (CPU): the sparse contact-row layout, the synthetic
stacked-box problem the bench solves, and the argument checks the native entry points make
before any GPU work.

The stacked-box problem is pinned by rigid-motion invariants rather than by restating its
construction: a velocity field common to every cube (one angular velocity w and one origin
velocity v for the whole block, each cube's origin moving at v + w x o_b) leaves every
cube-cube contact at zero relative velocity, while the ground contacts see exactly the block's
velocity at the contact point; and each A block gives the kinetic energy of a cube.
"""
import ctypes

import numpy as np
import pytest

from sap_warp_amd import _native as nat
from sap_warp_amd.large_world import block_diagonal, dense_rows, sparse_contact_rows, stacked_boxes_problem


def test_sparse_rows_round_trip():
    rng = np.random.default_rng(0)
    n, nc = 40, 9
    J = np.zeros((nc, 3, n))
    for c in range(nc):
        cols = rng.choice(n, size=rng.integers(1, 12), replace=False)
        J[c][:, cols] = rng.normal(size=(3, len(cols)))
    cols, vals = sparse_contact_rows(J, 16)
    assert cols.shape == (nc, 16) and vals.shape == (nc, 3, 16)
    np.testing.assert_array_equal(dense_rows(cols, vals.astype(np.float64), n), J.astype(np.float32))
    with pytest.raises(ValueError):
        sparse_contact_rows(J, 4)


def _rigid_field(P, nx, ny, nz, w, v, size=0.1):
    """SAP velocities [w, v_O] of every cube for the block moving rigidly with (w, v)."""
    nb = nx * ny * nz
    out = np.zeros(6 * nb)
    h = 0.5 * size
    for b in range(nb):
        k, rem = divmod(b, nx * ny)
        j, i = divmod(rem, nx)
        o = np.array([(i - 0.5 * (nx - 1)) * size, (j - 0.5 * (ny - 1)) * size, h + k * size])
        out[6 * b:6 * b + 3] = w
        out[6 * b + 3:6 * b + 6] = v + np.cross(w, o)
    return out


def test_stacked_boxes_rigid_motion_invariants():
    nx, ny, nz = 3, 2, 2
    P = stacked_boxes_problem(nx, ny, nz, seed=4)
    nb = nx * ny * nz
    n = 6 * nb
    cols, vals = P["J_cols"], P["J_vals"]
    J = dense_rows(cols, vals, n)
    nc = J.shape[0]
    # 4 contacts per face-neighbour pair and per bottom cube
    pairs = (nx - 1) * ny * nz + nx * (ny - 1) * nz + nx * ny * (nz - 1)
    assert nc == 4 * (pairs + nx * ny)
    rng = np.random.default_rng(1)
    w, v = rng.normal(size=3), rng.normal(size=3)
    vc = (J @ _rigid_field(P, nx, ny, nz, w, v)).reshape(nc, 3)
    ground = cols[:, 0] < 0
    np.testing.assert_allclose(vc[~ground], 0.0, atol=1e-12)
    # ground contacts see the block's own velocity: a pure vertical translation gives each of
    # them its speed on the normal row (+z) and nothing on the tangent rows
    assert np.any(ground)
    vc2 = (J @ _rigid_field(P, nx, ny, nz, np.zeros(3), np.array([0.0, 0.0, 0.7]))).reshape(nc, 3)
    np.testing.assert_allclose(vc2[ground], np.tile([0.0, 0.0, 0.7], (ground.sum(), 1)), atol=1e-12)
    np.testing.assert_allclose(vc2[~ground], 0.0, atol=1e-12)


def test_stacked_boxes_inertia_blocks_give_kinetic_energy():
    P = stacked_boxes_problem(2, 2, 2, seed=5)
    A = block_diagonal(P["A_blocks"])
    rng = np.random.default_rng(3)
    for b in range(8):
        blk = P["A_blocks"][b]
        m = blk[3, 3]
        inertia = blk[0, 0]
        # a cube of edge 0.1 and density 500-2000: m in [0.5, 2] kg, I = m s^2 / 6
        assert 0.5 <= m <= 2.0
        assert inertia == pytest.approx(m * 0.01 / 6.0, rel=1e-12)
        wv = rng.normal(size=6)
        T = 0.5 * wv @ blk @ wv
        assert T == pytest.approx(0.5 * m * wv[3:] @ wv[3:] + 0.5 * inertia * wv[:3] @ wv[:3], rel=1e-12)
    np.testing.assert_allclose(A, A.T)
    # v* = v0 + dt g on the vertical origin velocity only
    vs = P["v_star"].reshape(-1, 6)
    np.testing.assert_allclose(vs[:, 5], P["dt"] * -9.81)
    np.testing.assert_allclose(vs[:, :5], 0.0)


def test_native_create_rejects_bad_sizes_before_gpu_work():
    lib = nat.load()
    h = ctypes.c_void_p()
    with pytest.raises(ValueError):
        nat.check(lib.sapx_chol_create(0, 1, ctypes.byref(h)), "sapx_chol_create")
    with pytest.raises(ValueError):
        nat.check(lib.sapx_chol_create(64, 0, ctypes.byref(h)), "sapx_chol_create")
    with pytest.raises(ValueError, match="512"):
        nat.check(lib.sapx_chol_create(64 * 600, 1, ctypes.byref(h)), "sapx_chol_create")
    with pytest.raises(ValueError):
        nat.check(lib.sapx_large_create(0, 10, 16, ctypes.byref(h)), "sapx_large_create")
    with pytest.raises(ValueError):
        nat.check(lib.sapx_large_create(100, 10, 300, ctypes.byref(h)), "sapx_large_create")


def test_bin_scene_structure():
    """The bin world: one world of free bodies (one free joint and articulation each), static
    walls and ground as global shapes that never pair with each other, bodies stacked 2 mm apart
    so the oracle's collision finds contacts at once, and every contact inside one world."""
    from oracle import collide as oc
    from sap_warp_amd import hetero

    m = hetero.build_bin_scene(40, seed=3, columns=3)
    assert m.world_count == 1 and m.body_count == 40 and m.joint_dof_count == 240
    assert np.all(np.asarray(m.joint_type) == 4)  # free joints
    assert np.all(np.asarray(m.joint_parent) == -1)
    sw = np.asarray(m.shape_world)
    sb = np.asarray(m.shape_body)
    assert np.sum(sw == -1) == 5 and np.all(sb[sw == -1] == -1)  # ground plane + four walls
    pairs = np.asarray(m.shape_pairs).reshape(-1, 2)
    assert not np.any((sb[pairs[:, 0]] < 0) & (sb[pairs[:, 1]] < 0))
    # every body pair and every body-static pair is a candidate
    assert len(pairs) == 40 * 39 // 2 + 40 * 5
    c = oc.collide(m, m.body_q, 400)
    assert c["count"] > 40 and c["unsupported"] == 0
    b0 = sb[c["shape0"][:c["count"]]]
    b1 = sb[c["shape1"][:c["count"]]]
    assert np.all((b0 >= 0) | (b1 >= 0))

"""Blocked Cholesky of large SPD matrices (``sapx_chol_*``, the replacement of
``BlockCholeskySolverBatched``, ``sim/blocked_cholesky.py:725-901``).

The oracle is the mathematics: solutions are compared with numpy's float64 solve of the same
matrix, and a matrix that is not positive definite must be reported.  Matrices:

* random well-conditioned SPD matrices of every padding case (n = 1, 17, 64, 65, 200, 640,
  1000) and a batch of three;
* SAP-shaped Newton matrices: a block diagonal of body spatial inertias (6 x 6 at the body
  origin, with armature) plus J^T G J of random contacts between body pairs (G the 3 x 3
  contact Hessian blocks, 10-3000 x the pair's effective mass, the range of SAP's
  regularised friction and normal compliances), up to 1024 bodies plus a 40-dof articulation
  (6184 dofs, a 1024-body bin plus a humanoid);
* block-diagonal, block-banded and arrowhead SAP matrices (the factor skips the blocks outside
  the block structure of the factor);
* an indefinite matrix (info names the failing column);
* graph capture: factorize + solve recorded once and replayed on new matrix contents.

The GPU receives float32 matrices, so the reference is numpy's float64 solve of the float32-rounded
matrix.  Bars: 1e-10 relative (max-norm) for the well-conditioned matrices, 1e-8 relative for the
SAP matrices (1e-8 after three float64-residual sweeps on the scaled float32 factor).
"""
import numpy as np
import pytest

pytestmark = pytest.mark.gpu


def _spd(n, rng, cond=1e3):
    Q, _ = np.linalg.qr(rng.normal(size=(n, n)))
    ev = np.exp(rng.uniform(0.0, np.log(cond), size=n))
    return (Q * ev) @ Q.T


def _sap_matrix(bodies, contacts, rng, extra_dof=0):
    """A + J^T G J with A block-diagonal spatial inertias (SAP [w, v_O] order) and contacts
    coupling random body pairs; extra_dof appends a dense SPD articulation block."""
    n = 6 * bodies + extra_dof
    H = np.zeros((n, n))
    mass = np.zeros(bodies)
    for b in range(bodies):
        m = mass[b] = rng.uniform(0.1, 5.0)
        half = rng.uniform(0.03, 0.15, size=3)
        Ic = m / 3.0 * np.diag([half[1] ** 2 + half[2] ** 2, half[0] ** 2 + half[2] ** 2, half[0] ** 2 + half[1] ** 2])
        c = rng.uniform(-0.05, 0.05, size=3)
        cx = np.array([[0, -c[2], c[1]], [c[2], 0, -c[0]], [-c[1], c[0], 0]])
        blk = np.zeros((6, 6))
        blk[:3, :3] = Ic - m * cx @ cx
        blk[:3, 3:] = m * cx
        blk[3:, :3] = -m * cx
        blk[3:, 3:] = m * np.eye(3)
        blk += 1e-3 * np.eye(6)
        H[6 * b:6 * b + 6, 6 * b:6 * b + 6] = blk
    if extra_dof:
        H[6 * bodies:, 6 * bodies:] = _spd(extra_dof, rng, cond=1e4)
    for _ in range(contacts):
        a, b = rng.choice(bodies, size=2, replace=False)
        Jc = np.zeros((3, n))
        for body, sgn in ((a, -1.0), (b, 1.0)):
            r = rng.normal(scale=0.05, size=3)
            rx = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
            Jc[:, 6 * body:6 * body + 3] = -sgn * rx  # v_point = v_O + w x r = v_O - r x w
            Jc[:, 6 * body + 3:6 * body + 6] = sgn * np.eye(3)
        # contact Hessian blocks ~ 1 / r with r = sigma w (friction, sigma = 1e-3) down to
        # beta^2 / (4 pi^2) w (normal), w ~ 1 / m_eff (contact_solve.py:791-822)
        m_eff = 1.0 / (1.0 / mass[a] + 1.0 / mass[b])
        G = _spd(3, rng, cond=1e2) * m_eff * 10.0 ** rng.uniform(1.0, 3.5)
        H += Jc.T @ G @ Jc
    return 0.5 * (H + H.T)


def _solve(A_list, rhs_list, refine=2, max_n=None):
    import torch

    from sap_warp_amd.blocked_cholesky import BlockCholeskySolverBatched

    B = len(A_list)
    n = A_list[0].shape[0]
    solver = BlockCholeskySolverBatched(max_n or n, B, refine=refine)
    A = torch.tensor(np.stack(A_list), dtype=torch.float32, device="cuda:0")
    rhs = torch.tensor(np.stack(rhs_list), dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(rhs)
    print(f"chol batch {B} n {n} refine {refine}", flush=True)
    solver.factorize(A, n)
    solver.solve(rhs, x)
    info = solver.info()
    out = x.cpu().numpy()
    solver.close()
    return out, info


def _rel(x, ref):
    return float(np.max(np.abs(x - ref)) / np.max(np.abs(ref)))


def _f32(A):
    """The matrix the GPU receives (float32); its float64 solve is the reference."""
    return A.astype(np.float32).astype(np.float64)


@pytest.mark.parametrize("n", [1, 17, 64, 65, 200, 640, 1000])
def test_random_spd_solve(gpu_device, n):
    rng = np.random.default_rng(n)
    A = _spd(n, rng)
    b = rng.normal(size=n)
    x, info = _solve([A], [b])
    assert info[0] == 0
    assert _rel(x[0], np.linalg.solve(_f32(A), b)) <= 1e-10


def test_batch_of_three(gpu_device):
    rng = np.random.default_rng(5)
    As = [_spd(300, rng, cond=10.0 ** k) for k in (1, 3, 5)]
    bs = [rng.normal(size=300) for _ in range(3)]
    x, info = _solve(As, bs, refine=3)
    assert np.all(info == 0)
    for k in range(3):
        assert _rel(x[k], np.linalg.solve(_f32(As[k]), bs[k])) <= 1e-9, k


@pytest.mark.parametrize("bodies,contacts,extra", [(40, 120, 0), (200, 700, 35), (1024, 3000, 40)])
def test_sap_newton_matrix(gpu_device, bodies, contacts, extra):
    rng = np.random.default_rng(bodies)
    H = _sap_matrix(bodies, contacts, rng, extra_dof=extra)
    b = rng.normal(size=H.shape[0])
    ref = np.linalg.solve(_f32(H), b)
    x, info = _solve([H], [b], refine=3)
    print("n", H.shape[0], "rel err", _rel(x[0], ref))
    assert info[0] == 0
    assert _rel(x[0], ref) <= 1e-8


def _banded_sap_matrix(bodies, reach, rng, arrow=False):
    """A SAP matrix whose contacts couple bodies at most ``reach`` apart in index (block-banded
    factor, the structure of a pile numbered by position); ``arrow`` adds contacts of body 0 with
    every body (an arrowhead: the whole lower triangle fills)."""
    H = _sap_matrix(bodies, 0, rng)
    n = H.shape[0]
    pairs = [(a, b) for a in range(bodies) for b in range(a + 1, min(bodies, a + reach + 1)) if rng.random() < 0.5]
    if arrow:
        pairs += [(0, b) for b in range(1, bodies)]
    for a, b in pairs:
        Jc = np.zeros((3, 12))  # the contact's columns: body a's six dofs, then body b's
        for k, (body, sgn) in enumerate(((a, -1.0), (b, 1.0))):
            r = rng.normal(scale=0.05, size=3)
            rx = np.array([[0, -r[2], r[1]], [r[2], 0, -r[0]], [-r[1], r[0], 0]])
            Jc[:, 6 * k:6 * k + 3] = -sgn * rx
            Jc[:, 6 * k + 3:6 * k + 6] = sgn * np.eye(3)
        idx = np.r_[6 * a:6 * a + 6, 6 * b:6 * b + 6]
        H[np.ix_(idx, idx)] += Jc.T @ (_spd(3, rng, cond=1e2) * 10.0 ** rng.uniform(1.0, 3.0)) @ Jc
    return 0.5 * (H + H.T)


@pytest.mark.parametrize("bodies,reach,arrow", [(200, 0, False), (300, 12, False), (1024, 40, False),
                                                (150, 3, True)])
def test_block_sparse_matrix(gpu_device, bodies, reach, arrow):
    """Block-diagonal, block-banded and arrowhead structures: the factor skips the tiles outside
    the block structure (and its fill), and the solve is that of the whole matrix."""
    rng = np.random.default_rng(bodies + reach)
    H = _banded_sap_matrix(bodies, reach, rng, arrow)
    b = rng.normal(size=H.shape[0])
    ref = np.linalg.solve(_f32(H), b)
    x, info = _solve([H], [b], refine=3)
    print("n", H.shape[0], "reach", reach, "arrow", arrow, "rel err", _rel(x[0], ref))
    assert info[0] == 0
    assert _rel(x[0], ref) <= 1e-8


def test_indefinite_matrix_reported(gpu_device):
    rng = np.random.default_rng(9)
    A = _spd(150, rng)
    A[100, 100] = -5.0e3
    b = rng.normal(size=150)
    _, info = _solve([A], [b], refine=0)
    assert 1 <= info[0] <= 101


def test_graph_replay_matches_eager(gpu_device):
    """factorize + solve captured once; each replay on new contents equals an eager run."""
    import torch

    from sap_warp_amd.blocked_cholesky import BlockCholeskySolverBatched

    rng = np.random.default_rng(11)
    n = 333
    solver = BlockCholeskySolverBatched(n, 1, refine=2)
    A = torch.zeros((1, n, n), dtype=torch.float32, device="cuda:0")
    rhs = torch.zeros((1, n), dtype=torch.float64, device="cuda:0")
    x = torch.zeros_like(rhs)
    A.copy_(torch.tensor(_spd(n, rng)[None], dtype=torch.float32))
    rhs.copy_(torch.tensor(rng.normal(size=(1, n))))
    s = torch.cuda.Stream()
    s.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(s):
        solver.factorize(A, n)  # warm-up outside the capture
        solver.solve(rhs, x)
    torch.cuda.current_stream().wait_stream(s)
    torch.cuda.synchronize()
    g = torch.cuda.CUDAGraph()
    with torch.cuda.graph(g):
        solver.factorize(A, n)
        solver.solve(rhs, x)
    for rep in range(3):
        An = _spd(n, rng)
        bn = rng.normal(size=n)
        A.copy_(torch.tensor(An[None], dtype=torch.float32))
        rhs.copy_(torch.tensor(bn[None]))
        g.replay()
        torch.cuda.synchronize()
        assert _rel(x.cpu().numpy()[0], np.linalg.solve(_f32(An), bn)) <= 1e-10, rep
        assert solver.info()[0] == 0
    solver.close()

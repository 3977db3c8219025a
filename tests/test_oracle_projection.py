"""Pin the oracle's SAP projection, PD and joint-limit terms.

* The projection is the R-norm projection of y = R^-1 (vhat - vc) onto the
  friction cone (Castro et al. 2022; ``contact_solve.py:743-921``).  It is
  checked against an independent constrained minimiser (scipy SLSQP) and the
  closed-form region tests (stiction / sliding / separation, frictionless).
* The contact cost l(vc) = 1/2 gamma^T R gamma is the potential of the impulse:
  dl/dvc = -gamma, and d(-gamma)/dvc = G (the Hessian block) — finite differences.
* PD and limit terms (``contact_solve.py:490-741``, ``sap_helpers.py:2251-2340``):
  cost -> impulse -> Hessian by finite differences, the implicit-PD closed form
  and the near-rigid gain cap invariants.
"""
from __future__ import annotations

import math

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import sap_oracle as so


def _random_contact(rng):
    rt = 10 ** rng.uniform(-6, -2)
    rn = 10 ** rng.uniform(-5, -1)
    vhat = rng.normal(scale=0.3)
    mu = rng.uniform(0.2, 1.5)
    vc = rng.normal(scale=0.5, size=3)
    return vc, rt, rn, vhat, mu


def _cone_projection_reference(y, rt, rn, mu):
    """argmin_{gamma in K} 1/2 (gamma - y)^T R (gamma - y) with K = {|gamma_t| <= mu gamma_n}."""
    R = np.array([rt, rt, rn])
    scale = np.sqrt(R)
    # minimise in scaled coordinates for conditioning; start from the feasible apex
    f = lambda z: 0.5 * float(np.sum((z - scale * y) ** 2))  # noqa: E731
    cons = [dict(type="ineq", fun=lambda z: mu * z[2] / scale[2] - math.hypot(z[0] / scale[0], z[1] / scale[1]))]
    best = None
    for x0 in (np.zeros(3), scale * np.array([0.0, 0.0, max(y[2], 0.0) + 1e-3]), scale * y):
        res = minimize(f, x0, constraints=cons, method="SLSQP", options=dict(ftol=1e-16, maxiter=500))
        if best is None or res.fun < best.fun:
            best = res
    return best.x / scale


@pytest.mark.parametrize("seed", range(40))
def test_projection_matches_constrained_minimiser(seed):
    rng = np.random.default_rng(seed)
    vc, rt, rn, vhat, mu = _random_contact(rng)
    gamma, G, cost, mode = so.project(vc, rt, rn, vhat, mu)
    y = np.array([-vc[0] / rt, -vc[1] / rt, (vhat - vc[2]) / rn])
    ref = _cone_projection_reference(y, rt, rn, mu)
    scale = max(1.0, float(np.max(np.abs(y))))
    np.testing.assert_allclose(gamma, ref, atol=2e-6 * scale)
    # in the cone, impulse cost consistent
    assert gamma[2] >= 0.0
    assert math.hypot(gamma[0], gamma[1]) <= mu * gamma[2] * (1 + 1e-9) + 1e-300
    assert abs(cost - 0.5 * (rt * (gamma[0] ** 2 + gamma[1] ** 2) + rn * gamma[2] ** 2)) <= 1e-12 * max(cost, 1e-300)


def test_projection_regions():
    rt, rn, mu = 1e-3, 1e-4, 0.8
    # stiction: |y_t| <= mu y_n -> gamma = y
    vc = np.array([1e-5, -2e-5, -0.01])
    g, _, _, mode = so.project(vc, rt, rn, 0.0, mu)
    np.testing.assert_allclose(g, [-vc[0] / rt, -vc[1] / rt, -vc[2] / rn], rtol=1e-12)
    assert mode == so.MODE_STICTION
    # separation: moving apart faster than the cone's polar -> zero impulse
    g, G, cost, mode = so.project(np.array([0.3, 0.0, 0.5]), rt, rn, 0.0, mu)
    assert np.all(g == 0.0) and np.all(G == 0.0) and cost == 0.0 and mode == so.MODE_NONE
    # sliding: on the cone boundary, tangential impulse opposes the slip direction
    vc = np.array([0.2, -0.1, -0.001])
    g, _, _, mode = so.project(vc, rt, rn, 0.0, mu)
    assert mode == so.MODE_SLIDING
    assert abs(math.hypot(g[0], g[1]) - mu * g[2]) < 1e-9 * g[2]
    t = g[:2] / np.linalg.norm(g[:2])
    np.testing.assert_allclose(t, -vc[:2] / np.linalg.norm(vc[:2]), atol=1e-6)
    # frictionless: only the normal component survives
    g, G, _, mode = so.project(np.array([0.3, 0.1, -0.01]), rt, rn, 0.0, 0.0)
    assert g[0] == 0.0 and g[1] == 0.0 and abs(g[2] - 0.01 / rn) < 1e-9 and mode == so.MODE_FRICTIONLESS


def test_projection_is_continuous_across_regions():
    rt, rn, mu = 1e-3, 2e-4, 0.6
    rng = np.random.default_rng(1)
    for _ in range(200):
        vc = rng.normal(scale=0.05, size=3)
        g1 = so.project(vc, rt, rn, 0.0, mu)[0]
        g2 = so.project(vc + 1e-9 * rng.normal(size=3), rt, rn, 0.0, mu)[0]
        assert np.max(np.abs(g1 - g2)) < 1e-9 / min(rt, rn) * 10


@pytest.mark.parametrize("seed", range(40))
def test_cost_gradient_and_hessian_by_finite_differences(seed):
    rng = np.random.default_rng(1000 + seed)
    vc, rt, rn, vhat, mu = _random_contact(rng)
    gamma, G, cost, mode = so.project(vc, rt, rn, vhat, mu)
    h = 1e-7 * max(1.0, float(np.max(np.abs(vc))))
    grad = np.zeros(3)
    dgam = np.zeros((3, 3))
    for k in range(3):
        e = np.zeros(3)
        e[k] = h
        gp, _, cp, mp = so.project(vc + e, rt, rn, vhat, mu)
        gm, _, cm, mm = so.project(vc - e, rt, rn, vhat, mu)
        if mp != mode or mm != mode:
            pytest.skip("sample sits on a region boundary")
        grad[k] = (cp - cm) / (2 * h)
        dgam[:, k] = (gp - gm) / (2 * h)
    gs = max(1.0, float(np.max(np.abs(gamma))))
    np.testing.assert_allclose(grad, -gamma, atol=1e-5 * gs, rtol=1e-5)
    Gs = max(1.0, float(np.max(np.abs(G))))
    np.testing.assert_allclose(-dgam, G, atol=1e-5 * Gs, rtol=1e-4)


def _model_terms(rng, n=6, dt=2e-3):
    mt = so.ModelTerms(*[np.zeros(n) for _ in range(10)])
    mt.pd_active = rng.integers(0, 2, size=n)
    mt.pd_a = rng.normal(scale=5.0, size=n)
    mt.pd_gain = 10 ** rng.uniform(0, 3, size=n)
    mt.pd_limit = np.where(rng.uniform(size=n) < 0.5, rng.uniform(0.5, 20.0, size=n), np.inf)
    mt.lo_active = rng.integers(0, 2, size=n)
    mt.up_active = rng.integers(0, 2, size=n)
    mt.lo_vhat = rng.normal(scale=0.5, size=n)
    mt.up_vhat = rng.normal(scale=0.5, size=n)
    mt.lo_r = 10 ** rng.uniform(-4, -1, size=n)
    mt.up_r = 10 ** rng.uniform(-4, -1, size=n)
    return mt, dt


@pytest.mark.parametrize("seed", range(20))
def test_model_terms_cost_impulse_hessian(seed):
    rng = np.random.default_rng(seed)
    mt, dt = _model_terms(rng)
    n = len(mt.pd_a)
    v = rng.normal(scale=0.5, size=n)
    imp, hd, cost = so.eval_model_terms(mt, v, dt)
    h = 1e-7
    for i in range(n):
        e = np.zeros(n)
        e[i] = h
        ip, _, cp = so.eval_model_terms(mt, v + e, dt)
        im, _, cm = so.eval_model_terms(mt, v - e, dt)
        np.testing.assert_allclose((cp - cm) / (2 * h), -imp[i], rtol=1e-5, atol=1e-6 * max(1.0, abs(imp[i])))
        d = -(ip[i] - im[i]) / (2 * h)
        if abs(d - hd[i]) > 1e-4 * max(1.0, hd[i]):
            # a kink (clamp edge / limit activation) between v-h and v+h is the only excuse
            assert not np.isclose(ip[i] - imp[i], imp[i] - im[i], rtol=1e-3), (i, d, hd[i])


def test_pd_impulse_is_implicit_pd_force():
    """Unsaturated PD impulse = dt [kp (q_t - (q0 + dt v)) + kd (v_t - v) + u0] (implicit PD)."""
    from types import SimpleNamespace

    kp, kd, dt = 300.0, 12.0, 2e-3
    q0, qt, vt, u0 = 0.3, -0.1, 0.4, 1.5
    model = SimpleNamespace(joint_target_mode=np.array([1]), joint_target_ke=np.array([kp]),
                            joint_target_kd=np.array([kd]), joint_effort_limit=np.array([np.inf]),
                            joint_limit_ke=np.array([0.0]), joint_limit_kd=np.array([0.0]),
                            joint_limit_lower=np.array([-1e10]), joint_limit_upper=np.array([1e10]),
                            joint_count=1, joint_type=np.array([so.JOINT_REVOLUTE]), joint_qd_start=np.array([0, 1]),
                            joint_q_start=np.array([0, 1]), joint_dof_dim=np.array([[0, 1]]), joint_dof_count=1)
    wv = SimpleNamespace(n=1, d0=0, d1=1)
    A = np.array([[50.0]])  # heavy enough that the near-rigid cap does not engage
    ctl = dict(joint_target_pos=np.array([qt]), joint_target_vel=np.array([vt]), joint_act=np.array([u0]))
    mt = so.build_model_terms(model, wv, np.array([q0]), ctl, A, np.zeros(1), np.zeros(1), dt)
    kp_e, kd_e = so.effective_pd_gains(kp, kd, dt, 1.0 / A[0, 0])
    assert (kp_e, kd_e) == (kp, kd)
    for v in (-1.0, 0.0, 0.7):
        imp, hd, _ = so.eval_model_terms(mt, np.array([v]), dt)
        expect = dt * (kp * (qt - (q0 + dt * v)) + kd * (vt - v) + u0)
        assert abs(imp[0] - expect) < 1e-12 * max(1.0, abs(expect))
        assert abs(hd[0] - dt * (dt * kp + kd)) < 1e-12


@pytest.mark.parametrize("seed", range(50))
def test_effective_pd_gains_cap(seed):
    rng = np.random.default_rng(seed)
    kp = 10 ** rng.uniform(-1, 7) * (rng.uniform() > 0.1)
    kd = 10 ** rng.uniform(-2, 4) * (rng.uniform() > 0.1)
    dt = 10 ** rng.uniform(-4, -2)
    delassus = 10 ** rng.uniform(-3, 3)
    kp_e, kd_e = so.effective_pd_gains(kp, kd, dt, delassus)
    if kp <= 0 and kd <= 0:
        assert (kp_e, kd_e) == (0.0, 0.0)
        return
    r_nr = so.PD_BETA ** 2 / (4 * math.pi ** 2) * delassus
    r_eff = 1.0 / (dt * (dt * kp_e + kd_e))
    r_in = 1.0 / (dt * (dt * kp + kd))
    if r_in >= r_nr:
        assert (kp_e, kd_e) == (kp, kd)
    else:
        # capped: the regulariser sits exactly at the near-rigid bound, the time constant is kept
        assert abs(r_eff - r_nr) <= 1e-9 * r_nr
        if kp > 0 and kd > 0:
            assert abs(kd_e * kp - kp_e * kd) <= 1e-9 * kp_e * kd

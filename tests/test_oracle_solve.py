"""Pin the oracle's SAP Newton solve, line searches and contact model end to end.

* The SAP objective is strictly convex (A is SPD), so its minimiser is unique:
  the oracle's Newton + line-search result must match an independent
  trust-region minimiser (scipy ``trust-exact``) run to a tighter tolerance.
* The objective's gradient (A(v - v*) - J^T gamma - model impulses) is the
  finite-difference gradient of its cost.
* Exact-root line search: phi'(alpha) = 0 at the accepted step (or alpha_max
  with phi' <= 0); Armijo: sufficient decrease (``contact_solve.py:2670-3140``).
* Physics closed forms through the whole step (collision oracle + step): a
  resting sphere's normal impulse equals m g dt; static friction balances a
  push below mu m g; a frictionless box accelerates at F / m.  (Fast sliding
  is not a closed form here: the convex SAP model lifts a sliding body by
  about mu |v_t| per unit time, Anitescu's relaxation.)

Reference: ``sim/contact_solve.py:1493-1627, 1970-2087, 2234-3176, 6129-6285``.
"""
from __future__ import annotations

import numpy as np
import pytest
from scipy.optimize import minimize

from oracle import collide as oc
from oracle import sap_oracle as so
from sap_warp_amd.runtime import GEO_BOX, GEO_SPHERE
from sap_warp_amd.scenes import SceneBuilder, ShapeCfg, build_ant_scene, build_humanoid_scene, tf
from tests.helpers import control_dict_from_model, state_dict_from_model


def _world_problem(model, w, dt, C=64, preset="approx32", seed_state=None):
    st = state_dict_from_model(model) if seed_state is None else seed_state
    ctl = control_dict_from_model(model)
    contacts = oc.collide(model, st["body_q"], C)
    wv = so.WorldView(model, w)
    jq = st["joint_q"].astype(np.float64)
    v0 = so.public_to_sap_velocity(model, wv, st["body_q"], st["joint_qd"])
    f_sap = so.public_to_sap_force(model, wv, st["body_q"], ctl["joint_f"])
    bf = so.public_body_force_to_sap(model, wv, st["body_q"], st.get("body_f"))
    fm = so.free_motion(model, wv, jq[wv.c0:wv.c1], v0, f_sap, bf, dt)
    weight, point, _ = so.StepOptions(preset=preset).resolved()
    cs = so.world_contacts(model, wv, fm, contacts, max_contacts=C, contact_point_mode=point,
                           contact_weight_mode=weight, fallback_mu=so.infer_mu_fallback(model), fallback_k=1e10,
                           fallback_tau=so.infer_tau_fallback(model))
    mt = so.build_model_terms(model, wv, jq, ctl, fm.A_env, v0, fm.v_star, dt)
    return wv, fm, cs, mt, v0


def _cases():
    ant = build_ant_scene(num_worlds=4, seed=3)
    hum = build_humanoid_scene(num_worlds=3, seed=4)
    out = []
    for w in range(4):
        out.append(("ant", ant, w, 1e-3))
    for w in range(3):
        out.append(("humanoid", hum, w, 2e-3))
    return out


CASES = _cases()


@pytest.mark.parametrize("case", range(len(CASES)))
@pytest.mark.parametrize("ls", ["armijo_decay", "exact_root", "monotone_decay"])
def test_solution_is_the_unique_minimiser(case, ls):
    name, model, w, dt = CASES[case]
    wv, fm, cs, mt, v0 = _world_problem(model, w, dt)
    opts = so.SolveOptions.for_variant(ls)
    res = so.sap_solve(model, wv, fm.A_env, fm.v_star, v0, None, cs, mt, dt, opts)
    assert res.converged
    if cs.count == 0 and not np.any(mt.pd_active | mt.lo_active | mt.up_active):
        np.testing.assert_array_equal(res.v, fm.v_star)
        return
    prob = so.SapProblem(fm.A_env, fm.v_star, cs, mt, dt)
    ev0 = prob.evaluate(res.v)
    # independent minimiser: trust region on the same strictly convex objective
    ref = minimize(lambda x: prob.evaluate(x)["cost"], fm.v_star.copy(), jac=lambda x: prob.evaluate(x)["grad"],
                   hess=lambda x: prob.hessian(prob.evaluate(x)), method="trust-exact",
                   options=dict(gtol=1e-12 * max(1.0, float(np.max(np.abs(ev0["Av"])))), maxiter=500))
    # strong convexity (modulus lambda_min(A)) bounds the distance to the minimiser by the gradient
    lam = float(np.min(np.linalg.eigvalsh(fm.A_env)))
    g = float(np.linalg.norm(ev0["grad"]))
    assert float(np.linalg.norm(res.v - ref.x)) <= g / lam * (1 + 1e-6) + 1e-10, (name, w)
    # monotone_decay stops on a loose cost criterion by default (relative cost change per iteration
    # below 5e-3, contact_solve.py:6179-6181): only the strong-convexity bound above and the tight
    # solve below apply to it
    if ls != "monotone_decay":
        assert res.cost <= prob.cost(ref.x) + 1e-9 * max(1.0, abs(res.cost))
    # driven to a tight tolerance, the same Newton + line search lands on the minimiser
    tight = so.SolveOptions.for_variant(ls, optimality_rel_tol=1e-13, optimality_abs_tol=1e-15, cost_rel_tol=0.0)
    rt = so.sap_solve(model, wv, fm.A_env, fm.v_star, v0, None, cs, mt, dt, tight)
    scale = max(1.0, float(np.max(np.abs(ref.x))))
    assert np.max(np.abs(rt.v - ref.x)) <= 1e-9 * scale, (name, w, np.max(np.abs(rt.v - ref.x)))


@pytest.mark.parametrize("case", [0, 4, 5])
def test_objective_gradient_is_cost_derivative(case):
    name, model, w, dt = CASES[case]
    wv, fm, cs, mt, v0 = _world_problem(model, w, dt)
    prob = so.SapProblem(fm.A_env, fm.v_star, cs, mt, dt)
    rng = np.random.default_rng(case)
    v = fm.v_star + rng.normal(scale=0.05, size=wv.n)
    ev = prob.evaluate(v)
    h = 1e-6
    fd = np.array([(prob.cost(v + h * e) - prob.cost(v - h * e)) / (2 * h) for e in np.eye(wv.n)])
    np.testing.assert_allclose(fd, ev["grad"], rtol=1e-5, atol=1e-6 * max(1.0, float(np.max(np.abs(ev["grad"])))))
    # Hessian: finite differences of the gradient
    Hfd = np.column_stack([(prob.evaluate(v + h * e)["grad"] - prob.evaluate(v - h * e)["grad"]) / (2 * h)
                           for e in np.eye(wv.n)])
    H = prob.hessian(ev)
    np.testing.assert_allclose(Hfd, H, rtol=1e-4, atol=1e-5 * float(np.max(np.abs(H))))


@pytest.mark.parametrize("case", range(len(CASES)))
def test_line_search_conditions(case):
    name, model, w, dt = CASES[case]
    wv, fm, cs, mt, v0 = _world_problem(model, w, dt)
    if cs.count == 0:
        pytest.skip("no contacts")
    prob = so.SapProblem(fm.A_env, fm.v_star, cs, mt, dt)
    v = v0.copy()
    ev = prob.evaluate(v)
    dv = np.linalg.solve(prob.hessian(ev), -ev["grad"])
    d0 = float(ev["grad"] @ dv)
    assert d0 < 0.0
    # exact root: phi'(alpha) = grad(v + alpha dv) . dv vanishes (or alpha_max with phi' <= 0)
    eo = so.SolveOptions.for_variant("exact_root")
    ok, a, _, status = so.exact_line_search(prob, v, dv, ev, eo)
    assert ok and status == 0
    d_a = float(prob.evaluate(v + a * dv)["grad"] @ dv)
    if a < so.EXACT_ALPHA_MAX:
        assert abs(d_a) <= 1e-6 * abs(d0)
    else:
        assert d_a <= 0.0
    # armijo: sufficient decrease
    ao = so.SolveOptions.for_variant("armijo_decay")
    ok, a, _, status = so.armijo_line_search(prob, v, dv, ev, ao)
    assert ok
    assert prob.cost(v + a * dv) <= ev["cost"] + ao.armijo_c * a * d0 + 1e-12 * abs(ev["cost"])


@pytest.mark.parametrize("case", range(len(CASES)))
def test_monotone_decay_never_raises_the_cost(case):
    """monotone_decay: alpha in {1, 1/2, 1/4, ...}, the accepted cost never exceeds the current one
    (relaxed by 1e-14 + 1e-12 |cost|), and the first halving that decreases the cost is taken."""
    name, model, w, dt = CASES[case]
    wv, fm, cs, mt, v0 = _world_problem(model, w, dt)
    if cs.count == 0 and not np.any(mt.pd_active | mt.lo_active | mt.up_active):
        pytest.skip("unconstrained")
    opts = so.SolveOptions.for_variant("monotone_decay")
    res = so.sap_solve(model, wv, fm.A_env, fm.v_star, v0, None, cs, mt, dt, opts)
    assert res.converged
    costs = [h["cost"] for h in res.history]
    for a, b in zip(costs, costs[1:]):
        assert b <= a + 1e-14 + 1e-12 * abs(a)
    # one search by hand: the accepted alpha is the first power of 1/2 that does not raise the cost
    prob = so.SapProblem(fm.A_env, fm.v_star, cs, mt, dt)
    ev = prob.evaluate(v0)
    dv = np.linalg.solve(prob.hessian(ev), -ev["grad"])
    ok, a, trials, status = so.monotone_line_search(prob, v0, dv, ev, opts)
    tol = 1e-14 + 1e-12 * abs(ev["cost"])
    k = next(k for k in range(40) if prob.cost(v0 + 0.5 ** k * dv) <= ev["cost"] + tol)
    assert ok and status == 0 and a == 0.5 ** k and trials == k + 1


# ----------------------------------------------------------------------------
# physics through the whole step
# ----------------------------------------------------------------------------
def _single_body_scene(shape_type, scale, z, mu, vx=0.0):
    b = SceneBuilder(rigid_gap=0.01)
    b.current_world = 0
    cfg = ShapeCfg(density=1000.0, ke=2.0e5, tau=0.02, mu=mu, gap=0.01)
    body = b.add_body(tf((0.0, 0.0, z)))
    b.add_shape(body, shape_type, tf(), scale, cfg)
    j = b.add_free_joint(body)
    b.add_articulation([j])
    b.joint_qd[b.joint_qd_start[j]] = vx
    b.current_world = -1
    b.add_ground_plane(ShapeCfg(density=0.0, ke=2.0e5, tau=0.02, mu=mu, gap=0.01, collision_group=-1))
    from sap_warp_amd.scenes import _with_consistent_body_q

    return _with_consistent_body_q(b.finalize())


def _simulate(model, steps, dt, opts, body_f=None):
    st = state_dict_from_model(model)
    if body_f is not None:
        st["body_f"] = np.asarray(body_f, dtype=np.float32).reshape(-1, 6)
    ctl = control_dict_from_model(model)
    guess = None
    hist = []
    for _ in range(steps):
        c = oc.collide(model, st["body_q"], 16)
        out = so.step(model, st, ctl, c, dt, opts, v_guess=guess)
        r = out["per_world"][0]
        hist.append(dict(v=out["v_sap"].copy(), gamma=r["result"].gamma.copy(), nc=r["contacts"].count,
                         qd=out["joint_qd"].copy()))
        guess = out["v_sap"]
        st = dict(joint_q=out["joint_q"].astype(np.float32), joint_qd=out["joint_qd"].astype(np.float32),
                  body_q=out["body_q"].astype(np.float32), body_qd=out["body_qd"].astype(np.float32),
                  body_f=st["body_f"])
    return hist


def test_resting_sphere_carries_its_weight():
    r, dt = 0.1, 1e-3
    model = _single_body_scene(GEO_SPHERE, (r, 0.0, 0.0), r, mu=0.5)
    m = float(model.body_mass[0])
    hist = _simulate(model, 400, dt, so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=8))
    last = hist[-1]
    assert last["nc"] == 1
    gn = float(np.sum(last["gamma"][:, 2]))
    assert abs(gn - m * 9.81 * dt) <= 1e-3 * m * 9.81 * dt, (gn, m * 9.81 * dt)
    assert np.max(np.abs(last["gamma"][:, :2])) <= 1e-6 * gn
    assert np.max(np.abs(last["v"])) < 1e-4


def test_static_friction_holds_a_pushed_box():
    """A horizontal push below mu m g is balanced by friction: the box stays, sum gamma_t = -F dt."""
    h, dt, mu = 0.05, 1e-3, 0.6
    model = _single_body_scene(GEO_BOX, (h, h, h), h, mu=mu)
    m = float(model.body_mass[0])
    F = 0.5 * mu * m * 9.81
    hist = _simulate(model, 300, dt, so.StepOptions(preset="drake", line_search="armijo_decay", max_rigid_contact=8),
                     body_f=[F, 0, 0, 0, 0, 0])
    last = hist[-1]
    assert abs(last["qd"][0]) < 1e-3
    # frame of a +z normal: cx = -y, cy = +x, so the world-x friction is the cy component
    ft_x = float(np.sum(last["gamma"][:, 1]))
    assert abs(ft_x + F * dt) <= 2e-3 * F * dt, (ft_x, F * dt)


def test_frictionless_box_accelerates_at_f_over_m():
    h, dt = 0.05, 1e-3
    model = _single_body_scene(GEO_BOX, (h, h, h), h, mu=0.0)
    m = float(model.body_mass[0])
    F = 2.0
    hist = _simulate(model, 200, dt, so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=8),
                     body_f=[F, 0, 0, 0, 0, 0])
    vx = np.array([s["qd"][0] for s in hist])
    np.testing.assert_allclose(vx, F / m * dt * np.arange(1, 201), rtol=1e-5)
    assert np.max(np.abs(np.array([s["gamma"][:, :2] for s in hist]))) == 0.0


def test_optimality_on_solver_data():
    """optimality_on_data rebuilds the problem from data formed elsewhere: fed the oracle's own A,
    v*, J, weights and phi0 it gives optimality_at's norms; a v optimal for a perturbed problem
    (A scaled by 1 + 1e-4, as a float32-formed A would differ, only larger) passes on that problem
    and fails on the oracle's."""
    model = build_humanoid_scene(num_worlds=2, seed=4)
    st, ctl = state_dict_from_model(model), control_dict_from_model(model)
    contacts = oc.collide(model, st["body_q"], 64)
    dt = 2.0e-3
    opts = so.SolveOptions.for_variant("exact_root")
    for w in range(2):
        r = so.step_world(model, w, st, ctl, contacts, dt, so.StepOptions(line_search="exact_root"))
        fm, cs, res = r["free_motion"], r["contacts"], r["result"]
        assert res.iterations > 0
        gn, tol = so.optimality_on_data(model, w, r, res.v, dt, opts, A=fm.A_env, v_star=fm.v_star, J=cs.J,
                                        weight=cs.w, phi0=cs.phi0)
        gn0, tol0, _, _ = so.optimality_at(model, w, r, res.v, dt, opts)
        assert gn == pytest.approx(gn0, rel=1e-12) and tol == pytest.approx(tol0, rel=1e-12)
        assert gn <= tol
        # the optimum of a perturbed problem: optimal there, not on the original
        A2 = fm.A_env * (1.0 + 1.0e-4)
        wv = so.WorldView(model, w)
        mt2 = so.build_model_terms(model, wv, r["joint_q_in"], ctl, A2, r["v0"], fm.v_star, dt)
        tight = so.SolveOptions.for_variant("exact_root", optimality_rel_tol=1e-12, cost_rel_tol=0.0)
        r2 = so.sap_solve(model, wv, A2, fm.v_star, r["v0"], None, cs, mt2, dt, tight)
        gn2, tol2 = so.optimality_on_data(model, w, r, r2.v, dt, opts, A=A2, v_star=fm.v_star, J=cs.J,
                                          weight=cs.w, phi0=cs.phi0)
        assert gn2 <= tol2
        gn3, tol3, _, _ = so.optimality_at(model, w, r, r2.v, dt, opts)
        assert gn3 > 2.0 * tol3


def test_solve_on_data_reproduces_own_solve():
    """solve_on_data fed the oracle's own problem data returns the oracle's own solution."""
    model = build_humanoid_scene(num_worlds=2, seed=4)
    st, ctl = state_dict_from_model(model), control_dict_from_model(model)
    contacts = oc.collide(model, st["body_q"], 64)
    dt = 2.0e-3
    opts = so.SolveOptions.for_variant("exact_root")
    for w in range(2):
        r = so.step_world(model, w, st, ctl, contacts, dt, so.StepOptions(line_search="exact_root"))
        fm, cs = r["free_motion"], r["contacts"]
        res = so.solve_on_data(model, w, r, dt, opts, A=fm.A_env, v_star=fm.v_star, J=cs.J, weight=cs.w,
                               phi0=cs.phi0)
        np.testing.assert_array_equal(res.v, r["result"].v)
        np.testing.assert_array_equal(res.gamma, r["result"].gamma)

"""Learned warm start on the GPU vs the CPU oracle (oracle/learned_oracle.py).

* MLP forward, gradients and the Adam step of the HIP kernels vs the numpy
  restatement with the same bfloat16 rounding points (batch not a multiple of
  the 64-row tile, padded input / output widths).
* Features of the feature kernel vs the oracle on humanoid worlds with contacts.
* The training target the step kernel writes (converged v minus the plain start,
  free-joint rows in the body frame) vs the oracle step from the same input.
* A learned start keeps the answer: v within 1e-5 of the oracle step from the
  same start, within the optimality tolerance ball (1e-4) of the plain-start
  answer, every world converged, the cheaper start taken; the exact correction
  takes the Newton count down; a useless correction is rejected (the fallback)
  and the answer is the plain one to 1e-5.
* Training on plain rollouts reduces the mean Newton count on a held-out seed.
"""
from __future__ import annotations

import numpy as np
import pytest

from oracle import learned_oracle as lo
from oracle import sap_oracle as so
from tests.helpers import rel_err

pytestmark = pytest.mark.gpu


def _np(t):
    return t.detach().cpu().numpy()


def _oracle_mlp_from(m, flat, norm):
    o = lo.Mlp(list(m.dims), m.in_features, m.out_features)
    o.W, o.b = lo.unflat(o, flat.astype(np.float64))
    o.x_mean, o.x_std, o.y_mean, o.y_std = [np.asarray(a, np.float64) for a in norm]
    # the kernels normalise in float32: mirror the float32 statistics
    o.x_mean = o.x_mean.astype(np.float32).astype(np.float64)
    o.x_std = o.x_std.astype(np.float32).astype(np.float64)
    return o


def test_mlp_forward_gradients_adam_match_oracle(gpu_device):
    import torch

    from sap_warp_amd.learned import Mlp

    rng = np.random.default_rng(21)
    Fi, Fo, B = 50, 20, 100
    m = Mlp(Fi, Fo, hidden=(128, 128), seed=3, out_scale=1.0)
    flat = m.read()
    # non-zero biases (zero on padded outputs)
    o = lo.Mlp(list(m.dims), Fi, Fo)
    Ws, bs = lo.unflat(o, flat.astype(np.float64))
    for l, b in enumerate(bs):
        b[:] = rng.standard_normal(b.shape) * 0.1
        if l == len(bs) - 1:
            b[Fo:] = 0.0
    flat = np.concatenate([np.concatenate([w.ravel(), b]) for w, b in zip(Ws, bs)]).astype(np.float32)
    m.write_params(flat)
    norm = (rng.standard_normal(Fi), 0.5 + rng.random(Fi), rng.standard_normal(Fo), 0.5 + rng.random(Fo))
    m.set_normalization(*norm)
    x = rng.standard_normal((B, Fi)).astype(np.float32)
    t = rng.standard_normal((B, Fo)).astype(np.float32)
    xd = torch.as_tensor(x, device="cuda:0")
    td = torch.as_tensor(t, device="cuda:0")
    y = torch.zeros((B, Fo), dtype=torch.float32, device="cuda:0")
    m.forward(xd, y)
    ref = _oracle_mlp_from(m, flat, norm)
    y_ref = ref.predict(x.astype(np.float64))
    err = np.max(np.abs(_np(y) - y_ref)) / max(1.0, np.max(np.abs(y_ref)))
    print("forward rel err", err)
    assert err < 2e-3, err
    loss = torch.zeros(1, dtype=torch.float32, device="cuda:0")
    m.gradients(xd, td, loss)
    from sap_warp_amd import _native as nat

    g = m.read(nat.MLP_GRADS)
    l_ref, gW, gb = ref.loss_and_grads(x.astype(np.float64), t.astype(np.float64))
    assert abs(float(loss.item()) - l_ref) <= 1e-3 * l_ref
    gW_gpu, gb_gpu = lo.unflat(ref, g.astype(np.float64))
    for l in range(len(gW)):
        ew = np.linalg.norm(gW_gpu[l] - gW[l]) / max(np.linalg.norm(gW[l]), 1e-12)
        eb = np.linalg.norm(gb_gpu[l] - gb[l]) / max(np.linalg.norm(gb[l]), 1e-12)
        print("layer", l, "grad rel err W", ew, "b", eb)
        assert ew < 1e-2 and eb < 1e-2, (l, ew, eb)
    # one Adam step from the GPU's own gradients vs the oracle's Adam
    m.adam(1e-3)
    after = m.read()
    opt = lo.Adam(lr=1e-3)
    p = [flat.astype(np.float64).copy()]
    opt.step(p, [g.astype(np.float64)])
    np.testing.assert_allclose(after, p[0], rtol=1e-5, atol=1e-7)


def _scene(W, seed):
    from sap_warp_amd.scenes import build_humanoid_scene

    return build_humanoid_scene(W, seed=seed, device="cuda:0")


def _state_np(s):
    return dict(joint_q=_np(s.joint_q), joint_qd=_np(s.joint_qd), body_q=_np(s.body_q), body_qd=_np(s.body_qd),
                body_f=_np(s.body_f))


def _contacts_np(c):
    return dict(count=int(_np(c.rigid_contact_count)[0]), world_start=_np(c.rigid_contact_world_start),
                shape0=_np(c.rigid_contact_shape0), shape1=_np(c.rigid_contact_shape1),
                point0=_np(c.rigid_contact_point0), point1=_np(c.rigid_contact_point1),
                normal=_np(c.rigid_contact_normal), margin0=_np(c.rigid_contact_margin0),
                margin1=_np(c.rigid_contact_margin1))


def _control_np(c):
    return dict(joint_f=_np(c.joint_f), joint_target_pos=_np(c.joint_target_pos),
                joint_target_vel=_np(c.joint_target_vel), joint_act=_np(c.joint_act))


def _setup(W=6, seed=0, C=64):
    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.learned import LearnedWarmStart
    from sap_warp_amd.solver_sap import SolverSAP

    model = _scene(W, seed)
    solver = SolverSAP(model, max_rigid_contact=C, line_search_variant="exact_root",
                       capture_contact_jacobian_snapshots=True, collect_iteration_stats=True)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    lws = LearnedWarmStart(solver, hidden=(128, 128), seed=1)
    return model, solver, pipe, lws


def test_features_match_oracle(gpu_device):
    import torch

    model, solver, pipe, lws = _setup(W=6, seed=2)
    contacts = pipe.contacts()
    s0 = model.state()
    ctl = model.control()
    ctl.joint_target_pos.add_(0.1)
    pipe.collide(s0, contacts)
    f = _np(lws.compute_features(s0, ctl, contacts).clone())
    torch.cuda.synchronize()
    st, cn, cc = _state_np(s0), _contacts_np(contacts), _control_np(ctl)
    assert cn["count"] > 0
    for w in range(model.world_count):
        ref = lo.features(model, w, st, cc, cn)
        assert ref.shape == (lws.F,)
        np.testing.assert_allclose(f[w], ref, atol=2e-5, rtol=1e-5)


def _oracle_step(model, st_in, ctl, contacts, dt, v_guess):
    opts = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=64)
    return so.step(model, st_in, ctl, contacts, dt, opts, v_guess=v_guess)


def _plain_start(model, w, r, v_guess_w):
    """The solver's plain start of world w (contact_solve.py:316-393): guess or v0, v* where not participating."""
    wv = so.WorldView(model, w)
    cs, mt, vs = r["contacts"], r["model_terms"], r["v_star"]
    unconstrained = cs.count == 0 and not np.any((mt.pd_active == 1) | (mt.lo_active == 1) | (mt.up_active == 1))
    if unconstrained:
        return vs.copy()
    v = np.array(v_guess_w if v_guess_w is not None else r["v0"], dtype=np.float64)
    part = so._participating(model, wv, cs, mt)
    v[part == 0] = vs[part == 0]
    return v


def test_target_and_learned_start_parity(gpu_device):
    import torch

    model, solver, pipe, lws = _setup(W=6, seed=0)
    contacts = pipe.contacts()
    s0, s1 = model.state(), model.state()
    ctl = model.control()
    dt = 2.0e-3
    n = model.dof_per_env
    # two plain steps so that the plain start of the next one is the previous solution
    v_guess = None
    for _ in range(2):
        pipe.collide(s0, contacts)
        st_in, cn = _state_np(s0), _contacts_np(contacts)
        lws.plain_step_with_target(s0, s1, ctl, contacts, dt)
        torch.cuda.synchronize()
        ref = _oracle_step(model, st_in, _control_np(ctl), cn, dt, v_guess)
        tgt = _np(lws.last_target)
        for w in range(model.world_count):
            r = ref["per_world"][w]
            vp = _plain_start(model, w, r, None if v_guess is None else v_guess[w * n:(w + 1) * n])
            want = lo.local_delta(model, w, st_in["joint_q"], r["v_sap"] - vp)
            assert rel_err(tgt[w], want) < 2e-5, (w, rel_err(tgt[w], want))
        v_guess = _np(solver.read_buffer(0)).astype(np.float64)
        s0, s1 = s1, s0
    # learned start = the correction to this step's optimum (the oracle solved to a 1e-12 tolerance:
    # a start well inside the stopping tolerance, not on its boundary)
    pipe.collide(s0, contacts)
    st_in, cn, cc = _state_np(s0), _contacts_np(contacts), _control_np(ctl)
    ref = _oracle_step(model, st_in, cc, cn, dt, v_guess)
    tight = so.StepOptions(preset="approx32", line_search="exact_root", max_rigid_contact=64,
                           solve=so.SolveOptions.for_variant("exact_root", optimality_rel_tol=1e-12,
                                                             max_iterations=200))
    ref_t = so.step(model, st_in, cc, cn, dt, tight, v_guess=v_guess)
    exact = np.zeros((model.world_count, n), np.float32)
    for w in range(model.world_count):
        r = ref["per_world"][w]
        exact[w] = lo.local_delta(model, w, st_in["joint_q"],
                                  ref_t["per_world"][w]["v_sap"] - _plain_start(model, w, r, v_guess[w * n:(w + 1) * n]))
    delta = torch.as_tensor(exact, device="cuda:0")
    solver.iteration_histogram(reset=True)
    lws.set_io(delta, None)
    try:
        solver.step(s0, s1, ctl, contacts, dt)
    finally:
        lws.set_io(None, None)
    torch.cuda.synchronize()
    h = solver.iteration_histogram(reset=True)
    stats = solver.last_contact_solve_result
    v = _np(solver.read_buffer(0))
    # the oracle from the same start (plain start + the float32 correction) follows the same path
    starts = np.concatenate([lo.apply_delta(model, w, st_in["joint_q"],
                                            _plain_start(model, w, ref["per_world"][w], v_guess[w * n:(w + 1) * n]),
                                            exact[w].astype(np.float64)) for w in range(model.world_count)])
    ref_same = _oracle_step(model, st_in, cc, cn, dt, starts)
    constrained = 0
    for w in range(model.world_count):
        r = ref["per_world"][w]
        assert stats.converged_env[w] == 1
        assert ref_same["per_world"][w]["result"].converged
        assert rel_err(v[w * n:(w + 1) * n], ref_same["per_world"][w]["v_sap"]) <= 1e-5, w
        assert rel_err(v[w * n:(w + 1) * n], ref_t["per_world"][w]["v_sap"]) <= 1e-5, w
        # both answers satisfy the optimality condition: they agree to the tolerance ball
        assert rel_err(v[w * n:(w + 1) * n], r["v_sap"]) <= 1e-4, w
        plain_its = r["result"].iterations
        if plain_its > 0:
            constrained += 1
            assert stats.newton_iterations_env[w] <= max(1, plain_its - 1), (w, stats.newton_iterations_env[w], plain_its)
    assert h[258] + h[259] >= constrained
    assert h[258] >= constrained - 1
    # a large random correction is rejected: the plain start stays and the answer is unchanged
    solver.reset_runtime_state()
    bad = torch.as_tensor(np.random.default_rng(4).standard_normal((model.world_count, n)).astype(np.float32) * 50.0,
                          device="cuda:0")
    s2 = model.state()
    pipe.collide(s2, contacts)
    st_in, cn = _state_np(s2), _contacts_np(contacts)
    lws.set_io(bad, None)
    try:
        solver.step(s2, s1, ctl, contacts, dt)
    finally:
        lws.set_io(None, None)
    torch.cuda.synchronize()
    h = solver.iteration_histogram(reset=True)
    ref = _oracle_step(model, st_in, cc, cn, dt, None)
    v = _np(solver.read_buffer(0))
    for w in range(model.world_count):
        assert rel_err(v[w * n:(w + 1) * n], ref["per_world"][w]["v_sap"]) <= 1e-5, w
    assert h[259] >= 1 and h[258] == 0, h[256:]


def test_training_reduces_newton_iterations(gpu_device):
    import torch

    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.learned import LearnedWarmStart
    from sap_warp_amd.solver_sap import SolverSAP

    W, dt, steps = 512, 2.0e-3, 60

    def rollout(seed, lws_mode, lws_obj=None, collect=False):
        model = _scene(W, seed)
        solver = SolverSAP(model, max_rigid_contact=64, line_search_variant="exact_root")
        pipe = SapCollisionPipeline(model, max_contacts_per_world=64)
        contacts = pipe.contacts()
        ctl = model.control()
        s0, s1 = model.state(), model.state()
        gen = torch.Generator(device="cuda:0")
        gen.manual_seed(100 + seed)
        mask = torch.zeros(model.dof_per_env, device="cuda:0")
        mask[6:] = 1.0
        lws = LearnedWarmStart(solver, hidden=(512, 512, 512), seed=0,
                               mlp=None if lws_obj is None else lws_obj.mlp)

        def ctl_fn(k):
            ctl.joint_target_pos.add_((torch.randn((W, model.dof_per_env), generator=gen, device="cuda:0") * 0.05
                                       * mask).reshape(-1))

        if collect:
            x, y, _, _ = lws.collect(pipe, s0, s1, ctl, contacts, steps, dt, control_fn=ctl_fn)
            return lws, x, y
        solver.iteration_histogram(reset=True)
        for k in range(steps):
            ctl_fn(k)
            pipe.collide(s0, contacts)
            if lws_mode:
                lws.step(s0, s1, ctl, contacts, dt)
            else:
                solver.step(s0, s1, ctl, contacts, dt)
            s0, s1 = s1, s0
        h = solver.iteration_histogram(reset=True)
        its = (h[:256] * np.arange(256)).sum() / max(h[257], 1)
        return its, h

    xs, ys = [], []
    lws = None
    for seed in (11, 12, 13):
        l, x, y = rollout(seed, False, collect=True)
        lws = l if lws is None else lws
        xs.append(x)
        ys.append(y)
    x = torch.cat(xs)
    y = torch.cat(ys)
    lws.fit(x, y, epochs=6, batch=2048, lr=1e-3, lr_final=1e-4)
    print("trained", lws.train_steps, "steps, loss", lws.final_loss, "in", round(lws.train_seconds, 2), "s")
    plain, _ = rollout(21, False)
    learned, h = rollout(21, True, lws_obj=lws)
    print("held-out mean Newton iterations: plain", plain, "learned", learned, "taken", h[258], "rejected", h[259])
    assert lws.final_loss < 1.0
    assert learned < plain

#!/usr/bin/env python
"""World-steps/s of the SAP contact step on one MI355X.

Workload: the batched plain-SAP humanoid configuration — 4096 humanoid
worlds (floating base + 29 revolute joints, 35 DOF, 30 bodies, capsule /
sphere / box links), float32 state, dt = 2e-3, PD position targets driven by
a random walk (sigma = 0.05 rad per step), exact-root line search, approx32
preset, 64 contacts per world capacity.  One step = collide (AABB + explicit
pairs + primitive narrowphase) + SolverSAP.step (free motion, contact
Jacobian, Newton solve, integration, public FK) for every world, with all
inputs resident in HBM.

Prints one JSON line (metric / value / roofline / cpu_baseline ...).
"""
from __future__ import annotations

import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
if ROOT not in sys.path:
    sys.path.insert(0, ROOT)

PEAK_FP32_TFLOPS = 157.3   # MI355X dense FP32 (vector and matrix)
PEAK_HBM_GBS = 8000.0      # MI355X HBM3E
STEP_KERNELS = "sapx_wave_kernel<48,16> + sapx_wave_kernel<48,64>"


def step_flops(n, nb, nc, k, t):
    """Algorithmic FP operations of one world-step (free motion + Jacobian + k Newton iterations).

    Per-unit figures: free motion 2*6nb*6n + 2*6nb*n(n+1)/2 + n^3/3 + 2n^2 + 600nb;
    contact Jacobian 12*nc*n; per Newton iteration 3*(2*3nc*n) + 2*3nc*n(n+1)/2
    + 18nc*n + 4n^2 + n^3/3 + 2n^2 + 150nc; per line-search trial 150nc + 10n.
    """
    fm = 2 * 6 * nb * 6 * n + 2 * 6 * nb * n * (n + 1) / 2 + n ** 3 / 3 + 2 * n * n + 600 * nb
    jac = 12 * nc * n
    newton = 3 * (2 * 3 * nc * n) + 2 * 3 * nc * n * (n + 1) / 2 + 18 * nc * n + 4 * n * n + n ** 3 / 3 + 2 * n * n + 150 * nc
    return fm + jac + k * newton + t * (150 * nc + 10 * n)


def collide_bytes(model, contacts_written, S, P):
    """HBM bytes of one collide call as its two kernels lay them out: the collide kernel reads body
    poses, shape data and the pair template and writes each world's staged contacts (52 bytes each)
    and count; the placement kernel reads the counts and staged contacts and writes the final
    contacts and world_start."""
    W = model.world_count
    body = model.body_count * 28
    shapes = W * S * (28 + 12 + 4 * 5 + 4)
    pairs = P * 8
    per_contact = 4 * 2 + 12 * 3 + 4 * 2
    out = contacts_written * per_contact * 3 + W * 4 * 2 + (W + 1) * 4
    return body + shapes + pairs + out


def read_traffic():
    """HBM bytes per step-kernel launch from the committed rocprofv3 PMC summary (None if absent)."""
    path = os.path.join(ROOT, "profiles", "sapx_step_humanoid4096_traffic.md")
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            if line.startswith("hbm_bytes_per_launch"):
                return float(line.split("=")[1])
    return None


def _log(msg: str) -> None:
    """A progress line on stderr (stdout carries only the JSON result)."""
    sys.stderr.write(f"[bench {time.strftime('%H:%M:%S')}] {msg}\n")
    sys.stderr.flush()


def _cpu_model_name() -> str:
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def _cpu_rollout(model, C, steps, dt, line_search, threads, walk_seed):
    """collide + step on the CPU for every world, ``steps`` times; returns (world-steps/s, seconds,
    mean Newton iterations).  PD targets follow a random walk (sigma 0.05 rad per step)."""
    import numpy as np

    from oracle import collide as oc
    from oracle import cpu_step as cs
    from oracle import sap_oracle as so

    opts = so.StepOptions(preset="approx32", line_search=line_search, max_rigid_contact=C)
    stepper = cs.CpuStepper(model, opts)
    n = model.joint_dof_count // model.world_count
    mask = np.zeros(n, np.float32)
    mask[6:] = 1.0 if n > 6 else 0.0
    rng = np.random.default_rng(walk_seed)
    state = dict(joint_q=model.joint_q, joint_qd=model.joint_qd, body_q=model.body_q, body_qd=model.body_qd,
                 body_f=np.zeros((model.body_count, 6), np.float32))
    control = dict(joint_f=model.joint_f, joint_target_pos=model.joint_target_pos.copy(),
                   joint_target_vel=model.joint_target_vel, joint_act=model.joint_act)
    guess = None
    its = []
    t0 = time.perf_counter()
    for _ in range(steps):
        control["joint_target_pos"] = control["joint_target_pos"] + (
            rng.standard_normal((model.world_count, n)).astype(np.float32) * 0.05 * mask).reshape(-1)
        c = oc.collide(model, state["body_q"], C, threads=threads)
        out = stepper.step(state, control, c, dt, v_guess=guess, threads=threads)
        guess = out["v_sap"]
        its.append(float(np.mean(out["iterations"])))
        state = dict(joint_q=out["joint_q"], joint_qd=out["joint_qd"], body_q=out["body_q"], body_qd=out["body_qd"],
                     body_f=state["body_f"])
    el = time.perf_counter() - t0
    return model.world_count * steps / el, el, float(np.mean(its))


def cpu_baseline(worlds: int, steps: int, seed: int, dt: float, line_search: str, threads: int):
    """The C++17 restatement of the step (oracle/cpu_step.cpp, OpenMP over worlds; collision by the
    C oracle, threaded over worlds) on the host: a sample of the benchmark workload at ``threads``
    threads and at 1 thread, and the 64-ant CPU configuration (dt 1e-3, 200 steps) in full."""
    from sap_warp_amd.scenes import build_ant_scene, build_humanoid_scene

    model = build_humanoid_scene(worlds, seed=seed)
    v_all, el_all, it_all = _cpu_rollout(model, 64, steps, dt, line_search, threads, seed)
    v_one, el_one, it_one = _cpu_rollout(model, 64, steps, dt, line_search, 1, seed)
    ants = build_ant_scene(64, seed=1)
    a_all, a_el_all, _ = _cpu_rollout(ants, 16, 200, 1.0e-3, "armijo_decay", threads, 1)
    a_one, a_el_one, _ = _cpu_rollout(ants, 16, 200, 1.0e-3, "armijo_decay", 1, 1)
    return {
        "value": round(v_all, 1), "unit": "world-steps/s", "cores": threads, "kind": "port",
        "value_1_thread": round(v_one, 1),
        "cpu": _cpu_model_name(),
        "sample": (f"C++17 CPU restatement of the step (float64, OpenMP over worlds, {threads} threads) + C collide "
                   f"oracle: {worlds} humanoid worlds of the benchmark workload x {steps} steps (collide + step), "
                   f"{el_all:.2f} s; 1 thread {el_one:.2f} s; mean Newton iterations {it_all:.2f}"),
        "ants_64x200": {"value": round(a_all, 1), "value_1_thread": round(a_one, 1), "seconds": round(a_el_all, 3),
                        "seconds_1_thread": round(a_el_one, 3),
                        "workload": "64 ant worlds (14 DOF), dt 1e-3, 200 steps, armijo_decay, approx32"},
    }


def _scene_and_solver(W, C, seed, dev, line_search):
    from sap_warp_amd.collision import SapCollisionPipeline
    from sap_warp_amd.scenes import build_humanoid_scene
    from sap_warp_amd.solver_sap import SolverSAP

    model = build_humanoid_scene(W, seed=seed, device=dev)
    solver = SolverSAP(model, max_rigid_contact=C, line_search_variant=line_search)
    pipe = SapCollisionPipeline(model, max_contacts_per_world=C)
    return model, solver, pipe


def _target_walk(model, W, steps, seed, dev):
    """PD-target random walk, sigma = 0.05 rad per step on the revolute dofs, precomputed on the GPU."""
    import torch

    n = model.dof_per_env
    gen = torch.Generator(device=dev)
    gen.manual_seed(1234 + seed)
    mask = torch.zeros(n, device=dev)
    mask[6:] = 1.0
    incr = torch.randn((steps, W, n), generator=gen, device=dev) * 0.05 * mask
    return incr.reshape(steps, -1).contiguous()


def train_learned(args, W, C, dt, dev):
    """Training set from plain rollouts of held-in seeds, then GPU training of the warm-start MLP."""
    import torch

    from sap_warp_amd.learned import LearnedWarmStart

    t0 = time.perf_counter()
    seeds = [args.seed + 100 + k for k in range(args.train_seeds)]
    lws = None
    xs, ys = [], []
    for sd in seeds:
        model, solver, pipe = _scene_and_solver(W, C, sd, dev, args.line_search)
        contacts = pipe.contacts()
        ctl = model.control()
        s0, s1 = model.state(), model.state()
        incr = _target_walk(model, W, args.train_steps, sd, dev)
        obj = LearnedWarmStart(solver, seed=0, mlp=None if lws is None else lws.mlp)
        lws = obj if lws is None else lws

        def ctl_fn(k):
            ctl.joint_target_pos.add_(incr[k])

        x, y, _, _ = obj.collect(pipe, s0, s1, ctl, contacts, args.train_steps, dt, control_fn=ctl_fn)
        xs.append(x)
        ys.append(y)
        del obj
    x = torch.cat(xs)
    y = torch.cat(ys)
    torch.cuda.synchronize()
    t_data = time.perf_counter() - t0
    lws.fit(x, y, epochs=args.train_epochs, batch=4096, lr=args.train_lr, lr_final=args.train_lr * 0.1)
    info = {"samples": int(x.shape[0]), "seeds": seeds, "steps_per_seed": args.train_steps,
            "epochs": args.train_epochs, "adam_steps": lws.train_steps, "final_loss": round(lws.final_loss, 5),
            "data_seconds": round(t_data, 2), "train_seconds": round(lws.train_seconds, 2),
            "held_out_seed": args.seed, "network": {"dims": lws.mlp.dims, "features": lws.F}}
    del x, y, xs, ys
    torch.cuda.empty_cache()
    return lws, info


def timed_run(args, W, C, dt, dev, lws_trained, graph=False, schedule=None, tune=False):
    """Warmup + timed steps on the benchmark seed, plain (lws_trained None) or from the learned start.

    graph=True captures the timed steps (PD-target update, collide, step) into one HIP graph
    after the warmup and times one replay of it (no per-step events inside the graph).
    schedule sets the solver's work schedule; tune=True runs the autotuner on the warmed-up
    state first (sap_warp_amd/autotune.py) and keeps its choice.
    """
    import numpy as np
    import torch

    from sap_warp_amd.learned import LearnedWarmStart

    model, solver, pipe = _scene_and_solver(W, C, args.seed, dev, args.line_search)
    if schedule is not None:
        solver.set_schedule(**schedule)
    lws = None
    if lws_trained is not None:
        lws = LearnedWarmStart(solver, mlp=lws_trained.mlp)
    contacts = pipe.contacts()
    s0, s1 = model.state(), model.state()
    ctl = model.control()
    total = args.warmup + args.steps
    incr = _target_walk(model, W, total, args.seed, dev)
    contact_acc = torch.zeros(1, dtype=torch.int64, device=dev)

    def solve(a, b):
        if lws is None:
            solver.step(a, b, ctl, contacts, dt)
        else:
            lws.prepare(a, ctl, contacts)
            solver.step(a, b, ctl, contacts, dt)
            lws.finish()

    for k in range(args.warmup):
        ctl.joint_target_pos.add_(incr[k])
        pipe.collide(s0, contacts)
        solve(s0, s1)
        s0, s1 = s1, s0
    torch.cuda.synchronize()
    tuning = None
    if tune:
        from sap_warp_amd.autotune import autotune

        tuning = autotune(solver, pipe, s0, ctl, contacts, dt)
        # the tuner's own steps restarted from s0; the timed run continues from s0 as well,
        # with the warm-start guess of the last warmup step replaced by the plain start
    torch.cuda.synchronize()
    solver.iteration_histogram(reset=True)
    ev_start = torch.cuda.Event(enable_timing=True)
    ev_end = torch.cuda.Event(enable_timing=True)
    if graph:
        g = torch.cuda.CUDAGraph()
        side = torch.cuda.Stream()
        side.wait_stream(torch.cuda.current_stream())
        a, b = s0, s1
        with torch.cuda.stream(side):
            with torch.cuda.graph(g, stream=side):
                for i in range(args.steps):
                    ctl.joint_target_pos.add_(incr[args.warmup + i])
                    pipe.collide(a, contacts)
                    solver.step(a, b, ctl, contacts, dt)
                    contact_acc.add_(contacts.rigid_contact_count)
                    a, b = b, a
        torch.cuda.current_stream().wait_stream(side)
        torch.cuda.synchronize()
        t_wall0 = time.perf_counter()
        ev_start.record()
        g.replay()
        ev_end.record()
        torch.cuda.synchronize()
        wall = time.perf_counter() - t_wall0
        ms_total = ev_start.elapsed_time(ev_end)
        hist = solver.iteration_histogram(reset=True)
        ws = int(hist[257])
        iters = hist[:256].astype(np.float64)
        cdf = np.cumsum(iters) / max(ws, 1)
        return {"value": W * args.steps / (ms_total / 1000.0), "ms_per_step": ms_total / args.steps, "wall": wall,
                "mean_it": float((iters * np.arange(256)).sum() / max(ws, 1)),
                "p99_it": int(np.searchsorted(cdf, 0.99)), "world_steps": ws,
                "mean_nc": float(contact_acc.item()) / (args.steps * W)}
    ev_c = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev_p = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev_s = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    ev_e = [torch.cuda.Event(enable_timing=True) for _ in range(args.steps)]
    torch.cuda.synchronize()
    t_wall0 = time.perf_counter()
    ev_start.record()
    for i in range(args.steps):
        k = args.warmup + i
        ctl.joint_target_pos.add_(incr[k])
        ev_c[i].record()
        pipe.collide(s0, contacts)
        ev_p[i].record()
        if lws is not None:
            lws.prepare(s0, ctl, contacts)
        ev_s[i].record()
        solver.step(s0, s1, ctl, contacts, dt)
        ev_e[i].record()
        if lws is not None:
            lws.finish()
        contact_acc.add_(contacts.rigid_contact_count)
        s0, s1 = s1, s0
    ev_end.record()
    torch.cuda.synchronize()
    wall = time.perf_counter() - t_wall0
    ms_total = ev_start.elapsed_time(ev_end)
    step_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev_s, ev_e)]))
    pre_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev_p, ev_s)]))
    collide_ms = float(np.mean([a.elapsed_time(b) for a, b in zip(ev_c, ev_p)]))
    hist = solver.iteration_histogram(reset=True)
    ws = int(hist[257])
    iters = hist[:256].astype(np.float64)
    mean_it = float((iters * np.arange(256)).sum() / max(ws, 1))
    cdf = np.cumsum(iters) / max(ws, 1)
    stats = solver.iteration_stats()
    return {
        "schedule": solver.get_schedule(), "tuning": tuning,
        "model": model, "n": model.dof_per_env, "nb": model.bodies_per_env,
        "value": W * args.steps / (ms_total / 1000.0), "ms_per_step": ms_total / args.steps,
        "step_ms": step_ms, "pre_ms": pre_ms, "collide_ms": collide_ms, "wall": wall,
        "mean_it": mean_it, "p99_it": int(np.searchsorted(cdf, 0.99)), "ls_per_ws": float(hist[256]) / max(ws, 1),
        "world_steps": ws, "taken": int(hist[258]), "rejected": int(hist[259]),
        "mean_nc": float(contact_acc.item()) / (args.steps * W),
        "converged_frac": float(np.mean(stats.converged_env == 1)),
        "S_world": (model.shape_count - 1) // W, "P_world": model.shape_pairs.shape[0] // W,
    }


def hetero_run(worlds: int, steps: int, warmup: int, seed: int, dt: float, line_search: str, graph_steps: int,
               pile_sizes=(8,)):
    """The heterogeneous batch (arm+gripper+box 15 DOF, quadruped 18, humanoid 35, pile of 8 free
    bodies 48; kinds drawn per world by seed) in the batch mode: each kind's collide, then one step
    launch per contact variant over every kind's worlds.  The contact variants (16- or 32-slot small
    variant, forked or serial 64-slot variant) are chosen by timing eager steps; the timed steps are
    replays of one HIP graph holding ``graph_steps`` steps, timed with HIP events."""
    import numpy as np
    import torch

    from sap_warp_amd.hetero import HeteroBatch, build_hetero_models

    t0 = time.time()
    models = build_hetero_models(worlds, seed=seed, device="cuda:0", pile_sizes=tuple(pile_sizes))
    build_s = time.time() - t0
    _log(f"heterogeneous models built ({build_s:.1f} s)")
    batch = HeteroBatch(models, dt=dt, line_search_variant=line_search, collect_iteration_stats=True, mode="batch")
    for _ in range(max(warmup, 3)):  # three eager steps bind both parameter slots
        batch.step()
    torch.cuda.synchronize()
    batch.step()  # per-kind Newton iterations of one eager step (the batch kernel fills each solver's counters)
    torch.cuda.synchronize()
    iters = {n: float(np.mean(g["solver"].iteration_stats().newton_iterations_env)) for n, g in batch.groups.items()}
    for g in batch.groups.values():
        g["solver"].collect_iteration_stats = False  # no per-step host reads in the timed or captured steps
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    tuning = {}
    for small in (16, 32):
        for fork in (False, True):
            batch.set_schedule(small_capacity=small, fork=fork)
            for _ in range(2):
                batch.step()
            torch.cuda.synchronize()
            e0.record()
            for _ in range(10):
                batch.step()
            e1.record()
            torch.cuda.synchronize()
            tuning[f"small{small}_{'fork' if fork else 'serial'}"] = round(e0.elapsed_time(e1) / 10, 4)
    best = min(tuning, key=tuning.get)
    batch.set_schedule(small_capacity=int(best[5:7]), fork=best.endswith("fork"))
    for _ in range(2):
        batch.step()
    torch.cuda.synchronize()
    k = max(2, graph_steps - graph_steps % 2)  # even: the state slots return to where they started
    reps = max(1, steps // k)
    g = torch.cuda.CUDAGraph()
    side = torch.cuda.Stream()
    side.wait_stream(torch.cuda.current_stream())
    with torch.cuda.stream(side):
        with torch.cuda.graph(g, stream=side):
            for _ in range(k):
                batch.step()
    torch.cuda.current_stream().wait_stream(side)
    torch.cuda.synchronize()
    g.replay()  # first replay uploads the graph
    torch.cuda.synchronize()
    e0.record()
    for _ in range(reps):
        g.replay()
    e1.record()
    torch.cuda.synchronize()
    ms = e0.elapsed_time(e1) / (reps * k)
    finite = all(bool(torch.isfinite(gr["states"][0].joint_q).all()) and
                 bool(torch.isfinite(gr["states"][1].joint_q).all()) for gr in batch.groups.values())
    return {
        "value": round(batch.world_count * 1000.0 / ms, 1),
        "ms_per_step": round(ms, 4),
        "worlds": batch.world_count,
        "steps_timed": reps * k,
        "kinds": {n: {"worlds": gr["model"].world_count,
                      "dof_per_world": gr["model"].joint_dof_count // gr["model"].world_count,
                      "contact_slots": gr["solver"].max_rigid_contact,
                      "kernel": "batch launch" if n in batch.batched else "own solver (256-thread kernel)",
                      "mean_newton_iterations_one_step": round(iters[n], 3)} for n, gr in batch.groups.items()},
        "schedule": best,
        "autotune_ms_per_step": tuning,
        "finite_after": finite,
        "build_s": round(build_s, 2),
        "timing": f"{reps} replays of one HIP graph holding {k} batch steps (each kind's collide, one step "
                  "launch per contact variant over all kinds), HIP events around the replays",
    }


def bin_world_run(bodies: int, warmup: int, steps: int, line_search: str) -> dict:
    """One world of free bodies stepped over the whole GPU (BigWorldSolver): a bin of ``bodies``
    boxes, spheres and capsules stacked in 6 x 6 collapsing columns (hetero.build_bin_scene,
    synthetic, dt 1e-3), each step = collide (all explicit pairs, thread per pair) + free motion +
    contact rows + Newton solve + integration; wall clock around ``steps`` steps from the initial
    stacking (300 by default, the stress configuration's run) after ``warmup`` untimed collides of
    that state (code objects, first touches); the Newton loop is host-driven, one sync per
    iteration."""
    import time

    import torch

    from sap_warp_amd import hetero
    from sap_warp_amd.large_world import BigWorldSolver

    model = hetero.build_bin_scene(bodies, seed=17, device="cuda:0")
    C = 16 * bodies
    solver = BigWorldSolver(model, max_rigid_contact=C, line_search_variant=line_search, contact_tau_d=0.005)
    contacts = solver.allocate_contacts()
    s0, s1 = model.state(), model.state()
    ctl = model.control()

    def one():
        nonlocal s0, s1
        solver.collide(s0, contacts)
        solver.step(s0, s1, ctl, contacts, 1.0e-3)
        s0, s1 = s1, s0
        return solver.solve_stats()

    for _ in range(warmup):
        solver.collide(s0, contacts)
    torch.cuda.synchronize()
    its, ncs = [], []
    t0 = time.perf_counter()
    for k in range(steps):
        st = one()
        its.append(st.iterations)
        ncs.append(int(contacts.rigid_contact_count.item()))
        if (k + 1) % 100 == 0:
            _log(f"bin world: {k + 1} steps, {ncs[-1]} contacts, {st.iterations} Newton iterations")
    torch.cuda.synchronize()
    dt_s = time.perf_counter() - t0
    solver.close()
    return {
        "workload": f"one world: {bodies} free boxes / spheres / capsules in a bin (collapsing columns), "
                    f"{6 * bodies} dofs, {line_search}, dt 1e-3, synthetic (hetero.build_bin_scene)",
        "world_steps_per_s": round(steps / dt_s, 3),
        "ms_per_step": round(1e3 * dt_s / steps, 3),
        "steps": steps,
        "warmup_collides": warmup,
        "contacts_mean": round(sum(ncs) / len(ncs), 1),
        "contacts_last": ncs[-1],
        "newton_iterations_mean": round(sum(its) / len(its), 2),
        "newton_iterations_max": max(its),
        "timing": "wall clock around the steps from the initial stacking (collide + step, the host reads the "
                  "contact count and each Newton iteration's record)",
    }


def large_world_run(nx: int, ny: int, nz: int, repeats: int, line_search: str) -> dict:
    """One large world on the whole GPU: the SAP Newton solve of a resting block of nx*ny*nz free
    cubes (sap_warp_amd.large_world.stacked_boxes_problem, synthetic; 16x8x8 = 6144 dofs and
    11 520 contacts, the size of a settled thousand-body bin), timed with HIP events around whole
    solves, and the blocked Cholesky of that order timed on its own (its cost does not depend on
    the values: dense float32, 64 x 64 blocks)."""
    import numpy as np
    import torch

    from sap_warp_amd.blocked_cholesky import BlockCholeskySolverBatched
    from sap_warp_amd.large_world import LargeWorldSolver, stacked_boxes_problem

    dev = "cuda:0"
    P = stacked_boxes_problem(nx, ny, nz, seed=0)
    nb = nx * ny * nz
    n, nc = 6 * nb, len(P["phi0"])
    A = torch.zeros((n, n), dtype=torch.float32, device=dev)
    idx = torch.arange(nb, device=dev)
    A.view(nb, 6, nb, 6)[idx, :, idx, :] = torch.tensor(P["A_blocks"], dtype=torch.float32, device=dev)
    T = lambda x, d: torch.tensor(np.ascontiguousarray(x), dtype=d, device=dev)  # noqa: E731
    f64 = torch.float64
    args = dict(A=A, v_star=T(P["v_star"], f64), v_guess=T(P["v_guess"], f64),
                participating=T(P["participating"], torch.int32), J_cols=T(P["J_cols"], torch.int32),
                J_vals=T(P["J_vals"], torch.float32), phi0=T(P["phi0"], f64), weight=T(P["weight"], f64),
                stiffness=T(P["stiffness"], f64), tau=T(P["tau"], f64), mu=T(P["mu"], f64), dt=P["dt"])
    solver = LargeWorldSolver(n, nc, 12)
    _, _, st = solver.solve(line_search=line_search, **args)  # warm-up (code objects, first touches)
    torch.cuda.synchronize()
    e0, e1 = torch.cuda.Event(enable_timing=True), torch.cuda.Event(enable_timing=True)
    e0.record()
    for _ in range(repeats):
        _, _, st = solver.solve(line_search=line_search, **args)
    e1.record()
    torch.cuda.synchronize()
    solve_ms = e0.elapsed_time(e1) / repeats
    solver.close()
    chol = BlockCholeskySolverBatched(n, 1, refine=1)
    Ab = A.unsqueeze(0)
    chol.factorize(Ab, n)
    torch.cuda.synchronize()
    e0.record()
    for _ in range(repeats):
        chol.factorize(Ab, n)
    e1.record()
    torch.cuda.synchronize()
    factor_ms = e0.elapsed_time(e1) / repeats
    chol.close()
    flops = n ** 3 / 3.0
    return {
        "workload": f"one world: {nb} resting free cubes ({nx}x{ny}x{nz}), {n} dofs, {nc} contacts, "
                    f"{line_search}, synthetic (large_world.stacked_boxes_problem)",
        "solve_ms": round(solve_ms, 3),
        "newton_iterations": st.iterations,
        "line_search_trials": st.ls_iterations,
        "converged": st.converged,
        "ms_per_newton_iteration": round(solve_ms / max(st.iterations, 1), 3),
        "factor_ms": round(factor_ms, 3),
        "factor_tflops": round(flops / (factor_ms * 1e-3) / 1e12, 3),
        "factor_frac_of_fp32_peak": round(flops / (factor_ms * 1e-3) / 1e12 / PEAK_FP32_TFLOPS, 5),
        "timing": f"HIP events around {repeats} whole solves (host-driven Newton loop, one sync per iteration) "
                  f"and {repeats} factorisations; flops = n^3/3",
    }


def copy_calibration() -> int:
    """Three copies of a 512 MiB buffer: the HBM counters of a kernel of known traffic."""
    import torch

    a = torch.ones(128 * 1024 * 1024, dtype=torch.float32, device="cuda:0")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_bytes": a.numel() * 4, "copies": 3}))
    return 0


def main() -> int:
    ap = argparse.ArgumentParser(description=__doc__)
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=500)
    ap.add_argument("--warmup", type=int, default=20)
    ap.add_argument("--worlds", type=int, default=4096)
    ap.add_argument("--seed", type=int, default=2)
    ap.add_argument("--line-search", default="exact_root", choices=["exact_root", "armijo_decay"])
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-worlds", type=int, default=1024)
    ap.add_argument("--cpu-steps", type=int, default=30)
    ap.add_argument("--cpu-threads", type=int, default=int(os.environ.get("OMP_NUM_THREADS", "0")) or os.cpu_count())
    ap.add_argument("--learned", type=int, default=1, help="train and time the learned warm start (1) or not (0)")
    ap.add_argument("--graph", type=int, default=1, help="time the plain steps as one HIP graph replay (1) or not (0)")
    ap.add_argument("--autotune", type=int, default=1, help="choose the step's work schedule by timing (1) or keep the default (0)")
    ap.add_argument("--train-seeds", type=int, default=2)
    ap.add_argument("--train-steps", type=int, default=250)
    ap.add_argument("--train-epochs", type=int, default=12)
    ap.add_argument("--train-lr", type=float, default=1e-3)
    ap.add_argument("--hetero", type=int, default=1, help="also time the heterogeneous batch (1) or not (0)")
    ap.add_argument("--hetero-worlds", type=int, default=8192)
    ap.add_argument("--hetero-steps", type=int, default=1000)
    ap.add_argument("--hetero-graph-steps", type=int, default=100,
                    help="batch steps captured in the heterogeneous batch's HIP graph (replayed to cover the steps)")
    ap.add_argument("--hetero-large", type=int, default=1,
                    help="also time the heterogeneous batch with piles of 8, 16 and 32 bodies (1) or not (0)")
    ap.add_argument("--hetero-large-steps", type=int, default=200)
    ap.add_argument("--large-world", type=int, default=1,
                    help="also time the single large world (1024 cubes, 6144 dofs) solve (1) or not (0)")
    ap.add_argument("--large-world-repeats", type=int, default=3)
    ap.add_argument("--bin-bodies", type=int, default=1024, help="bodies of the bin world (0: skip it)")
    ap.add_argument("--bin-steps", type=int, default=300)
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 512 MiB device copies (scales rocprofv3 FETCH_SIZE / WRITE_SIZE)")
    args = ap.parse_args()
    if args.copy_calibration:
        return copy_calibration()
    if args.gpus != 1 or int(os.environ.get("WORLD_SIZE", "1")) != 1:
        print(json.dumps({"error": "this build runs on one GPU only; worlds shard trivially across GPUs "
                                   "(see DESIGN.md) but no multi-GPU path is built", "n_gpus": args.gpus}))
        return 2

    import numpy as np
    import torch

    dev = "cuda:0"
    torch.cuda.set_device(0)
    dt = 2.0e-3
    C = 64
    W = args.worlds

    lws = None
    train_info = None
    if args.learned:
        lws, train_info = train_learned(args, W, C, dt, dev)
        _log("learned warm start trained")
    plain = timed_run(args, W, C, dt, dev, None, tune=bool(args.autotune))
    sched = plain["schedule"]
    _log(f"plain eager run done ({plain['ms_per_step']:.4f} ms/step)")
    plain_graph = timed_run(args, W, C, dt, dev, None, graph=True, schedule=sched) if args.graph else None
    if plain_graph is not None:
        _log(f"plain graph run done ({plain_graph['ms_per_step']:.4f} ms/step)")
    learned = timed_run(args, W, C, dt, dev, lws, schedule=sched) if lws is not None else None
    if learned is not None:
        _log("learned run done")
    head = plain  # the benchmark workload is plain SAP; the learned start is reported beside it
    n, nb = head["n"], head["nb"]
    # roofline of the dominant kernel, the step (the wave kernel pair sapx_step launches on the
    # solver's stream and a forked second stream: the 16-contact variant, and the 64-contact
    # variant for the worlds above 16): algorithmic FP32-class operations per step over the
    # HIP-event time of the pair
    flops_launch = W * step_flops(n, nb, head["mean_nc"], head["mean_it"], head["ls_per_ws"])
    achieved_tflops = flops_launch / (head["step_ms"] / 1000.0) / 1.0e12
    cbytes = collide_bytes(head["model"], head["mean_nc"] * W, head["S_world"], head["P_world"])
    collide_gbs = cbytes / (head["collide_ms"] / 1000.0) / 1.0e9
    kernel_ms = {"sapx_step": round(head["step_ms"], 4), "collide_total": round(head["collide_ms"], 4)}
    out = {
        "metric": "world-steps/s at 4096 humanoid worlds; mean Newton iters, learned vs plain",
        "value": round(plain_graph["value"] if plain_graph is not None else head["value"], 1),
        "unit": "world-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(plain_graph["ms_per_step"] if plain_graph is not None else head["ms_per_step"], 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32/f64",
        "data": "synthetic (seeded humanoid batch, random-walk PD targets)",
        "config": {"workload": "batched SAP, 4096 humanoid worlds (35 DOF), dt=2e-3, approx32 preset, "
                               f"{args.line_search} line search, C=64 contacts/world, plain warm start "
                               "(previous step's solution)",
                   "worlds": W, "dof_per_world": n, "bodies_per_world": nb, "dt": dt, "steps_timed": args.steps},
        "newton_iterations": {"mean": round(head["mean_it"], 3), "p99": head["p99_it"],
                              "line_search_trials_per_world_step": round(head["ls_per_ws"], 3),
                              "world_steps": head["world_steps"],
                              "warm_start": "previous step (plain)"},
        "mean_contacts_per_world": round(head["mean_nc"], 3),
        "converged_fraction_last_step": head["converged_frac"],
        "kernel_ms": kernel_ms,
        "roofline": {"bound": "mfma", "achieved": round(achieved_tflops, 4), "peak": PEAK_FP32_TFLOPS,
                     "unit": "TFLOP/s", "frac": round(achieved_tflops / PEAK_FP32_TFLOPS, 6),
                     "traffic": read_traffic(), "kernel": STEP_KERNELS,
                     "flops_per_launch": round(flops_launch)},
        "roofline_collide": {"bound": "hbm", "achieved": round(collide_gbs, 2), "peak": PEAK_HBM_GBS, "unit": "GB/s",
                             "frac": round(collide_gbs / PEAK_HBM_GBS, 6), "bytes_per_launch": round(cbytes)},
        "timing": ("one replay of a HIP graph holding all timed steps (PD-target update, collide, step); "
                   "kernel_ms and roofline from a second, eager run of the same steps with per-step HIP events")
                  if plain_graph is not None else "eager launches, HIP events around the timed steps",
        "eager": {"value": round(plain["value"], 1), "ms_per_step": round(plain["ms_per_step"], 4)},
        "schedule": sched,
        "autotune_ms_per_step": plain["tuning"]["ms_per_step"] if plain["tuning"] else None,
        "wall_s": round(head["wall"], 3),
    }
    if plain_graph is not None:
        out["graph"] = {"value": round(plain_graph["value"], 1), "ms_per_step": round(plain_graph["ms_per_step"], 4),
                        "mean_newton_iterations": round(plain_graph["mean_it"], 3),
                        "mean_contacts_per_world": round(plain_graph["mean_nc"], 3)}
    if learned is not None:
        taken, rejected = learned["taken"], learned["rejected"]
        out["learned"] = {
            "value": round(learned["value"], 1),
            "ms_per_step": round(learned["ms_per_step"], 4),
            "step_ms": round(learned["step_ms"], 4),
            "features_mlp_ms": round(learned["pre_ms"], 4),
            "speedup_vs_plain": round(learned["value"] / plain["value"], 4),
            "mean_newton_iterations": round(learned["mean_it"], 3),
            "p99_newton_iterations": learned["p99_it"],
            "iteration_reduction": round(1.0 - learned["mean_it"] / max(plain["mean_it"], 1e-12), 4),
            "fallback_rate": round(rejected / max(taken + rejected, 1), 4),
            "train": train_info,
        }
    if args.hetero:
        out["hetero"] = hetero_run(args.hetero_worlds, args.hetero_steps, min(args.warmup, 20), args.seed, dt,
                                   args.line_search, args.hetero_graph_steps)
        _log(f"heterogeneous batch done ({out['hetero']['ms_per_step']:.4f} ms/step)")
        if args.hetero_large:
            out["hetero_large_piles"] = hetero_run(args.hetero_worlds, args.hetero_large_steps, min(args.warmup, 20),
                                                   args.seed, dt, args.line_search, args.hetero_graph_steps,
                                                   pile_sizes=(8, 16, 32))
            _log(f"heterogeneous batch with large piles done ({out['hetero_large_piles']['ms_per_step']:.4f} ms/step)")
    if args.large_world:
        out["large_world"] = large_world_run(16, 8, 8, args.large_world_repeats, args.line_search)
        _log(f"large world done ({out['large_world']['solve_ms']:.1f} ms/solve, "
             f"factor {out['large_world']['factor_ms']:.2f} ms)")
        if args.bin_bodies > 0:
            out["large_world"]["bin_step"] = bin_world_run(args.bin_bodies, 2, args.bin_steps, args.line_search)
            _log(f"bin world done ({out['large_world']['bin_step']['ms_per_step']:.1f} ms/step)")
    if not args.no_cpu_baseline:
        out["cpu_baseline"] = cpu_baseline(args.cpu_worlds, args.cpu_steps, args.seed, dt, args.line_search,
                                           args.cpu_threads)
        _log("CPU baseline done")
    else:
        out["cpu_baseline"] = None
    print(json.dumps(out))
    return 0


if __name__ == "__main__":
    sys.exit(main())

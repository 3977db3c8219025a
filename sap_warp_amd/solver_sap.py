"""``SolverSAP`` — drop-in for the reference's step facade (``sim/solver_sap.py:736-1512``).

Same constructor keywords, presets and ``step(state_in, state_out, control,
contacts, dt)`` signature.  The whole step (boundary conversion, free motion,
contact Jacobian, SAP Newton solve, integration, public FK) runs in one fused
HIP kernel per call (``sapx_step`` in ``libsapx.so``); this module validates
arguments, expands presets and passes device pointers.
"""
from __future__ import annotations

import ctypes
import warnings

import numpy as np

from . import _native as nat
from .runtime import SapContacts, SapControl, SapModel, SapState

SAP_CONTACT_PRESET_VARIANTS = ("approx32", "approx64", "drake")

# solver_sap.py:75-108
_CONTACT_PRESET_KWARGS = {
    "approx32": dict(use_f64_boundary_pose=False, free_motion_solve_precision="fp32", contact_solve_precision="fp64",
                     contact_linear_solve_precision="fp32", sap_contact_weight_precision="fp32",
                     contact_weight_mode="body_inertia", contact_point_mode="witness_point",
                     position_integration="midpoint"),
    "approx64": dict(use_f64_boundary_pose=True, free_motion_solve_precision="fp64", contact_solve_precision="fp64",
                     contact_linear_solve_precision="fp64", sap_contact_weight_precision="fp64",
                     contact_weight_mode="body_inertia", contact_point_mode="witness_point",
                     position_integration="midpoint"),
    "drake": dict(use_f64_boundary_pose=True, free_motion_solve_precision="fp64", contact_solve_precision="fp64",
                  contact_linear_solve_precision="fp64", sap_contact_weight_precision="fp64",
                  contact_weight_mode="diag_delassus", contact_point_mode="contact_midpoint",
                  position_integration="sap_euler"),
}


def normalize_sap_line_search_mode(value: str) -> str:
    """contact_solve.py:68-79."""
    mode = str(value).strip().lower().replace("-", "_")
    if mode in ("monotone_decay", "armijo_decay", "exact_root"):
        return mode
    raise ValueError(
        "line_search_variant must be 'monotone_decay', 'armijo_decay', or 'exact_root', " f"got {value!r}."
    )


def _normalize_precision_knob(value: str, *, option_name: str) -> str:
    p = str(value).strip().lower()
    p = {"f32": "fp32", "f64": "fp64"}.get(p, p)
    if p not in ("fp32", "fp64"):
        raise ValueError(f"{option_name} must be 'fp32'/'f32' or 'fp64'/'f64', got {value!r}.")
    return p


def _normalize_contact_preset_variant(value: str | None) -> str:
    if value is None:
        return "approx32"
    v = str(value).strip().lower().replace("-", "_")
    v = {"approx_32": "approx32", "approx_64": "approx64"}.get(v, v)
    if v not in _CONTACT_PRESET_KWARGS:
        raise ValueError(f"contact_preset_variant must be one of {', '.join(SAP_CONTACT_PRESET_VARIANTS)}, got {value!r}.")
    return v


def _normalize_weight_mode(v: str) -> str:
    m = str(v).strip().lower().replace("-", "_")
    if m not in ("body_inertia", "diag_delassus"):
        raise ValueError(f"contact_weight_mode must be 'body_inertia' or 'diag_delassus', got {v!r}.")
    return m


def _normalize_point_mode(v: str) -> str:
    m = str(v).strip().lower().replace("-", "_")
    if m not in ("witness_point", "contact_midpoint"):
        raise ValueError(f"contact_point_mode must be 'witness_point' or 'contact_midpoint', got {v!r}.")
    return m


def infer_contact_mu_fallback(model: SapModel, default: float = 0.1) -> float:
    """``_infer_contact_mu_fallback`` (sap_helpers.py:40-45)."""
    mu = np.asarray(model.shape_material_mu, dtype=np.float32).reshape(-1)
    pos = mu[mu > 0.0]
    return float(default) if pos.size == 0 else float(np.min(pos))


def infer_contact_tau_d_fallback(model: SapModel) -> float:
    """``_infer_contact_tau_d_fallback`` (solver_sap.py:128-142)."""
    vals = np.asarray(model.shape_material_tau, dtype=np.float64).reshape(-1)
    ok = vals[np.isfinite(vals) & (vals >= 0.0)]
    return float(ok[0]) if ok.size else 0.0


_WARNED: set = set()


def _warn_once(msg: str) -> None:
    if msg not in _WARNED:
        _WARNED.add(msg)
        warnings.warn(msg, stacklevel=3)


class SapContactSolveResult:
    """Per-world statistics of the last solve (``SapContactSolveResult`` contact_solve.py:188-232)."""

    def __init__(self, newton_iterations, line_search_iterations, contact_count, flags):
        self.newton_iterations_env = newton_iterations
        self.line_search_iterations_env = line_search_iterations
        self.contact_count_env = contact_count
        self.converged_env = (flags & 1).astype(np.int32)
        self.optimality_reached_env = ((flags >> 1) & 1).astype(np.int32)
        self.cost_reached_env = ((flags >> 2) & 1).astype(np.int32)
        self.max_iterations_reached_env = ((flags >> 3) & 1).astype(np.int32)
        self.line_search_status_env = ((flags >> 8) & 0xFF).astype(np.int8).astype(np.int32)


class SolverSAP:
    """SAP-native solver (``SolverSAP``, solver_sap.py:736-1494) on the MI355X native step."""

    def __init__(
        self,
        model: SapModel,
        *,
        max_rigid_contact: int = 128,
        fallback_mu: float | None = None,
        fallback_stiffness: float = 1.0e10,
        contact_beta: float = 1.0,
        contact_sigma: float = 1.0e-3,
        contact_tau_d: float | None = None,
        block_size: int | None = None,
        diag_shift: float = 0.0,
        max_iterations: int = 100,
        optimality_abs_tol: float = 1.0e-14,
        optimality_rel_tol: float = 1.0e-6,
        cost_abs_tol: float | None = None,
        cost_rel_tol: float | None = None,
        line_search_max_iterations: int = 40,
        armijo_c: float = 1.0e-4,
        rho: float = 0.8,
        line_search_relative_slop: float | None = None,
        line_search_variant: str = "armijo_decay",
        contact_preset_variant: str | None = None,
        contact_weight_mode: str | None = None,
        contact_point_mode: str | None = None,
        capture_contact_jacobian_snapshots: bool = False,
        use_f64_boundary_pose: bool | None = None,
        free_motion_solve_precision: str | None = None,
        contact_solve_precision: str | None = None,
        contact_linear_solve_precision: str | None = None,
        sap_contact_weight_precision: str | None = None,
        collect_iteration_stats: bool = False,
        check_line_search_errors: bool = False,
        graph_conditional: bool = True,
        static_substep: bool = False,
        static_substep_iterations: int = 8,
        static_substep_line_search_iterations: int = 6,
        position_integration: str | None = None,
    ):
        if not isinstance(model, SapModel):
            raise TypeError("SolverSAP requires SapModel; convert frontend data before constructing the solver.")
        if int(model.joint_count) <= 0 or int(model.joint_dof_count) <= 0:
            raise ValueError("SolverSAP requires a model with articulated joint DOFs.")
        self._device = nat.require_gpu_device(model.device, "SolverSAP")
        preset = _CONTACT_PRESET_KWARGS[_normalize_contact_preset_variant(contact_preset_variant)]
        self.contact_preset_variant = _normalize_contact_preset_variant(contact_preset_variant)
        self.use_f64_boundary_pose = bool(preset["use_f64_boundary_pose"] if use_f64_boundary_pose is None
                                          else use_f64_boundary_pose)
        self.free_motion_solve_precision = _normalize_precision_knob(
            free_motion_solve_precision or preset["free_motion_solve_precision"], option_name="free_motion_solve_precision")
        self.contact_solve_precision = _normalize_precision_knob(
            contact_solve_precision or preset["contact_solve_precision"], option_name="contact_solve_precision")
        self.contact_linear_solve_precision = _normalize_precision_knob(
            contact_linear_solve_precision or preset["contact_linear_solve_precision"],
            option_name="contact_linear_solve_precision")
        self.sap_contact_weight_precision = _normalize_precision_knob(
            sap_contact_weight_precision or preset["sap_contact_weight_precision"],
            option_name="sap_contact_weight_precision")
        fp64_data = [k for k in ("free_motion_solve_precision", "contact_linear_solve_precision")
                     if getattr(self, k) == "fp64"]
        if fp64_data:
            # the reference forms H, J and solves in float64 for these knobs (solver_sap.py:88-108)
            _warn_once(
                "SolverSAP: " + ", ".join(f"{k}='fp64'" for k in fp64_data) + " requested; the native step forms "
                "the dynamics matrix and contact Jacobian in float32 and solves Newton systems with a float32 "
                "factor refined by float64 residuals (gradients, projection, costs and line search stay float64). "
                "Velocities agree with a float64 solve within 1e-5 relative; the SAP optimality test holds on the "
                "float32-formed problem.")
        self.contact_weight_mode = _normalize_weight_mode(contact_weight_mode or preset["contact_weight_mode"])
        self.contact_point_mode = _normalize_point_mode(contact_point_mode or preset["contact_point_mode"])
        pi = str(position_integration or preset["position_integration"]).strip().lower().replace("-", "_")
        if pi not in ("sap_euler", "midpoint"):
            raise ValueError(f"position_integration must be 'sap_euler' or 'midpoint', got {position_integration!r}.")
        self.position_integration = pi

        self.model = model
        self.max_rigid_contact = int(max_rigid_contact)
        self.max_iterations = int(max_iterations)
        self.optimality_abs_tol = float(optimality_abs_tol)
        self.optimality_rel_tol = float(optimality_rel_tol)
        self.armijo_c = float(armijo_c)
        self.rho = float(rho)
        self.line_search_variant = normalize_sap_line_search_mode(line_search_variant)
        if self.line_search_variant == "armijo_decay" and not (0.0 < float(rho) < 1.0):
            raise ValueError(f"rho must lie in (0, 1), got {rho!r}.")
        # cost-criterion defaults per variant (contact_solve.py:6179-6182)
        if cost_abs_tol is None:
            cost_abs_tol = 0.0 if self.line_search_variant == "monotone_decay" else 1.0e-30
        if cost_rel_tol is None:
            cost_rel_tol = 5.0e-3 if self.line_search_variant == "monotone_decay" else 1.0e-15
        self.cost_abs_tol = float(cost_abs_tol)
        self.cost_rel_tol = float(cost_rel_tol)
        self.line_search_max_iterations = int(line_search_max_iterations)
        if self.line_search_variant == "exact_root" and self.line_search_max_iterations == 40:
            self.line_search_max_iterations = 100
        self.line_search_relative_slop = (1000.0 * np.finfo(np.float64).eps if line_search_relative_slop is None
                                          else float(line_search_relative_slop))
        if not bool(graph_conditional):
            raise ValueError(
                "SolverSAP only supports graph_conditional=True; the Python contact-solve loop has been removed."
            )
        self.graph_conditional = True
        self.static_substep = bool(static_substep)
        self.static_substep_iterations = int(static_substep_iterations)
        self.static_substep_line_search_iterations = int(static_substep_line_search_iterations)
        self.collect_iteration_stats = bool(collect_iteration_stats)
        self.check_line_search_errors = bool(check_line_search_errors)
        self.capture_contact_jacobian_snapshots = bool(capture_contact_jacobian_snapshots)
        self.block_size = block_size
        self.diag_shift = float(diag_shift)
        self.contact_beta = float(contact_beta)
        self.contact_sigma = float(contact_sigma)

        shape_tau = infer_contact_tau_d_fallback(model) if contact_tau_d is None else float(contact_tau_d)
        if contact_tau_d is None and shape_tau == 0.0:
            warnings.warn(
                "No contact dissipation authored on the model and no contact_tau_d given: tau_d falls back to 0.0, "
                "making contacts undamped springs at near-rigid stiffness (bounce/rattle). contact_tau_d is a "
                "PER-SHAPE fallback (contacting pairs use the sum of both shapes' tau); pass a value on the order "
                "of the simulation step (Drake guidance: tau_d ~ dt) for stable contact.",
                stacklevel=2,
            )
        self.contact_tau_d_shape = shape_tau
        self.fallback_mu = float(fallback_mu) if fallback_mu is not None else infer_contact_mu_fallback(model)
        self.fallback_stiffness = float(fallback_stiffness)

        # _resolve_contact_shape_mu (sap_helpers.py:48-58): an explicit mu overrides a uniform model mu
        native_model_src = model
        mu = np.asarray(model.shape_material_mu, dtype=np.float32)
        if fallback_mu is not None and mu.size and np.allclose(mu, mu[0], rtol=0.0, atol=1.0e-7):
            from dataclasses import replace

            native_model_src = replace(model, shape_material_mu=np.full_like(mu, self.fallback_mu))
        self._native_model = nat.NativeModel(native_model_src, self._device)
        self._lib = nat.load()
        opts = nat.SapxSolverOpts()
        opts.max_rigid_contact = self.max_rigid_contact
        opts.fallback_mu = self.fallback_mu
        opts.fallback_stiffness = self.fallback_stiffness
        opts.fallback_tau_d = self.contact_tau_d_shape
        opts.contact_beta = self.contact_beta
        opts.contact_sigma = self.contact_sigma
        opts.diag_shift = self.diag_shift
        # static_substep: the reference unrolls fixed-count Newton / line-search loops with finished
        # worlds masked to no-ops (contact_solve.py:5601-5604, 5803-5805, 5208-5212), i.e. per world
        # the conditional loop capped at the static counts; the fused step has no per-iteration
        # graph nodes, so the caps are the whole difference
        opts.max_iterations = self.static_substep_iterations if self.static_substep else self.max_iterations
        opts.optimality_abs_tol = self.optimality_abs_tol
        opts.optimality_rel_tol = self.optimality_rel_tol
        opts.cost_abs_tol = self.cost_abs_tol
        opts.cost_rel_tol = self.cost_rel_tol
        opts.line_search = {"exact_root": nat.LS_EXACT_ROOT, "armijo_decay": nat.LS_ARMIJO_DECAY,
                            "monotone_decay": nat.LS_MONOTONE_DECAY}[self.line_search_variant]
        opts.line_search_max_iterations = (self.static_substep_line_search_iterations if self.static_substep
                                           else self.line_search_max_iterations)
        opts.armijo_c = self.armijo_c
        opts.rho = self.rho
        opts.line_search_relative_slop = self.line_search_relative_slop
        opts.contact_weight_mode = (nat.WEIGHT_DIAG_DELASSUS if self.contact_weight_mode == "diag_delassus"
                                    else nat.WEIGHT_BODY_INERTIA)
        opts.contact_point_mode = nat.POINT_MIDPOINT if self.contact_point_mode == "contact_midpoint" else nat.POINT_WITNESS
        opts.position_integration = (nat.INTEGRATE_SAP_EULER if self.position_integration == "sap_euler"
                                     else nat.INTEGRATE_MIDPOINT)
        opts.capture_snapshots = int(self.capture_contact_jacobian_snapshots)
        self._opts = opts
        self._handle = self._create_native(opts)
        self.dof_count = int(model.joint_dof_count)
        self.num_envs = int(model.world_count)
        self.dof_per_env = self.dof_count // max(self.num_envs, 1)
        import torch

        # per-world timestep, solver-owned (_update_dt_world, solver_sap.py:975-992)
        self._dt_world = torch.zeros(max(self.num_envs, 1), dtype=torch.float64, device=self._device)
        self.last_contacts = None
        self.last_contact_solve_result = None
        self.reset_runtime_state()

    def _create_native(self, opts):
        h = ctypes.c_void_p()
        with nat.on_device(self._device):
            nat.check(self._lib.sapx_solver_create(self._native_model.handle, ctypes.byref(opts), ctypes.byref(h)),
                      "sapx_solver_create")
        return h

    # ------------------------------------------------------------------
    @property
    def device(self):
        return self.model.device

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.sapx_solver_destroy(self._handle)
            self._handle = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def get_max_contact_count(self) -> int:
        return int(self.max_rigid_contact) * max(int(self.num_envs), 1)

    def reset_runtime_state(self, state: SapState | None = None) -> None:
        """solver_sap.py:1004-1017: clears counters and the warm-start guess."""
        self.sim_time = 0.0
        self.frame_id = 0
        self.last_contact_count = 0
        self.last_truncated_contact_count = 0
        self.last_solve_iterations = 0
        self.last_line_search_iterations = 0
        self.last_converged = True
        self._has_contact_solve_v_guess = False
        nat.check(self._lib.sapx_solver_reset(self._handle, None), "sapx_solver_reset")
        self.last_contacts = None
        self.last_contact_solve_result = None

    def notify_model_changed(self, flags: int) -> None:
        return

    def update_contacts(self, contacts: SapContacts) -> None:
        raise NotImplementedError("SolverSAP does not expose contact-force writeback yet.")

    # ------------------------------------------------------------------
    def _check_state(self, st: SapState, name: str) -> None:
        import torch

        m = self.model
        for f in ("joint_q", "joint_qd", "body_q"):
            if not isinstance(getattr(st, f), torch.Tensor):
                raise TypeError(f"{name}.{f} must be a torch tensor on {self._device}")
        if st.joint_q.dtype != torch.float32 or st.joint_qd.dtype != torch.float32:
            raise TypeError(f"{name}: joint_q / joint_qd must be float32 tensors")
        if st.body_q.dtype != torch.float32:
            raise TypeError(f"{name}: body_q must be a float32 transform array")
        if st.joint_q.numel() != m.joint_coord_count or st.joint_qd.numel() != m.joint_dof_count:
            raise ValueError(f"{name}: joint_q / joint_qd sizes do not match the model")
        if st.body_q.numel() != m.body_count * 7:
            raise ValueError(f"{name}: body_q size does not match the model")
        for f, size in (("body_qd", m.body_count * 6), ("body_f", m.body_count * 6)):
            t = getattr(st, f)
            if t is not None and (not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or t.numel() != size):
                raise TypeError(f"{name}.{f} must be a float32 tensor of {size} entries (or None)")
        for f in ("joint_q", "joint_qd", "body_q", "body_qd", "body_f"):
            t = getattr(st, f)
            if t is not None:
                nat.require_on(t, self._device, f"{name}.{f}")
                if not t.is_contiguous():
                    raise ValueError(f"{name}: tensors must be contiguous")
        if getattr(st, "joint_qd_order", "public") != "public":
            raise NotImplementedError("the native step takes public-order joint_qd (joint_qd_order='public')")
        if st.body_f is not None and getattr(st, "body_f_order", "public") != "public":
            raise NotImplementedError("the native step takes public-order body_f (body_f_order='public')")

    def _check_control(self, control: SapControl) -> None:
        import torch

        n = self.model.joint_dof_count
        for f in ("joint_f", "joint_target_pos", "joint_target_vel", "joint_act"):
            t = getattr(control, f)
            if t is None:
                continue
            if not isinstance(t, torch.Tensor) or t.dtype != torch.float32 or t.numel() != n:
                raise TypeError(f"control.{f} must be a float32 tensor of joint_dof_count={n} entries (or None)")
            nat.require_on(t, self._device, f"control.{f}")
            if not t.is_contiguous():
                raise ValueError(f"control.{f} must be contiguous")

    def _set_dt_world(self, dt):
        """Per-world dt into the solver's float64 buffer (``_update_dt_world``, solver_sap.py:975-992)."""
        if dt.numel() != self.num_envs:
            raise ValueError(f"per-world dt length {dt.numel()} != num_envs {self.num_envs}")
        nat.require_on(dt, self._device, "dt")
        if not dt.dtype.is_floating_point:
            raise TypeError("per-world dt must be a float32 or float64 tensor")
        self._dt_world.copy_(dt.reshape(-1))
        return self._dt_world

    def step(self, state_in: SapState, state_out: SapState, control: SapControl | None,
             contacts: SapContacts | None, dt) -> SapState:
        """Advance one SAP timestep (solver_sap.py:1388-1494)."""
        import torch

        if getattr(state_in, "requires_grad", False):
            raise NotImplementedError("SolverSAP does not support grad states.")
        if not isinstance(state_in, SapState):
            raise TypeError("SolverSAP.step requires SapState input.")
        if not isinstance(state_out, SapState):
            raise TypeError("SolverSAP.step requires SapState output.")
        if control is not None and not isinstance(control, SapControl):
            raise TypeError("SolverSAP.step requires SapControl.")
        if contacts is not None and not isinstance(contacts, SapContacts):
            raise TypeError("SolverSAP.step requires SapContacts.")
        if self.collect_iteration_stats and torch.cuda.is_current_stream_capturing():
            raise RuntimeError("collect_iteration_stats reads the statistics back to the host after every step, "
                               "which a HIP graph capture cannot record; construct the solver with "
                               "collect_iteration_stats=False for captured steps (iteration_histogram() reads "
                               "the accumulated statistics after a replay)")
        self._check_state(state_in, "state_in")
        self._check_state(state_out, "state_out")
        if control is not None and getattr(control, "joint_f_order", "public") != "public":
            raise NotImplementedError("the native step takes public-order joint_f (joint_f_order='public')")
        if control is not None:
            self._check_control(control)
        if contacts is not None:
            nat.check_contacts(contacts, self._device, self.model.world_count)
        sin = nat.SapxStateView(nat._ptr(state_in.joint_q), nat._ptr(state_in.joint_qd), nat._ptr(state_in.body_q),
                                nat._ptr(state_in.body_qd), nat._ptr(state_in.body_f))
        sout = nat.SapxStateView(nat._ptr(state_out.joint_q), nat._ptr(state_out.joint_qd),
                                 nat._ptr(state_out.body_q), nat._ptr(state_out.body_qd), nat._ptr(state_out.body_f))
        ctl = nat.SapxControlView()
        if control is not None:
            ctl = nat.SapxControlView(nat._ptr(control.joint_f), nat._ptr(control.joint_target_pos),
                                      nat._ptr(control.joint_target_vel), nat._ptr(control.joint_act))
        cview = nat.contacts_view(contacts if contacts is not None else SapContacts())
        dt_world = None
        if isinstance(dt, torch.Tensor):
            dt_world = self._set_dt_world(dt)
            dt_scalar = 0.0
        else:
            dt_scalar = float(dt)
            if not dt_scalar > 0.0:
                raise ValueError("dt must be positive")
        with nat.on_device(self._device):
            rc = self._lib.sapx_step(self._handle, ctypes.byref(sin), ctypes.byref(sout), ctypes.byref(ctl),
                                     ctypes.byref(cview), dt_scalar, nat._ptr(dt_world), nat.stream_ptr())
        nat.check(rc, "sapx_step")
        self.last_contacts = contacts
        self._has_contact_solve_v_guess = True
        if self.collect_iteration_stats:
            self.last_contact_solve_result = self.iteration_stats()
            r = self.last_contact_solve_result
            self.last_contact_count = int(r.contact_count_env.sum())
            self.last_solve_iterations = int(r.newton_iterations_env.max()) if r.newton_iterations_env.size else 0
            self.last_line_search_iterations = int(r.line_search_iterations_env.sum())
            self.last_converged = bool(np.all(r.converged_env == 1))
        else:
            self.last_solve_iterations = -1
            self.last_line_search_iterations = -1
            self.last_converged = True
        if not isinstance(dt, torch.Tensor):
            self.sim_time += float(dt)
        self.frame_id += 1
        return state_out

    # ------------------------------------------------------------------
    def get_schedule(self) -> dict:
        """Current work schedule of the native step (``sapx_solver_get_schedule``)."""
        sc = nat.SapxSchedule()
        nat.check(self._lib.sapx_solver_get_schedule(self._handle, ctypes.byref(sc)), "sapx_solver_get_schedule")
        return {"kernel": "wave" if sc.kernel == nat.KERNEL_WAVE else "fused", "small_capacity": sc.small_capacity,
                "fork": bool(sc.fork), "world_order": bool(sc.world_order)}

    def set_schedule(self, kernel: str = "wave", small_capacity: int = 16, fork: bool = True,
                     world_order: bool = True) -> None:
        """Set the work schedule (raises NotImplementedError when the model does not fit it)."""
        sc = nat.SapxSchedule(nat.KERNEL_WAVE if kernel == "wave" else nat.KERNEL_FUSED, int(small_capacity),
                              int(bool(fork)), int(bool(world_order)))
        nat.check(self._lib.sapx_solver_set_schedule(self._handle, ctypes.byref(sc)), "sapx_solver_set_schedule")

    def iteration_stats(self) -> SapContactSolveResult:
        """Per-world Newton/line-search statistics of the last step (synchronises)."""
        W = self.num_envs
        arrs = [np.zeros(W, dtype=np.int32) for _ in range(4)]
        nat.check(self._lib.sapx_get_stats(self._handle, *[a.ctypes.data for a in arrs], nat.stream_ptr()),
                  "sapx_get_stats")
        return SapContactSolveResult(*arrs)

    def iteration_histogram(self, reset: bool = True) -> np.ndarray:
        """Accumulated Newton-iteration histogram (``sapx_get_histogram``): 260 int64 entries."""
        h = np.zeros(nat.HIST_SIZE, dtype=np.int64)
        nat.check(self._lib.sapx_get_histogram(self._handle, h.ctypes.data, int(bool(reset)), nat.stream_ptr()),
                  "sapx_get_histogram")
        return h

    def read_buffer(self, which: int):
        """Copy an internal buffer of the last step (``sapx_solver_read``) into a new float64 tensor."""
        import torch

        W, n, C = self.num_envs, self.dof_per_env, self.max_rigid_contact
        shapes = {nat.READ_V: (W * n,), nat.READ_V_STAR: (W * n,), nat.READ_V0: (W * n,),
                  nat.READ_GAMMA: (W, C, 3), nat.READ_PHI0: (W, C), nat.READ_WEIGHT: (W, C), nat.READ_A_DIAG: (W * n,),
                  nat.READ_A: (W, n, n), nat.READ_J: (W, C, 3, n)}
        out = torch.zeros(shapes[which], dtype=torch.float64, device=self.model.device)
        nat.check(self._lib.sapx_solver_read(self._handle, int(which), nat._ptr(out), out.numel() * 8,
                                             nat.stream_ptr()), "sapx_solver_read")
        return out


__all__ = ["SolverSAP", "SapContactSolveResult", "SAP_CONTACT_PRESET_VARIANTS", "normalize_sap_line_search_mode"]

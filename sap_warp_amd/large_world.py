"""SAP Newton solve of one large world on the whole GPU (the native ``sapx_large_*`` entry points).

The per-world step kernels hold a world in one wave or one workgroup (up to 256 dofs).  A single
world of thousands of dofs -- a bin of free bodies, a robot among them -- is solved here with every
stage spread over the GPU: contact evaluation one thread per contact, gradient and line-search
products one wave per dof row, the Newton matrix assembled densely and factored by the blocked
Cholesky (``blocked_cholesky.py``), and every line-search trial run on the device.  The solve
mirrors the reference's ``SapContactSolve`` iteration (``contact_solve.py:6129-6285``) on the
problem a world's free motion and contact Jacobian produced (``contact_jacobian.py:51-84``).

Contact rows are passed sparse: each contact lists at most ``kmax`` dofs (those of the two
articulations it touches) with its three rows over them; :func:`sparse_contact_rows` converts
dense rows.  Tensors are torch CUDA tensors on the current stream.
"""
from __future__ import annotations

import ctypes
from dataclasses import dataclass

import numpy as np

from . import _native as nat
from .runtime import SapContacts, SapControl, SapState, allocate_contacts
from .solver_sap import SolverSAP

LINE_SEARCHES = {"armijo_decay": 0, "exact_root": 1}


def sparse_contact_rows(J: np.ndarray, kmax: int):
    """Dense contact rows (nc, 3, n) -> (cols (nc, kmax) int32, -1 padded; vals (nc, 3, kmax)
    float32).  Raises if a contact touches more than kmax dofs."""
    J = np.asarray(J)
    nc = J.shape[0]
    cols = np.full((nc, kmax), -1, dtype=np.int32)
    vals = np.zeros((nc, 3, kmax), dtype=np.float32)
    for c in range(nc):
        nz = np.nonzero(np.any(J[c] != 0.0, axis=0))[0]
        if len(nz) > kmax:
            raise ValueError(f"contact {c} touches {len(nz)} dofs > kmax {kmax}")
        cols[c, :len(nz)] = nz
        vals[c, :, :len(nz)] = J[c][:, nz]
    return cols, vals


@dataclass
class LargeSolveStats:
    iterations: int
    ls_iterations: int
    converged: bool
    optimality_reached: bool
    status: int
    cost: float
    grad_norm: float
    opt_tol: float


class LargeWorldSolver:
    """Workspace for worlds up to ``max_dofs`` dofs and ``max_contacts`` contacts of at most
    ``kmax`` dofs each."""

    def __init__(self, max_dofs: int, max_contacts: int, kmax: int = 64):
        self.lib = nat.load()
        self.max_dofs, self.max_contacts, self.kmax = int(max_dofs), int(max_contacts), int(kmax)
        h = ctypes.c_void_p()
        nat.check(self.lib.sapx_large_create(self.max_dofs, self.max_contacts, self.kmax, ctypes.byref(h)),
                  "sapx_large_create")
        self._h = h

    def close(self):
        if getattr(self, "_h", None) is not None and self._h.value:
            self.lib.sapx_large_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def solve(self, *, A, v_star, v_guess, participating, J_cols, J_vals, phi0, weight, stiffness, tau, mu, dt,
              model_terms: dict | None = None, line_search: str = "exact_root", max_iterations: int = 100,
              ls_max_iterations: int | None = None, refine: int = 1, optimality_abs_tol: float = 1.0e-14,
              optimality_rel_tol: float = 1.0e-6, cost_abs_tol: float = 1.0e-30, cost_rel_tol: float = 1.0e-15,
              armijo_c: float = 1.0e-4, rho: float = 0.8, relative_slop: float = 1000.0 * np.finfo(np.float64).eps,
              beta: float = 1.0, sigma: float = 1.0e-3, diag_shift: float = 0.0):
        """Returns (v (n,) float64 tensor, gamma (nc, 3) float64 tensor, LargeSolveStats).

        ``model_terms``: dict of per-dof device tensors pd_active, pd_a, pd_gain, pd_limit, lo_active,
        up_active, lo_vhat, up_vhat, lo_r, up_r (int32 for the masks, float64 otherwise), or None."""
        import torch

        if line_search not in LINE_SEARCHES:
            raise NotImplementedError(f"line search {line_search!r}: exact_root and armijo_decay only")
        n = int(A.shape[0])
        nc = int(phi0.shape[0]) if phi0 is not None else 0
        checks = dict(A=(A, torch.float32), v_star=(v_star, torch.float64), v_guess=(v_guess, torch.float64),
                      participating=(participating, torch.int32))
        if nc:
            checks.update(J_cols=(J_cols, torch.int32), J_vals=(J_vals, torch.float32), phi0=(phi0, torch.float64),
                          weight=(weight, torch.float64), stiffness=(stiffness, torch.float64),
                          tau=(tau, torch.float64), mu=(mu, torch.float64))
        for name, (t, dt_) in checks.items():
            if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == dt_ and t.is_contiguous()):
                raise TypeError(f"{name} must be a contiguous {dt_} CUDA tensor")
        if nc and (J_cols.shape != (nc, self.kmax) or J_vals.shape != (nc, 3, self.kmax)):
            raise ValueError("J_cols / J_vals must be (nc, kmax) / (nc, 3, kmax)")
        ptr = lambda t: None if t is None else ctypes.c_void_p(t.data_ptr())  # noqa: E731
        prob = nat.SapxLargeProblem()
        prob.n, prob.contact_count, prob.kmax = n, nc, self.kmax
        prob.A, prob.v_star, prob.v_guess, prob.participating = ptr(A), ptr(v_star), ptr(v_guess), ptr(participating)
        if nc:
            prob.J_cols, prob.J_vals = ptr(J_cols), ptr(J_vals)
            prob.phi0, prob.weight, prob.stiffness, prob.tau, prob.mu = (ptr(phi0), ptr(weight), ptr(stiffness),
                                                                         ptr(tau), ptr(mu))
        if model_terms is not None:
            for k in ("pd_active", "pd_a", "pd_gain", "pd_limit", "lo_active", "up_active", "lo_vhat", "up_vhat",
                      "lo_r", "up_r"):
                t = model_terms[k]
                want = torch.int32 if k.endswith("active") else torch.float64
                if not (isinstance(t, torch.Tensor) and t.is_cuda and t.dtype == want and t.numel() == n):
                    raise TypeError(f"model term {k} must be a {want} CUDA tensor of n values")
                setattr(prob, k, ptr(t))
        prob.dt = float(dt)
        o = nat.SapxLargeOpts()
        o.line_search = LINE_SEARCHES[line_search]
        o.max_iterations = int(max_iterations)
        o.ls_max_iterations = int(ls_max_iterations if ls_max_iterations is not None else
                                  (100 if line_search == "exact_root" else 40))
        o.refine = int(refine)
        o.optimality_abs_tol, o.optimality_rel_tol = optimality_abs_tol, optimality_rel_tol
        o.cost_abs_tol, o.cost_rel_tol = cost_abs_tol, cost_rel_tol
        o.armijo_c, o.rho, o.relative_slop = armijo_c, rho, relative_slop
        o.beta, o.sigma, o.diag_shift = beta, sigma, diag_shift
        v = torch.empty(n, dtype=torch.float64, device=A.device)
        gamma = torch.zeros((max(nc, 1), 3), dtype=torch.float64, device=A.device)
        st = nat.SapxLargeStats()
        stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
        nat.check(self.lib.sapx_large_solve(self._h, ctypes.byref(prob), ctypes.byref(o), ptr(v),
                                            ptr(gamma) if nc else None, ctypes.byref(st), stream), "sapx_large_solve")
        stats = LargeSolveStats(int(st.iterations), int(st.ls_iterations), bool(st.converged),
                                bool(st.optimality_reached), int(st.status), float(st.cost), float(st.grad_norm),
                                float(st.opt_tol))
        return v, gamma[:nc], stats


def stacked_boxes_problem(nx: int = 16, ny: int = 8, nz: int = 8, *, seed: int = 0, dt: float = 1.0e-3,
                          size: float = 0.1, stiffness: float = 1.0e5, tau: float = 0.01, mu: float = 0.7,
                          penetration: float = 1.0e-4, gravity: float = -9.81):
    """The SAP problem of one world holding a resting block of nx x ny x nz free cubes on the ground.

    Cubes of edge ``size`` (density U(500, 2000) kg/m^3, COM at the body origin, unrotated) sit on a
    touching lattice; every pair of face neighbours and every bottom cube against the ground
    touches at the four corners of the shared face (inset by 10 % of the edge), ``penetration``
    deep.  Starting at rest, v* = dt g for every body.  Returns numpy arrays: A_blocks (nb, 6, 6:
    each body's spatial inertia in SAP [w, v_O] order; A is their block diagonal), v_star,
    v_guess (zeros), participating, the contact rows sparse (J_cols (nc, 12): body 0's then body
    1's six dofs, -1 for the ground; J_vals (nc, 3, 12): rows t1, t2, n of the relative velocity
    of body 1 minus body 0 at the contact point), phi0, weight (the body_inertia contact weight,
    sap_helpers.py:307-338), stiffness, tau, mu.  1024 cubes (16 x 8 x 8) give 6144 dofs and
    11 520 contacts: the size of a settled bin of a thousand bodies."""
    rng = np.random.Generator(np.random.PCG64(seed))
    nb = nx * ny * nz
    n = 6 * nb
    h = 0.5 * size
    idx = lambda i, j, k: (k * ny + j) * nx + i  # noqa: E731
    pos = np.zeros((nb, 3))
    mass = np.zeros(nb)
    inert = np.zeros(nb)
    blocks = np.zeros((nb, 6, 6))
    for k in range(nz):
        for j in range(ny):
            for i in range(nx):
                b = idx(i, j, k)
                pos[b] = ((i - 0.5 * (nx - 1)) * size, (j - 0.5 * (ny - 1)) * size, h + k * size)
                m = rng.uniform(500.0, 2000.0) * size ** 3
                mass[b] = m
                inert[b] = m * size * size / 6.0
                blocks[b, :3, :3] = np.eye(3) * inert[b]
                blocks[b, 3:, 3:] = np.eye(3) * m
    v_star = np.zeros(n)
    v_star[5::6] = dt * gravity
    contacts = []  # (body0, body1, point, normal)
    inset = 0.4 * size
    axes = np.eye(3)
    for b in range(nb):
        k, rem = divmod(b, nx * ny)
        j, i = divmod(rem, nx)
        for ax, (di, dj, dk) in enumerate(((1, 0, 0), (0, 1, 0), (0, 0, 1))):
            ii, jj, kk = i + di, j + dj, k + dk
            if ii >= nx or jj >= ny or kk >= nz:
                continue
            nrm = axes[ax]
            u, w2 = axes[(ax + 1) % 3], axes[(ax + 2) % 3]
            centre = pos[b] + h * nrm
            for su in (-1.0, 1.0):
                for sw in (-1.0, 1.0):
                    contacts.append((b, idx(ii, jj, kk), centre + inset * (su * u + sw * w2), nrm))
        if k == 0:
            centre = pos[b] - h * axes[2]
            for su in (-1.0, 1.0):
                for sw in (-1.0, 1.0):
                    contacts.append((-1, b, centre + inset * (su * axes[0] + sw * axes[1]), axes[2]))
    nc = len(contacts)
    cols = np.full((nc, 12), -1, dtype=np.int32)
    vals = np.zeros((nc, 3, 12), dtype=np.float64)
    weight = np.zeros(nc)
    for c, (b0, b1, p, nrm) in enumerate(contacts):
        t1 = np.cross(nrm, axes[0] if abs(nrm[0]) < 0.9 else axes[1])
        t1 /= np.linalg.norm(t1)
        frame = (t1, np.cross(nrm, t1), nrm)
        wsum = 0.0
        for slot, (b, sgn) in enumerate(((b0, -1.0), (b1, 1.0))):
            if b < 0:
                continue
            cols[c, 6 * slot:6 * slot + 6] = np.arange(6 * b, 6 * b + 6)
            r = p - pos[b]
            for row, e in enumerate(frame):
                cb = np.cross(r, e)
                vals[c, row, 6 * slot:6 * slot + 3] = sgn * cb
                vals[c, row, 6 * slot + 3:6 * slot + 6] = sgn * e
                wsum += 1.0 / mass[b] + float(cb @ cb) / inert[b]
        weight[c] = max(wsum / 3.0, 1.0e-12)
    return dict(A_blocks=blocks, v_star=v_star, v_guess=np.zeros(n), participating=np.ones(n, dtype=np.int32),
                J_cols=cols, J_vals=vals, phi0=np.full(nc, -penetration), weight=weight,
                stiffness=np.full(nc, stiffness), tau=np.full(nc, tau), mu=np.full(nc, mu), dt=dt)


def dense_rows(cols: np.ndarray, vals: np.ndarray, n: int) -> np.ndarray:
    """Sparse contact rows back to dense (nc, 3, n)."""
    nc, kmax = cols.shape
    J = np.zeros((nc, 3, n), dtype=vals.dtype)
    for c in range(nc):
        for k in range(kmax):
            if cols[c, k] >= 0:
                J[c, :, cols[c, k]] += vals[c, :, k]
    return J


def block_diagonal(blocks: np.ndarray) -> np.ndarray:
    """(nb, 6, 6) blocks -> the dense (6 nb, 6 nb) block diagonal."""
    nb = blocks.shape[0]
    A = np.zeros((6 * nb, 6 * nb), dtype=blocks.dtype)
    for b in range(nb):
        A[6 * b:6 * b + 6, 6 * b:6 * b + 6] = blocks[b]
    return A


class BigWorldSolver(SolverSAP):
    """``SolverSAP`` plus ``SapCollisionPipeline.collide`` for ONE world of free rigid bodies
    spread over the whole GPU (the native ``sapx_bigworld_*`` entry points): a bin of a thousand
    boxes, spheres and capsules.

    Same keyword arguments and presets as :class:`SolverSAP` (``solver_sap.py:749-968``); the
    model must hold one world whose joints are all free joints rooted at the world with identity
    joint frames, one per body, without drives or limits, and the line search must be
    ``exact_root`` or ``armijo_decay``.  :meth:`collide` runs the model's explicit candidate pairs
    (``pipeline.py``: AABB test, narrowphase, per-world capacity rule), :meth:`step` one
    ``SolverSAP.step`` (``solver_sap.py:1388-1494``) with the Newton solve of
    :class:`LargeWorldSolver`, started from the previous step's solution after the first step
    (the warm start of ``SolverSAP``; :meth:`reset_runtime_state` clears it).
    """

    def _create_native(self, opts):
        if int(self.model.world_count) != 1:
            raise NotImplementedError("BigWorldSolver steps a model of exactly one world")
        h = ctypes.c_void_p()
        with nat.on_device(self._device):
            nat.check(self._lib.sapx_bigworld_create(self._native_model.handle, ctypes.byref(opts),
                                                     ctypes.byref(h)), "sapx_bigworld_create")
        self._stats = None
        return h

    def close(self) -> None:
        if getattr(self, "_handle", None):
            self._lib.sapx_bigworld_destroy(self._handle)
            self._handle = None

    def reset_runtime_state(self, state: SapState | None = None) -> None:
        """Clears the counters and the warm start (solver_sap.py:1004-1017)."""
        if getattr(self, "_handle", None):
            self._lib.sapx_bigworld_reset(self._handle)
        self.sim_time = 0.0
        self.frame_id = 0
        self.last_contact_count = 0
        self.last_solve_iterations = 0
        self.last_line_search_iterations = 0
        self.last_converged = True
        self.last_contacts = None
        self.last_contact_solve_result = None

    def allocate_contacts(self) -> SapContacts:
        return allocate_contacts(self.max_rigid_contact, 1, self._device)

    def collide(self, state: SapState, contacts: SapContacts) -> SapContacts:
        """Contacts of the model's explicit pairs at ``state.body_q`` (asynchronous)."""
        nat.check_contacts(contacts, self._device, 1, need_world_start=True)
        if nat.contact_capacity(contacts) < self.max_rigid_contact:
            raise ValueError("contact buffer smaller than max_rigid_contact")
        self._check_state(state, "state")
        cview = nat.contacts_view(contacts)
        with nat.on_device(self._device):
            nat.check(self._lib.sapx_bigworld_collide(self._handle, nat._ptr(state.body_q), ctypes.byref(cview),
                                                      nat.stream_ptr()), "sapx_bigworld_collide")
        return contacts

    def step(self, state_in: SapState, state_out: SapState, control: SapControl | None,
             contacts: SapContacts | None, dt, v_out=None, gamma_out=None) -> SapState:
        """One step; ``v_out`` (dofs, float64) and ``gamma_out`` (max_rigid_contact x 3, float64)
        optionally receive the SAP velocity and the contact impulses."""
        import torch

        if not isinstance(state_in, SapState) or not isinstance(state_out, SapState):
            raise TypeError("BigWorldSolver.step requires SapState input and output.")
        if isinstance(dt, torch.Tensor) or not float(dt) > 0.0:
            raise ValueError("dt must be a positive scalar")
        if contacts is None:
            contacts = allocate_contacts(1, 1, self._device)
        self._check_state(state_in, "state_in")
        self._check_state(state_out, "state_out")
        if control is not None:
            self._check_control(control)
        nat.check_contacts(contacts, self._device, 1)
        for t, shape, name in ((v_out, (self.dof_count,), "v_out"),
                               (gamma_out, (self.max_rigid_contact, 3), "gamma_out")):
            if t is not None and (t.dtype != torch.float64 or tuple(t.shape) != shape or not t.is_contiguous()
                                  or t.device != torch.device(self._device)):
                raise ValueError(f"{name} must be a contiguous float64 tensor of shape {shape} on {self._device}")
        sin = nat.SapxStateView(nat._ptr(state_in.joint_q), nat._ptr(state_in.joint_qd),
                                nat._ptr(state_in.body_q), nat._ptr(state_in.body_qd), nat._ptr(state_in.body_f))
        sout = nat.SapxStateView(nat._ptr(state_out.joint_q), nat._ptr(state_out.joint_qd),
                                 nat._ptr(state_out.body_q), nat._ptr(state_out.body_qd),
                                 nat._ptr(state_out.body_f))
        ctl = nat.SapxControlView()
        if control is not None:
            ctl = nat.SapxControlView(nat._ptr(control.joint_f), nat._ptr(control.joint_target_pos),
                                      nat._ptr(control.joint_target_vel), nat._ptr(control.joint_act))
        cview = nat.contacts_view(contacts)
        st = nat.SapxLargeStats()
        with nat.on_device(self._device):
            rc = self._lib.sapx_bigworld_step(self._handle, ctypes.byref(sin), ctypes.byref(sout),
                                              ctypes.byref(ctl), ctypes.byref(cview), float(dt), nat._ptr(v_out),
                                              nat._ptr(gamma_out), ctypes.byref(st), nat.stream_ptr())
        nat.check(rc, "sapx_bigworld_step")
        self._stats = LargeSolveStats(int(st.iterations), int(st.ls_iterations), bool(st.converged),
                                      bool(st.optimality_reached), int(st.status), float(st.cost),
                                      float(st.grad_norm), float(st.opt_tol))
        self.last_solve_iterations = self._stats.iterations
        self.last_line_search_iterations = self._stats.ls_iterations
        self.last_converged = self._stats.converged
        self.last_contacts = contacts
        self.sim_time += float(dt)
        self.frame_id += 1
        return state_out

    def solve_stats(self) -> LargeSolveStats | None:
        """The Newton record of the last step."""
        return self._stats

"""SAP Newton solve of one large world on the whole GPU (the native ``sapx_large_*`` entry points).

The per-world step kernels hold a world in one wave or one workgroup (up to 256 dofs).  A single
world of thousands of dofs -- a bin of free bodies, a robot among them -- is solved here with every
stage spread over the GPU: contact evaluation one thread per contact, gradient and line-search
products one wave per dof row, the Newton matrix assembled densely and factored by the blocked
Cholesky (``blocked_cholesky.py``), and every line-search trial run on the device.  The solve
mirrors the reference's ``SapContactSolver`` iteration (``contact_solve.py:6129-6285``) on the
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

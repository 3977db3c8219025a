"""Learned warm start for SolverSAP (no counterpart in the reference).

An MLP maps per-world state and contact features to a correction of the
solver's plain initial velocity (the previous step's solution, or v0 on the
first step).  The step kernel evaluates the SAP cost at both starts and keeps
the lower one, so the prediction can only help; the Newton solve still runs to
the unchanged tolerance, so the answer stays SAP's own.

Everything numeric runs in ``libsapx.so``: features (``sapx_features``), the
network's forward / backward / Adam kernels (``sapx_mlp_*``), the training
target (converged v minus the plain start, written by the step kernel through
``sapx_solver_set_warm_start_io``).  torch only holds the device buffers.

Network (the benchmark configuration): inputs padded to a multiple of 64,
three hidden layers of 512 ReLU units, outputs = dofs padded to 64; bfloat16
weights and activations with float32 accumulation.
"""
from __future__ import annotations

import ctypes
import time

import numpy as np

from . import _native as nat


def _round_up(x: int, m: int) -> int:
    return (x + m - 1) // m * m


class Mlp:
    """Handle of a library-owned MLP (``sapx_mlp_create``)."""

    def __init__(self, in_features: int, out_features: int, hidden=(512, 512, 512), seed: int = 0,
                 out_scale: float = 0.1):
        self._lib = nat.load()
        self.in_features = int(in_features)
        self.out_features = int(out_features)
        self.dims = [_round_up(self.in_features, 64), *[int(h) for h in hidden], _round_up(self.out_features, 64)]
        dims = (ctypes.c_int32 * len(self.dims))(*self.dims)
        h = ctypes.c_void_p()
        nat.check(self._lib.sapx_mlp_create(len(self.dims) - 1, dims, self.in_features, self.out_features,
                                            ctypes.byref(h)), "sapx_mlp_create")
        self._h = h
        self.param_count = int(self._lib.sapx_mlp_param_count(self._h))
        self.write_params(self.init_params(seed, out_scale))
        self.steps = 0

    def init_params(self, seed: int, out_scale: float) -> np.ndarray:
        """He-normal weights (zero on padded inputs / outputs; the last layer scaled by out_scale), zero biases."""
        rng = np.random.default_rng(seed)
        parts = []
        L = len(self.dims) - 1
        for l in range(L):
            n_in, n_out = self.dims[l], self.dims[l + 1]
            fan_in = self.in_features if l == 0 else n_in
            w = rng.standard_normal((n_out, n_in)) * np.sqrt(2.0 / fan_in)
            if l == 0:
                w[:, self.in_features:] = 0.0
            if l == L - 1:
                w *= out_scale
                w[self.out_features:, :] = 0.0
            parts.append(w.ravel())
            parts.append(np.zeros(n_out))
        flat = np.concatenate(parts).astype(np.float32)
        assert flat.size == self.param_count
        return flat

    def close(self):
        if getattr(self, "_h", None):
            self._lib.sapx_mlp_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def write_params(self, flat: np.ndarray):
        a = np.ascontiguousarray(flat, dtype=np.float32)
        nat.check(self._lib.sapx_mlp_write(self._h, nat.MLP_PARAMS, a.ctypes.data, a.size), "sapx_mlp_write")

    def read(self, which: int = nat.MLP_PARAMS) -> np.ndarray:
        out = np.zeros(self.param_count, dtype=np.float32)
        nat.check(self._lib.sapx_mlp_read(self._h, int(which), out.ctypes.data, out.size), "sapx_mlp_read")
        return out

    def set_normalization(self, x_mean, x_std, y_mean, y_std):
        for which, arr in ((nat.MLP_X_MEAN, x_mean), (nat.MLP_X_STD, x_std), (nat.MLP_Y_MEAN, y_mean),
                           (nat.MLP_Y_STD, y_std)):
            a = np.ascontiguousarray(arr, dtype=np.float32)
            nat.check(self._lib.sapx_mlp_write(self._h, which, a.ctypes.data, a.size), "sapx_mlp_write")

    def reserve(self, batch: int):
        nat.check(self._lib.sapx_mlp_reserve(self._h, int(batch)), "sapx_mlp_reserve")

    def _check_io(self, B, **tensors):
        import torch

        widths = dict(x=self.in_features, y=self.out_features, target=self.out_features, loss=1)
        for name, t in tensors.items():
            if t is None:
                continue
            if not isinstance(t, torch.Tensor) or not t.is_cuda or t.dtype != torch.float32 or not t.is_contiguous():
                raise TypeError(f"{name} must be a contiguous float32 GPU tensor")
            need = 1 if name == "loss" else B * widths[name]
            if t.numel() < need:
                raise ValueError(f"{name} holds {t.numel()} floats, the batch needs {need}")

    def forward(self, x, y, batch: int | None = None):
        """y[:B] = network(x[:B]) (device float32 tensors, row-major)."""
        B = int(batch if batch is not None else x.shape[0])
        self._check_io(B, x=x, y=y)
        nat.check(self._lib.sapx_mlp_forward(self._h, nat._ptr(x), B, nat._ptr(y), nat.stream_ptr()),
                  "sapx_mlp_forward")

    def gradients(self, x, target, loss=None, batch: int | None = None):
        B = int(batch if batch is not None else x.shape[0])
        self._check_io(B, x=x, target=target, loss=loss)
        nat.check(self._lib.sapx_mlp_gradients(self._h, nat._ptr(x), nat._ptr(target), B, nat._ptr(loss),
                                               nat.stream_ptr()), "sapx_mlp_gradients")

    def adam(self, lr: float, beta1: float = 0.9, beta2: float = 0.999, eps: float = 1e-8):
        nat.check(self._lib.sapx_mlp_adam(self._h, float(lr), float(beta1), float(beta2), float(eps),
                                          nat.stream_ptr()), "sapx_mlp_adam")
        self.steps += 1


class LearnedWarmStart:
    """Predicted start for ``solver`` (a wave-kernel SolverSAP) from an :class:`Mlp`.

    Usage per step (after collision)::

        lws.step(state_in, state_out, control, contacts, dt)

    which computes the features, runs the network and steps the solver from the
    cheaper of the two starts.  ``collect`` / ``fit`` build the training set
    from plain rollouts and train the network on the GPU.
    """

    def __init__(self, solver, hidden=(512, 512, 512), seed: int = 0, mlp: Mlp | None = None):
        import torch

        self.solver = solver
        self._lib = nat.load()
        self.W = solver.num_envs
        self.n = solver.dof_per_env
        self.F = int(self._lib.sapx_solver_feature_count(solver._handle))
        self.mlp = mlp if mlp is not None else Mlp(self.F, self.n, hidden=hidden, seed=seed)
        dev = solver.model.device
        self.features_buf = torch.zeros((self.W, self.F), dtype=torch.float32, device=dev)
        self.delta = torch.zeros((self.W, self.n), dtype=torch.float32, device=dev)
        # the step writes this step's correction into `target`; `prev` holds the previous step's
        # (an input feature); the two swap after every step
        self.target = torch.zeros((self.W, self.n), dtype=torch.float32, device=dev)
        self.prev = torch.zeros((self.W, self.n), dtype=torch.float32, device=dev)
        self.last_target = self.prev
        self.mlp.reserve(self.W)
        self.trained = False

    # ------------------------------------------------------------------ plumbing
    def compute_features(self, state, control, contacts):
        self.solver._check_state(state, "state")
        if control is not None:
            self.solver._check_control(control)
        nat.check_contacts(contacts, self.solver._device, self.W, need_world_start=True)
        sin = nat.SapxStateView(nat._ptr(state.joint_q), nat._ptr(state.joint_qd), nat._ptr(state.body_q),
                                nat._ptr(state.body_qd), nat._ptr(state.body_f))
        ctl = nat.SapxControlView()
        if control is not None:
            ctl = nat.SapxControlView(nat._ptr(control.joint_f), nat._ptr(control.joint_target_pos),
                                      nat._ptr(control.joint_target_vel), nat._ptr(control.joint_act))
        cview = nat.contacts_view(contacts)
        nat.check(self._lib.sapx_features(self.solver._handle, ctypes.byref(sin), ctypes.byref(ctl),
                                          ctypes.byref(cview), nat._ptr(self.prev), nat._ptr(self.features_buf),
                                          nat.stream_ptr()), "sapx_features")
        return self.features_buf

    def reset(self):
        """Start of an episode: no previous correction."""
        self.prev.zero_()

    def set_io(self, correction=None, target=None):
        nat.check(self._lib.sapx_solver_set_warm_start_io(self.solver._handle, nat._ptr(correction),
                                                          nat._ptr(target)), "sapx_solver_set_warm_start_io")

    # ------------------------------------------------------------------ stepping
    def prepare(self, state_in, control, contacts, learned: bool = True):
        """Features, network and solver IO for the next ``solver.step`` (enqueued on the current stream)."""
        self.compute_features(state_in, control, contacts)
        if learned:
            self.mlp.forward(self.features_buf, self.delta)
        self.set_io(self.delta if learned else None, self.target)

    def finish(self):
        """After the step: this step's correction becomes the next step's input feature."""
        self.set_io(None, None)
        self.prev, self.target = self.target, self.prev
        self.last_target = self.prev

    def step(self, state_in, state_out, control, contacts, dt):
        """One solver step from the learned start (falls back to the plain start where it costs more)."""
        self.prepare(state_in, control, contacts, learned=True)
        try:
            return self.solver.step(state_in, state_out, control, contacts, dt)
        finally:
            self.finish()

    def plain_step_with_target(self, state_in, state_out, control, contacts, dt):
        """Plain step that also records (features, converged v - plain start) for training."""
        self.prepare(state_in, control, contacts, learned=False)
        try:
            return self.solver.step(state_in, state_out, control, contacts, dt)
        finally:
            self.finish()

    # ------------------------------------------------------------------ training
    def collect(self, pipeline, state0, state1, control, contacts, steps: int, dt: float, control_fn=None,
                out_x=None, out_y=None, offset: int = 0):
        """Plain rollout of ``steps`` steps; rows [offset, offset + steps * W) of out_x / out_y get the samples."""
        import torch

        W = self.W
        if out_x is None:
            out_x = torch.empty((steps * W, self.F), dtype=torch.float32, device=self.features_buf.device)
            out_y = torch.empty((steps * W, self.n), dtype=torch.float32, device=self.features_buf.device)
        s0, s1 = state0, state1
        for k in range(steps):
            if control_fn is not None:
                control_fn(k)
            pipeline.collide(s0, contacts)
            self.plain_step_with_target(s0, s1, control, contacts, dt)
            r0 = offset + k * W
            out_x[r0:r0 + W].copy_(self.features_buf)
            out_y[r0:r0 + W].copy_(self.last_target)
            s0, s1 = s1, s0
        return out_x, out_y, s0, s1

    def fit(self, x, y, epochs: int = 4, batch: int = 4096, lr: float = 1e-3, lr_final: float = 1e-4,
            seed: int = 0, log_every: int = 0):
        """Train on (x, y) device tensors: normalisation from the data, Adam with a cosine-decayed rate."""
        import torch

        S = int(x.shape[0])
        # normalisation statistics from a host sample of the data set
        rng = np.random.default_rng(seed)
        sample = rng.choice(S, size=min(S, 200_000), replace=False)
        idx = torch.as_tensor(np.sort(sample), device=x.device)
        xs = x.index_select(0, idx).cpu().numpy().astype(np.float64)
        ys = y.index_select(0, idx).cpu().numpy().astype(np.float64)
        x_mean, x_std = xs.mean(axis=0), xs.std(axis=0)
        y_mean, y_std = ys.mean(axis=0), ys.std(axis=0)
        x_std = np.where(x_std < 1e-3, 1.0, x_std)
        y_std = np.where(y_std < 1e-6, 1.0, y_std)
        self.mlp.set_normalization(x_mean, x_std, y_mean, y_std)
        self.mlp.reserve(batch)
        nb = S // batch
        total = max(1, epochs * nb)
        loss = torch.zeros(1, dtype=torch.float32, device=x.device)
        gen = torch.Generator(device=x.device)
        gen.manual_seed(seed)
        xb = torch.empty((batch, x.shape[1]), dtype=torch.float32, device=x.device)
        yb = torch.empty((batch, y.shape[1]), dtype=torch.float32, device=x.device)
        history = []
        t0 = time.perf_counter()
        it = 0
        for ep in range(epochs):
            perm = torch.randperm(S, generator=gen, device=x.device)
            for bi in range(nb):
                sel = perm[bi * batch:(bi + 1) * batch]
                torch.index_select(x, 0, sel, out=xb)
                torch.index_select(y, 0, sel, out=yb)
                frac = it / total
                rate = lr_final + 0.5 * (lr - lr_final) * (1.0 + np.cos(np.pi * frac))
                self.mlp.gradients(xb, yb, loss)
                self.mlp.adam(rate)
                it += 1
                if log_every and it % log_every == 0:
                    history.append((it, float(loss.item())))
        torch.cuda.synchronize()
        self.trained = True
        self.train_seconds = time.perf_counter() - t0
        self.train_steps = it
        self.final_loss = float(loss.item())
        return history

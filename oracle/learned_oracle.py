"""CPU restatement of the learned warm start (test infrastructure only).

The learned warm start has no counterpart in the reference: it is the
north-star addition on top of ``SolverSAP`` (an MLP that predicts a correction
to the solver's plain initial velocity, trained on (state, converged v) pairs
the solver generates).  This module is the oracle the HIP kernels are checked
against:

* ``features``: per-world input features from public state, control, the
  contact list and the previous step's correction (yaw-invariant: free-joint
  quantities in the joint's body frame);
* ``local_delta`` / ``apply_delta``: the output convention (generalized
  velocity correction, free-joint rows in the body frame);
* ``Mlp``: forward pass, mean-squared-error loss, backward pass and Adam with
  the same bfloat16 rounding points as the kernels (inputs, weights and hidden
  activations rounded to bfloat16; products accumulated in float32 or wider).

Only tests, the smoke check and the benchmark's CPU-baseline leg import it.
Pinned by tests/test_oracle_learned.py: backward vs central finite differences
(float64, rounding off), Adam's first step vs its closed form, feature
invariance under horizontal translation and yaw rotation, and the output
round trip local_delta -> apply_delta.
"""
from __future__ import annotations

from dataclasses import dataclass, field

import numpy as np

JOINT_PRISMATIC = 0
JOINT_REVOLUTE = 1
JOINT_FIXED = 3
JOINT_FREE = 4
JOINT_DISTANCE = 5

PHI_CLIP = 0.05          # signed distances are clipped to +-5 cm; bodies without contacts read +PHI_CLIP
FEATURES_PER_FREE = 10   # height, gravity direction (3), angular (3) and linear (3) velocity in the body frame
FEATURES_PER_DOF = 4     # q, qd, target_pos - q, target_vel - qd
FEATURES_PER_BODY = 5    # contact count, min signed distance, summed contact normal (3)


def round_up(x: int, m: int) -> int:
    return (x + m - 1) // m * m


# =============================================================================
# bfloat16
# =============================================================================
def bf16_round(x):
    """Round float32 values to the nearest bfloat16 (ties to even), returned as float32."""
    a = np.ascontiguousarray(np.asarray(x, dtype=np.float32))
    u = a.view(np.uint32).astype(np.uint64)
    r = ((u + 0x7FFF + ((u >> 16) & 1)) & 0xFFFF0000).astype(np.uint32)
    out = r.view(np.float32).copy()
    nan = np.isnan(a)
    out[nan] = np.nan
    return out


# =============================================================================
# features / outputs
# =============================================================================
def _quat_matrix(q):
    """Rotation matrix of quaternion (x, y, z, w) after normalisation."""
    x, y, z, w = np.asarray(q, dtype=np.float64) / max(np.linalg.norm(q), 1e-30)
    return np.array([
        [1 - 2 * (y * y + z * z), 2 * (x * y - z * w), 2 * (x * z + y * w)],
        [2 * (x * y + z * w), 1 - 2 * (x * x + z * z), 2 * (y * z - x * w)],
        [2 * (x * z - y * w), 2 * (y * z + x * w), 1 - 2 * (x * x + y * y)],
    ])


def feature_count(model) -> int:
    jpw = model.joint_count // model.world_count
    nfree = 0
    ndof = 0
    for j in range(jpw):
        jt = int(model.joint_type[j])
        nd = int(model.joint_qd_start[j + 1] - model.joint_qd_start[j])
        if jt in (JOINT_FREE, JOINT_DISTANCE):
            nfree += 1
        else:
            ndof += nd
    return (FEATURES_PER_FREE * nfree + FEATURES_PER_DOF * ndof + FEATURES_PER_BODY * model.bodies_per_env
            + model.dof_per_env)


def _world_frame_rotation(model, w, joint_q):
    """Rotation of the world's first free joint (identity if it has none): frame of the contact normals."""
    jpw = model.joint_count // model.world_count
    for jl in range(jpw):
        j = w * jpw + jl
        if int(model.joint_type[j]) in (JOINT_FREE, JOINT_DISTANCE):
            qs = int(model.joint_q_start[j])
            return _quat_matrix(joint_q[qs + 3:qs + 7])
    return np.eye(3)


def features(model, w, state: dict, control: dict, contacts: dict, prev_target=None) -> np.ndarray:
    """Input features of world ``w`` (float64), in the kernel's order.

    ``prev_target`` (world_count x dofs, or None = zeros) is the previous step's
    correction in the output convention (``local_delta``), appended last.
    """
    jq = np.asarray(state["joint_q"], dtype=np.float64)
    jqd = np.asarray(state["joint_qd"], dtype=np.float64)
    bq = np.asarray(state["body_q"], dtype=np.float64)
    tpos = np.asarray(control["joint_target_pos"], dtype=np.float64)
    tvel = np.asarray(control["joint_target_vel"], dtype=np.float64)
    jpw = model.joint_count // model.world_count
    nb = model.bodies_per_env
    out = []
    for jl in range(jpw):
        j = w * jpw + jl
        jt = int(model.joint_type[j])
        qs, ds = int(model.joint_q_start[j]), int(model.joint_qd_start[j])
        nd = int(model.joint_qd_start[j + 1]) - ds
        if jt in (JOINT_FREE, JOINT_DISTANCE):
            R = _quat_matrix(jq[qs + 3:qs + 7])
            out.append(jq[qs + 2])
            out.extend(R.T @ np.array([0.0, 0.0, 1.0]))
            out.extend(R.T @ jqd[ds:ds + 3])
            out.extend(R.T @ jqd[ds + 3:ds + 6])
        else:
            for k in range(nd):
                q = jq[qs + k]
                out.extend([q, jqd[ds + k], tpos[ds + k] - q, tvel[ds + k] - jqd[ds + k]])
    cnt = np.zeros(nb)
    phimin = np.full(nb, PHI_CLIP)
    nsum = np.zeros((nb, 3))
    ws = contacts["world_start"]
    b0w = w * nb
    for c in range(int(ws[w]), int(ws[w + 1])):
        s0, s1 = int(contacts["shape0"][c]), int(contacts["shape1"][c])
        g0 = int(model.shape_body[s0]) if s0 >= 0 else -1
        g1 = int(model.shape_body[s1]) if s1 >= 0 else -1
        p0 = np.asarray(contacts["point0"][c], dtype=np.float64)
        p1 = np.asarray(contacts["point1"][c], dtype=np.float64)
        if g0 >= 0:
            p0 = bq[g0, :3] + _quat_matrix(bq[g0, 3:7]) @ p0
        if g1 >= 0:
            p1 = bq[g1, :3] + _quat_matrix(bq[g1, 3:7]) @ p1
        n = np.asarray(contacts["normal"][c], dtype=np.float64)
        phi = float((p1 - p0) @ n) - float(contacts["margin0"][c]) - float(contacts["margin1"][c])
        phi = min(max(phi, -PHI_CLIP), PHI_CLIP)
        for g, sgn in ((g0, 1.0), (g1, -1.0)):
            if g < 0:
                continue
            b = g - b0w
            cnt[b] += 1.0
            phimin[b] = min(phimin[b], phi)
            nsum[b] += sgn * n
    Rw = _world_frame_rotation(model, w, jq)
    for b in range(nb):
        out.append(cnt[b])
        out.append(phimin[b])
        out.extend(Rw.T @ nsum[b])
    n = model.dof_per_env
    out.extend(np.zeros(n) if prev_target is None else np.asarray(prev_target[w], dtype=np.float64))
    return np.asarray(out, dtype=np.float64)


def _free_blocks(model, w):
    jpw = model.joint_count // model.world_count
    d0 = int(model.joint_qd_start[w * jpw])
    for jl in range(jpw):
        j = w * jpw + jl
        if int(model.joint_type[j]) in (JOINT_FREE, JOINT_DISTANCE):
            yield int(model.joint_q_start[j]), int(model.joint_qd_start[j]) - d0


def local_delta(model, w, joint_q, v_delta) -> np.ndarray:
    """Express a world's generalized velocity change with free-joint rows in the joint's body frame."""
    jq = np.asarray(joint_q, dtype=np.float64)
    out = np.asarray(v_delta, dtype=np.float64).copy()
    for qs, dl in _free_blocks(model, w):
        R = _quat_matrix(jq[qs + 3:qs + 7])
        out[dl:dl + 3] = R.T @ out[dl:dl + 3]
        out[dl + 3:dl + 6] = R.T @ out[dl + 3:dl + 6]
    return out


def apply_delta(model, w, joint_q, v_plain, delta_local) -> np.ndarray:
    """Predicted start: plain start plus the correction rotated back to the world frame."""
    jq = np.asarray(joint_q, dtype=np.float64)
    d = np.asarray(delta_local, dtype=np.float64).copy()
    for qs, dl in _free_blocks(model, w):
        R = _quat_matrix(jq[qs + 3:qs + 7])
        d[dl:dl + 3] = R @ d[dl:dl + 3]
        d[dl + 3:dl + 6] = R @ d[dl + 3:dl + 6]
    return np.asarray(v_plain, dtype=np.float64) + d


# =============================================================================
# MLP
# =============================================================================
@dataclass
class Mlp:
    """Fully connected ReLU network; weights W[l] are (out, in), biases b[l] (out,).

    ``dims`` are the padded widths (input a multiple of 32, hidden / output multiples
    of 64); ``in_features`` / ``out_features`` are the real widths.  Normalisation:
    x_n = (x - x_mean) / x_std, targets y_n = (y - y_mean) / y_std.
    """

    dims: list
    in_features: int
    out_features: int
    W: list = field(default_factory=list)
    b: list = field(default_factory=list)
    x_mean: np.ndarray = None
    x_std: np.ndarray = None
    y_mean: np.ndarray = None
    y_std: np.ndarray = None
    rounding: bool = True

    def _r(self, x):
        return bf16_round(x).astype(np.float64) if self.rounding else np.asarray(x, dtype=np.float64)

    def normalized_input(self, x):
        B = x.shape[0]
        xn = np.zeros((B, self.dims[0]))
        xn[:, :self.in_features] = (np.asarray(x, dtype=np.float64) - self.x_mean) / self.x_std
        return self._r(xn.astype(np.float32) if self.rounding else xn)

    def forward(self, x):
        """Normalized output (B, dims[-1]) and the cache of rounded activations."""
        h = self.normalized_input(x)
        cache = [h]
        L = len(self.W)
        for l in range(L):
            Wr = self._r(np.asarray(self.W[l], dtype=np.float32) if self.rounding else self.W[l])
            z = h @ Wr.T + np.asarray(self.b[l], dtype=np.float64)
            if l < L - 1:
                a = np.maximum(z, 0.0)
                h = self._r(a.astype(np.float32) if self.rounding else a)
                cache.append(h)
            else:
                h = z
        return h, cache

    def predict(self, x):
        y, _ = self.forward(x)
        return y[:, :self.out_features] * self.y_std + self.y_mean

    def loss_and_grads(self, x, target):
        """Mean squared error over the real outputs in normalized units, and its gradients."""
        y, cache = self.forward(x)
        B = x.shape[0]
        No = self.out_features
        t = (np.asarray(target, dtype=np.float64) - self.y_mean) / self.y_std
        diff = np.zeros_like(y)
        diff[:, :No] = y[:, :No] - t
        loss = float((diff[:, :No] ** 2).sum() / (B * No))
        dz = 2.0 * diff / (B * No)
        L = len(self.W)
        gW = [None] * L
        gb = [None] * L
        for l in range(L - 1, -1, -1):
            dzr = self._r(dz.astype(np.float32) if self.rounding else dz)
            gW[l] = dzr.T @ cache[l]
            gb[l] = dz.sum(axis=0)
            if l > 0:
                Wr = self._r(np.asarray(self.W[l], dtype=np.float32) if self.rounding else self.W[l])
                dh = dzr @ Wr
                dz = dh * (cache[l] > 0.0)
        return loss, gW, gb


@dataclass
class Adam:
    lr: float = 1e-3
    beta1: float = 0.9
    beta2: float = 0.999
    eps: float = 1e-8
    t: int = 0
    m: list = None
    v: list = None

    def step(self, params: list, grads: list):
        """In-place update of ``params`` (list of float arrays) with bias-corrected moments."""
        if self.m is None:
            self.m = [np.zeros_like(p, dtype=np.float64) for p in params]
            self.v = [np.zeros_like(p, dtype=np.float64) for p in params]
        self.t += 1
        c1 = 1.0 - self.beta1 ** self.t
        c2 = 1.0 - self.beta2 ** self.t
        for p, g, m, v in zip(params, grads, self.m, self.v):
            m *= self.beta1
            m += (1.0 - self.beta1) * g
            v *= self.beta2
            v += (1.0 - self.beta2) * g * g
            p -= self.lr * (m / c1) / (np.sqrt(v / c2) + self.eps)


def init_mlp(dims, in_features, out_features, seed=0, out_scale=0.1) -> Mlp:
    """He-normal weights (zero on padded inputs / outputs), zero biases, identity normalisation."""
    rng = np.random.default_rng(seed)
    W, b = [], []
    L = len(dims) - 1
    for l in range(L):
        fan_in = in_features if l == 0 else dims[l]
        w = rng.standard_normal((dims[l + 1], dims[l])) * np.sqrt(2.0 / fan_in)
        if l == 0:
            w[:, in_features:] = 0.0
        if l == L - 1:
            w *= out_scale
            w[out_features:, :] = 0.0
        W.append(w.astype(np.float32).astype(np.float64))
        b.append(np.zeros(dims[l + 1]))
    return Mlp(list(dims), in_features, out_features, W, b, np.zeros(in_features), np.ones(in_features),
               np.zeros(out_features), np.ones(out_features))


def flat_params(mlp: Mlp) -> np.ndarray:
    """Parameters in the library's flat order: for each layer W (out x in, row-major) then b."""
    parts = []
    for W, b in zip(mlp.W, mlp.b):
        parts.append(np.asarray(W, dtype=np.float64).ravel())
        parts.append(np.asarray(b, dtype=np.float64).ravel())
    return np.concatenate(parts)


def unflat(mlp: Mlp, flat: np.ndarray):
    """Split a flat parameter-shaped vector into per-layer (W, b) lists."""
    Ws, bs = [], []
    o = 0
    for l in range(len(mlp.dims) - 1):
        n_out, n_in = mlp.dims[l + 1], mlp.dims[l]
        Ws.append(np.asarray(flat[o:o + n_out * n_in], dtype=np.float64).reshape(n_out, n_in))
        o += n_out * n_in
        bs.append(np.asarray(flat[o:o + n_out], dtype=np.float64))
        o += n_out
    return Ws, bs

"""ctypes binding of ``libsapx.so`` (the C ABI declared in ``include/sapx.h``).

The library is built in-tree by ``sap_warp_amd/build.py``.  There is no
fallback: if the library is missing or does not load, every native entry
point raises ``RuntimeError``.
"""
from __future__ import annotations

import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "libsapx.so")

SAPX_OK = 0
SAPX_ERR_INVALID = -1
SAPX_ERR_UNSUPPORTED = -2
SAPX_ERR_HIP = -3
SAPX_ERR_CAPACITY = -4

LS_ARMIJO_DECAY = 0
LS_EXACT_ROOT = 1
LS_MONOTONE_DECAY = 2
WEIGHT_BODY_INERTIA = 0
WEIGHT_DIAG_DELASSUS = 1
POINT_WITNESS = 0
POINT_MIDPOINT = 1
INTEGRATE_MIDPOINT = 0
INTEGRATE_SAP_EULER = 1

# buffer ids of sapx_solver_read
READ_V = 0
READ_V_STAR = 1
READ_V0 = 2
READ_GAMMA = 3
READ_PHI0 = 4
READ_WEIGHT = 5
READ_A_DIAG = 6
READ_A = 7
READ_J = 8

_P = ctypes.c_void_p
_I = ctypes.c_int32
_D = ctypes.c_double


class SapxModelDesc(ctypes.Structure):
    _fields_ = [
        ("world_count", _I), ("joint_count", _I), ("dof_count", _I), ("coord_count", _I),
        ("body_count", _I), ("shape_count", _I), ("articulation_count", _I), ("pair_count", _I),
        ("joint_type", _P), ("joint_parent", _P), ("joint_child", _P), ("joint_q_start", _P),
        ("joint_qd_start", _P), ("joint_dof_dim", _P), ("joint_ancestor", _P), ("articulation_start", _P),
        ("joint_axis", _P), ("joint_X_p", _P), ("joint_X_c", _P), ("joint_armature", _P),
        ("body_mass", _P), ("body_inertia", _P), ("body_com", _P), ("gravity", _P),
        ("joint_target_mode", _P), ("joint_target_ke", _P), ("joint_target_kd", _P), ("joint_effort_limit", _P),
        ("joint_limit_lower", _P), ("joint_limit_upper", _P), ("joint_limit_ke", _P), ("joint_limit_kd", _P),
        ("shape_body", _P), ("shape_type", _P), ("shape_world", _P), ("shape_transform", _P),
        ("shape_scale", _P), ("shape_margin", _P), ("shape_gap", _P), ("shape_collision_radius", _P),
        ("shape_material_mu", _P), ("shape_material_ke", _P), ("shape_material_tau", _P), ("shape_pairs", _P),
    ]


class SapxSolverOpts(ctypes.Structure):
    _fields_ = [
        ("max_rigid_contact", _I),
        ("fallback_mu", _D), ("fallback_stiffness", _D), ("fallback_tau_d", _D),
        ("contact_beta", _D), ("contact_sigma", _D), ("diag_shift", _D),
        ("max_iterations", _I),
        ("optimality_abs_tol", _D), ("optimality_rel_tol", _D), ("cost_abs_tol", _D), ("cost_rel_tol", _D),
        ("line_search", _I), ("line_search_max_iterations", _I),
        ("armijo_c", _D), ("rho", _D), ("line_search_relative_slop", _D),
        ("contact_weight_mode", _I), ("contact_point_mode", _I), ("position_integration", _I),
        ("capture_snapshots", _I),
    ]


class SapxStateView(ctypes.Structure):
    _fields_ = [("joint_q", _P), ("joint_qd", _P), ("body_q", _P), ("body_qd", _P), ("body_f", _P)]


class SapxControlView(ctypes.Structure):
    _fields_ = [("joint_f", _P), ("joint_target_pos", _P), ("joint_target_vel", _P), ("joint_act", _P)]


class SapxContactsView(ctypes.Structure):
    _fields_ = [
        ("count", _P), ("world_start", _P), ("shape0", _P), ("shape1", _P), ("point0", _P), ("point1", _P),
        ("normal", _P), ("margin0", _P), ("margin1", _P), ("capacity", _I),
    ]


class SapxSchedule(ctypes.Structure):
    _fields_ = [("kernel", _I), ("small_capacity", _I), ("fork", _I), ("world_order", _I)]


class SapxLargeProblem(ctypes.Structure):
    _fields_ = [("n", _I), ("contact_count", _I), ("kmax", _I), ("pad_", _I),
                ("A", _P), ("v_star", _P), ("v_guess", _P), ("participating", _P), ("J_cols", _P), ("J_vals", _P),
                ("phi0", _P), ("weight", _P), ("stiffness", _P), ("tau", _P), ("mu", _P),
                ("pd_active", _P), ("pd_a", _P), ("pd_gain", _P), ("pd_limit", _P), ("lo_active", _P),
                ("up_active", _P), ("lo_vhat", _P), ("up_vhat", _P), ("lo_r", _P), ("up_r", _P), ("dt", _D)]


class SapxLargeOpts(ctypes.Structure):
    _fields_ = [("line_search", _I), ("max_iterations", _I), ("ls_max_iterations", _I), ("refine", _I),
                ("optimality_abs_tol", _D), ("optimality_rel_tol", _D), ("cost_abs_tol", _D), ("cost_rel_tol", _D),
                ("armijo_c", _D), ("rho", _D), ("relative_slop", _D), ("beta", _D), ("sigma", _D),
                ("diag_shift", _D)]


class SapxLargeStats(ctypes.Structure):
    _fields_ = [("iterations", _I), ("ls_iterations", _I), ("converged", _I), ("optimality_reached", _I),
                ("status", _I), ("pad_", _I), ("cost", _D), ("grad_norm", _D), ("opt_tol", _D)]


KERNEL_WAVE = 0
KERNEL_FUSED = 1

# declared entry points: name -> (restype, argtypes); tests check the export table
EXPORTS = {
    "sapx_version": (_I, []),
    "sapx_last_error": (ctypes.c_char_p, []),
    "sapx_model_create": (_I, [ctypes.POINTER(SapxModelDesc), ctypes.POINTER(_P)]),
    "sapx_model_destroy": (None, [_P]),
    "sapx_solver_create": (_I, [_P, ctypes.POINTER(SapxSolverOpts), ctypes.POINTER(_P)]),
    "sapx_solver_destroy": (None, [_P]),
    "sapx_solver_reset": (_I, [_P, _P]),
    "sapx_step": (_I, [_P, ctypes.POINTER(SapxStateView), ctypes.POINTER(SapxStateView),
                       ctypes.POINTER(SapxControlView), ctypes.POINTER(SapxContactsView), _D, _P, _P]),
    "sapx_get_stats": (_I, [_P, _P, _P, _P, _P, _P]),
    "sapx_get_histogram": (_I, [_P, _P, _I, _P]),
    "sapx_solver_read": (_I, [_P, _I, _P, ctypes.c_int64, _P]),
    "sapx_pipeline_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "sapx_pipeline_destroy": (None, [_P]),
    "sapx_pipeline_create_all_pairs": (_I, [_P, _I, _P, _P, _P, _I, ctypes.POINTER(_P)]),
    "sapx_collide": (_I, [_P, _P, ctypes.POINTER(SapxContactsView), _P]),
    "sapx_collide_noalloc": (_I, [_P, _P, ctypes.POINTER(SapxContactsView), _P]),
    "sapx_solver_feature_count": (_I, [_P]),
    "sapx_features": (_I, [_P, ctypes.POINTER(SapxStateView), ctypes.POINTER(SapxControlView),
                           ctypes.POINTER(SapxContactsView), _P, _P, _P]),
    "sapx_solver_set_warm_start_io": (_I, [_P, _P, _P]),
    "sapx_solver_get_schedule": (_I, [_P, ctypes.POINTER(SapxSchedule)]),
    "sapx_solver_set_schedule": (_I, [_P, ctypes.POINTER(SapxSchedule)]),
    "sapx_batch_create": (_I, [_P, _I, ctypes.POINTER(_P)]),
    "sapx_batch_destroy": (None, [_P]),
    "sapx_batch_set_schedule": (_I, [_P, _I, _I]),
    "sapx_batch_bind": (_I, [_P, _I, ctypes.POINTER(SapxStateView), ctypes.POINTER(SapxStateView),
                             ctypes.POINTER(SapxControlView), ctypes.POINTER(SapxContactsView), _D]),
    "sapx_batch_step": (_I, [_P, _I, _P]),
    "sapx_mlp_create": (_I, [_I, _P, _I, _I, ctypes.POINTER(_P)]),
    "sapx_mlp_destroy": (None, [_P]),
    "sapx_mlp_param_count": (ctypes.c_int64, [_P]),
    "sapx_mlp_write": (_I, [_P, _I, _P, ctypes.c_int64]),
    "sapx_mlp_read": (_I, [_P, _I, _P, ctypes.c_int64]),
    "sapx_mlp_reserve": (_I, [_P, _I]),
    "sapx_mlp_forward": (_I, [_P, _P, _I, _P, _P]),
    "sapx_mlp_gradients": (_I, [_P, _P, _P, _I, _P, _P]),
    "sapx_mlp_adam": (_I, [_P, ctypes.c_float, ctypes.c_float, ctypes.c_float, ctypes.c_float, _P]),
    "sapx_chol_create": (_I, [_I, _I, ctypes.POINTER(_P)]),
    "sapx_chol_destroy": (None, [_P]),
    "sapx_chol_factorize": (_I, [_P, _P, _I, _I, ctypes.c_int64, _P]),
    "sapx_chol_solve": (_I, [_P, _P, _P, _I, _P]),
    "sapx_chol_info": (_I, [_P, _P]),
    "sapx_large_create": (_I, [_I, _I, _I, ctypes.POINTER(_P)]),
    "sapx_large_destroy": (None, [_P]),
    "sapx_large_solve": (_I, [_P, ctypes.POINTER(SapxLargeProblem), ctypes.POINTER(SapxLargeOpts), _P, _P,
                              ctypes.POINTER(SapxLargeStats), _P]),
    "sapx_bigworld_create": (_I, [_P, ctypes.POINTER(SapxSolverOpts), ctypes.POINTER(_P)]),
    "sapx_bigworld_destroy": (None, [_P]),
    "sapx_bigworld_reset": (None, [_P]),
    "sapx_bigworld_collide": (_I, [_P, _P, ctypes.POINTER(SapxContactsView), _P]),
    "sapx_bigworld_step": (_I, [_P, ctypes.POINTER(SapxStateView), ctypes.POINTER(SapxStateView),
                                ctypes.POINTER(SapxControlView), ctypes.POINTER(SapxContactsView), ctypes.c_float,
                                _P, _P, ctypes.POINTER(SapxLargeStats), _P]),
}

HIST_SIZE = 260  # sapx_get_histogram entries
MLP_PARAMS = 0
MLP_X_MEAN = 1
MLP_X_STD = 2
MLP_Y_MEAN = 3
MLP_Y_STD = 4
MLP_GRADS = 5

_lib = None


def load(path: str | None = None):
    """Load the native library (raises if it is missing: there is no fallback)."""
    global _lib
    if _lib is not None and path is None:
        return _lib
    p = path or LIB_PATH
    if not os.path.exists(p):
        raise RuntimeError(f"libsapx.so not found at {p}; run `python -m sap_warp_amd.build` (or __graft_entry__.build())")
    lib = ctypes.CDLL(p)
    for name, (res, args) in EXPORTS.items():
        fn = getattr(lib, name)
        fn.restype = res
        fn.argtypes = args
    if path is None:
        _lib = lib
    return lib


class NativeError(RuntimeError):
    pass


def check(rc: int, what: str) -> None:
    """Map a status code to the reference's Python exception types."""
    if rc == SAPX_OK:
        return
    msg = load().sapx_last_error().decode(errors="replace")
    text = f"{what} failed: {msg}"
    if rc == SAPX_ERR_INVALID:
        raise ValueError(text)
    if rc == SAPX_ERR_UNSUPPORTED:
        raise NotImplementedError(text)
    if rc == SAPX_ERR_CAPACITY:
        raise ValueError(text)
    raise NativeError(text)


def _ptr(t):
    """Device/host data pointer of a torch tensor or numpy array (None -> NULL)."""
    if t is None:
        return None
    if isinstance(t, np.ndarray):
        return t.ctypes.data
    return t.data_ptr()


class HostArrays:
    """Keeps contiguous host copies alive while a descriptor points at them."""

    def __init__(self):
        self._keep = []

    def arr(self, a, dtype):
        c = np.ascontiguousarray(np.asarray(a), dtype=dtype)
        self._keep.append(c)
        return c.ctypes.data


def model_desc(model, keep: HostArrays) -> SapxModelDesc:
    i32, f32 = np.int32, np.float32
    d = SapxModelDesc()
    d.world_count = model.world_count
    d.joint_count = model.joint_count
    d.dof_count = model.joint_dof_count
    d.coord_count = model.joint_coord_count
    d.body_count = model.body_count
    d.shape_count = model.shape_count
    d.articulation_count = model.articulation_count
    d.pair_count = int(np.asarray(model.shape_pairs).reshape(-1, 2).shape[0])
    d.joint_type = keep.arr(model.joint_type, i32)
    d.joint_parent = keep.arr(model.joint_parent, i32)
    d.joint_child = keep.arr(model.joint_child, i32)
    d.joint_q_start = keep.arr(model.joint_q_start, i32)
    d.joint_qd_start = keep.arr(model.joint_qd_start, i32)
    d.joint_dof_dim = keep.arr(model.joint_dof_dim, i32)
    d.joint_ancestor = keep.arr(model.joint_ancestor, i32)
    d.articulation_start = keep.arr(model.articulation_start, i32)
    d.joint_axis = keep.arr(model.joint_axis, f32)
    d.joint_X_p = keep.arr(model.joint_X_p, f32)
    d.joint_X_c = keep.arr(model.joint_X_c, f32)
    d.joint_armature = keep.arr(model.joint_armature, f32)
    d.body_mass = keep.arr(model.body_mass, f32)
    d.body_inertia = keep.arr(model.body_inertia, f32)
    d.body_com = keep.arr(model.body_com, f32)
    d.gravity = keep.arr(model.gravity, f32)
    d.joint_target_mode = keep.arr(model.joint_target_mode, i32)
    d.joint_target_ke = keep.arr(model.joint_target_ke, f32)
    d.joint_target_kd = keep.arr(model.joint_target_kd, f32)
    d.joint_effort_limit = keep.arr(model.joint_effort_limit, f32)
    d.joint_limit_lower = keep.arr(model.joint_limit_lower, f32)
    d.joint_limit_upper = keep.arr(model.joint_limit_upper, f32)
    d.joint_limit_ke = keep.arr(model.joint_limit_ke, f32)
    d.joint_limit_kd = keep.arr(model.joint_limit_kd, f32)
    d.shape_body = keep.arr(model.shape_body, i32)
    d.shape_type = keep.arr(model.shape_type, i32)
    d.shape_world = keep.arr(model.shape_world, i32)
    d.shape_transform = keep.arr(model.shape_transform, f32)
    d.shape_scale = keep.arr(model.shape_scale, f32)
    d.shape_margin = keep.arr(model.shape_margin, f32)
    d.shape_gap = keep.arr(model.shape_gap, f32)
    d.shape_collision_radius = keep.arr(model.shape_collision_radius, f32)
    d.shape_material_mu = keep.arr(model.shape_material_mu, f32)
    d.shape_material_ke = keep.arr(model.shape_material_ke, f32)
    d.shape_material_tau = keep.arr(model.shape_material_tau, f32)
    d.shape_pairs = keep.arr(model.shape_pairs, i32)
    return d


def require_gpu_device(device, who: str):
    """The torch device a native object runs on; raises unless ``device`` names a visible GPU."""
    import torch

    d = torch.device(device)
    if d.type != "cuda":
        raise ValueError(f"{who} runs on the GPU, but the model lives on {str(device)!r}: build it with a GPU "
                         "device (e.g. device='cuda:0')")
    if not torch.cuda.is_available():
        raise RuntimeError(f"{who}: no GPU is visible to this process")
    if d.index is None:
        d = torch.device("cuda", torch.cuda.current_device())
    if d.index >= torch.cuda.device_count():
        raise ValueError(f"{who}: device {d} does not exist")
    return d


def require_on(t, device, name: str) -> None:
    """Raise before any launch unless tensor ``t`` lives on ``device`` (a GPU)."""
    import torch

    if not isinstance(t, torch.Tensor):
        raise TypeError(f"{name} must be a torch tensor on {device}")
    if t.device != device:
        raise ValueError(f"{name} is on {t.device}, but the native object runs on {device}")


def on_device(device):
    """Context that makes ``device`` current for native calls (allocations, launches)."""
    import torch

    return torch.cuda.device(device)


_CONTACT_FIELDS = (("rigid_contact_shape0", 1, "int32"), ("rigid_contact_shape1", 1, "int32"),
                   ("rigid_contact_point0", 3, "float32"), ("rigid_contact_point1", 3, "float32"),
                   ("rigid_contact_normal", 3, "float32"), ("rigid_contact_margin0", 1, "float32"),
                   ("rigid_contact_margin1", 1, "float32"))


def contact_capacity(contacts) -> int:
    if contacts.rigid_contact_max:
        return int(contacts.rigid_contact_max)
    return int(contacts.rigid_contact_shape0.shape[0]) if contacts.rigid_contact_shape0 is not None else 0


def check_contacts(contacts, device, world_count: int, need_world_start: bool = False) -> None:
    """Device, dtype and size checks of a ``SapContacts`` bundle before its pointers reach a kernel."""
    import torch

    if contacts.rigid_contact_count is None:
        return
    require_on(contacts.rigid_contact_count, device, "contacts.rigid_contact_count")
    if contacts.rigid_contact_count.dtype != torch.int32 or contacts.rigid_contact_count.numel() < 1:
        raise TypeError("contacts.rigid_contact_count must be an int32 tensor with one entry")
    cap = contact_capacity(contacts)
    for f, width, dt in _CONTACT_FIELDS:
        t = getattr(contacts, f)
        if t is None:
            raise ValueError(f"contacts.{f} is missing")
        require_on(t, device, f"contacts.{f}")
        if str(t.dtype) != f"torch.{dt}":
            raise TypeError(f"contacts.{f} must be {dt}")
        if not t.is_contiguous():
            raise ValueError(f"contacts.{f} must be contiguous")
        if t.numel() < cap * width:
            raise ValueError(f"contacts.{f} holds {t.numel() // width} contacts, fewer than the capacity {cap}")
    ws = contacts.rigid_contact_world_start
    if ws is None:
        if need_world_start:
            raise ValueError("contacts need rigid_contact_world_start; allocate them with pipeline.contacts()")
        return
    require_on(ws, device, "contacts.rigid_contact_world_start")
    if ws.dtype != torch.int32 or ws.numel() < world_count + 1 or not ws.is_contiguous():
        raise ValueError(f"contacts.rigid_contact_world_start must be a contiguous int32 tensor of "
                         f"world_count + 1 = {world_count + 1} entries")


class NativeModel:
    """Owns a ``SapxModel*`` (device-resident copy of a SapModel) on ``device``."""

    def __init__(self, model, device=None):
        lib = load()
        keep = HostArrays()
        desc = model_desc(model, keep)
        h = _P()
        if device is None:
            device = require_gpu_device(model.device, "NativeModel")
        with on_device(device):
            check(lib.sapx_model_create(ctypes.byref(desc), ctypes.byref(h)), "sapx_model_create")
        self.handle = h
        self.device = device
        self._lib = lib

    def __del__(self):
        if getattr(self, "handle", None):
            self._lib.sapx_model_destroy(self.handle)
            self.handle = None


def stream_ptr(stream=None):
    """hipStream_t of a torch stream (torch's current stream when None)."""
    import torch

    if stream is None:
        if not torch.cuda.is_available():
            return None
        stream = torch.cuda.current_stream()
    return ctypes.c_void_p(stream.cuda_stream)


def contacts_view(contacts) -> SapxContactsView:
    v = SapxContactsView()
    v.count = _ptr(contacts.rigid_contact_count)
    v.world_start = _ptr(contacts.rigid_contact_world_start)
    v.shape0 = _ptr(contacts.rigid_contact_shape0)
    v.shape1 = _ptr(contacts.rigid_contact_shape1)
    v.point0 = _ptr(contacts.rigid_contact_point0)
    v.point1 = _ptr(contacts.rigid_contact_point1)
    v.normal = _ptr(contacts.rigid_contact_normal)
    v.margin0 = _ptr(contacts.rigid_contact_margin0)
    v.margin1 = _ptr(contacts.rigid_contact_margin1)
    v.capacity = contact_capacity(contacts)
    return v

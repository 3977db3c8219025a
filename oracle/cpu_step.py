"""Python side of the C++ CPU restatement of the SAP step (TEST INFRASTRUCTURE AND CPU BASELINE ONLY).

Loads ``oracle/_build/libsapx_cpu_step.so`` (built from ``oracle/cpu_step.cpp`` by
``oracle/Makefile``): the float64 step of ``oracle/sap_oracle.py`` in C++17 with
OpenMP over worlds.  It validates against the numpy oracle on the CPU and is the
CPU baseline that ``bench.py`` times.  Contacts come in world-grouped (as
``oracle/collide.py`` writes them).
"""
from __future__ import annotations

import ctypes
import os
import subprocess

import numpy as np

from . import sap_oracle as so

_HERE = os.path.dirname(os.path.abspath(__file__))
_LIB_PATH = os.path.join(_HERE, "_build", "libsapx_cpu_step.so")
_lib = None

_I = ctypes.POINTER(ctypes.c_int32)
_F = ctypes.POINTER(ctypes.c_float)
_D = ctypes.POINTER(ctypes.c_double)


class _Model(ctypes.Structure):
    _fields_ = [(k, ctypes.c_int32) for k in ("world_count", "joint_count", "dof_count", "coord_count", "body_count",
                                                "articulation_count")] + \
               [(k, _I) for k in ("joint_type", "joint_parent", "joint_child", "joint_q_start", "joint_qd_start",
                                  "joint_dof_dim", "joint_ancestor", "articulation_start")] + \
               [(k, _F) for k in ("joint_axis", "joint_X_p", "joint_X_c", "joint_armature", "body_mass",
                                  "body_inertia", "body_com", "gravity")] + \
               [("joint_target_mode", _I)] + \
               [(k, _F) for k in ("joint_target_ke", "joint_target_kd", "joint_effort_limit", "joint_limit_lower",
                                  "joint_limit_upper", "joint_limit_ke", "joint_limit_kd")] + \
               [("shape_body", _I)] + [(k, _F) for k in ("shape_mu", "shape_ke", "shape_tau")]


class _State(ctypes.Structure):
    _fields_ = [(k, _F) for k in ("joint_q", "joint_qd", "body_q", "body_f")]


class _Control(ctypes.Structure):
    _fields_ = [(k, _F) for k in ("joint_f", "joint_target_pos", "joint_target_vel", "joint_act")]


class _Contacts(ctypes.Structure):
    _fields_ = [(k, _I) for k in ("world_start", "shape0", "shape1")] + \
               [(k, _F) for k in ("point0", "point1", "normal", "margin0", "margin1")]


class _Opts(ctypes.Structure):
    _fields_ = [("max_rigid_contact", ctypes.c_int32)] + \
               [(k, ctypes.c_double) for k in ("fallback_mu", "fallback_k", "fallback_tau", "beta", "sigma",
                                               "diag_shift")] + \
               [("max_iterations", ctypes.c_int32)] + \
               [(k, ctypes.c_double) for k in ("opt_abs", "opt_rel", "cost_abs", "cost_rel")] + \
               [("line_search", ctypes.c_int32), ("ls_max_iterations", ctypes.c_int32)] + \
               [(k, ctypes.c_double) for k in ("armijo_c", "rho", "relative_slop")] + \
               [(k, ctypes.c_int32) for k in ("weight_mode", "point_mode", "integration")]


class _Out(ctypes.Structure):
    _fields_ = [(k, _F) for k in ("joint_q", "joint_qd", "body_q", "body_qd")] + [("v_sap", _D)] + \
               [(k, _I) for k in ("iterations", "ls_iterations", "contact_count", "converged")]


def build() -> str:
    subprocess.run(["make", "-s", "-C", _HERE], check=True)
    return _LIB_PATH


def _load():
    global _lib
    if _lib is None:
        if not os.path.exists(_LIB_PATH):
            build()
        _lib = ctypes.CDLL(_LIB_PATH)
        _lib.sapx_cpu_step.restype = ctypes.c_int
        _lib.sapx_cpu_step.argtypes = [ctypes.POINTER(_Model), ctypes.POINTER(_State), ctypes.POINTER(_Control),
                                       ctypes.POINTER(_Contacts), ctypes.POINTER(_Opts), ctypes.c_double, _D,
                                       ctypes.POINTER(_Out), ctypes.c_int]
    return _lib


class CpuStepper:
    """Holds the model's contiguous host arrays and steps every world on the CPU."""

    def __init__(self, model, opts: "so.StepOptions"):
        keep = {}

        def arr(name, dtype, value=None):
            a = np.ascontiguousarray(np.asarray(getattr(model, name) if value is None else value, dtype=dtype))
            keep[name] = a
            return a.ctypes.data_as(_I if dtype == np.int32 else _F)

        m = _Model()
        m.world_count, m.joint_count = int(model.world_count), int(model.joint_count)
        m.dof_count, m.coord_count = int(model.joint_dof_count), int(model.joint_coord_count)
        m.body_count, m.articulation_count = int(model.body_count), int(model.articulation_count)
        for k in ("joint_type", "joint_parent", "joint_child", "joint_q_start", "joint_qd_start", "joint_dof_dim",
                  "joint_ancestor", "articulation_start", "joint_target_mode", "shape_body"):
            setattr(m, k, arr(k, np.int32))
        for k in ("joint_axis", "joint_X_p", "joint_X_c", "joint_armature", "body_mass", "body_inertia", "body_com",
                  "gravity", "joint_target_ke", "joint_target_kd", "joint_effort_limit", "joint_limit_lower",
                  "joint_limit_upper", "joint_limit_ke", "joint_limit_kd"):
            setattr(m, k, arr(k, np.float32))
        m.shape_mu = arr("shape_material_mu", np.float32)
        m.shape_ke = arr("shape_material_ke", np.float32)
        m.shape_tau = arr("shape_material_tau", np.float32)
        self._model, self._keep = m, keep
        weight, point, integration = opts.resolved()
        sol = opts.solve or so.SolveOptions.for_variant(opts.line_search)
        o = _Opts()
        o.max_rigid_contact = int(opts.max_rigid_contact)
        o.fallback_mu = opts.fallback_mu if opts.fallback_mu is not None else so.infer_mu_fallback(model)
        o.fallback_k = float(opts.fallback_stiffness)
        o.fallback_tau = opts.contact_tau_d if opts.contact_tau_d is not None else so.infer_tau_fallback(model)
        o.beta, o.sigma, o.diag_shift = sol.beta, sol.sigma, sol.diag_shift
        o.max_iterations = sol.max_iterations
        o.opt_abs, o.opt_rel, o.cost_abs, o.cost_rel = (sol.optimality_abs_tol, sol.optimality_rel_tol,
                                                        sol.cost_abs_tol, sol.cost_rel_tol)
        o.line_search = {"armijo_decay": 0, "exact_root": 1, "monotone_decay": 2}[sol.line_search]
        o.ls_max_iterations = sol.ls_max_iterations
        o.armijo_c, o.rho, o.relative_slop = sol.armijo_c, sol.rho, sol.relative_slop
        o.weight_mode = 0 if weight == "body_inertia" else 1
        o.point_mode = 0 if point == "witness_point" else 1
        o.integration = 0 if integration == "midpoint" else 1
        self._opts = o
        self.model = model

    def step(self, state: dict, control: dict, contacts: dict, dt: float, v_guess=None, threads: int = 0) -> dict:
        model = self.model
        keep = []

        def f32(a):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.float32))
            keep.append(a)
            return a.ctypes.data_as(_F)

        def i32(a):
            a = np.ascontiguousarray(np.asarray(a, dtype=np.int32))
            keep.append(a)
            return a.ctypes.data_as(_I)

        st = _State(f32(state["joint_q"]), f32(state["joint_qd"]), f32(state["body_q"]),
                    f32(state["body_f"]) if state.get("body_f") is not None else None)
        ct = _Control(*[f32(control[k]) for k in ("joint_f", "joint_target_pos", "joint_target_vel", "joint_act")])
        ws = contacts.get("world_start")
        if ws is None:
            raise ValueError("the CPU step takes world-grouped contacts (world_start)")
        cc = _Contacts(i32(ws), i32(contacts["shape0"]), i32(contacts["shape1"]), f32(contacts["point0"]),
                       f32(contacts["point1"]), f32(contacts["normal"]), f32(contacts["margin0"]),
                       f32(contacts["margin1"]))
        W = int(model.world_count)
        out = dict(joint_q=np.zeros(model.joint_coord_count, np.float32),
                   joint_qd=np.zeros(model.joint_dof_count, np.float32),
                   body_q=np.zeros((model.body_count, 7), np.float32),
                   body_qd=np.zeros((model.body_count, 6), np.float32),
                   v_sap=np.zeros(model.joint_dof_count, np.float64),
                   iterations=np.zeros(W, np.int32), ls_iterations=np.zeros(W, np.int32),
                   contact_count=np.zeros(W, np.int32), converged=np.zeros(W, np.int32))
        o = _Out(*[out[k].ctypes.data_as(_F) for k in ("joint_q", "joint_qd", "body_q", "body_qd")],
                 out["v_sap"].ctypes.data_as(_D),
                 *[out[k].ctypes.data_as(_I) for k in ("iterations", "ls_iterations", "contact_count", "converged")])
        guess = None
        if v_guess is not None:
            g = np.ascontiguousarray(np.asarray(v_guess, dtype=np.float64))
            keep.append(g)
            guess = g.ctypes.data_as(_D)
        rc = _load().sapx_cpu_step(ctypes.byref(self._model), ctypes.byref(st), ctypes.byref(ct), ctypes.byref(cc),
                                   ctypes.byref(self._opts), float(dt), guess, ctypes.byref(o), int(threads))
        if rc != 0:
            raise RuntimeError(f"sapx_cpu_step failed ({rc})")
        return out

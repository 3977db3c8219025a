"""C ABI checks that need no GPU: exports, struct layouts, validation paths.

* every function declared in ``include/sapx.h`` is exported by ``libsapx.so``;
* the ctypes mirrors in ``sap_warp_amd/_native.py`` have the C layout (sizes
  and field offsets printed by a tiny C program compiled against the header);
* ``sapx_model_create`` rejects bad models before touching the device, with
  the error classes the Python layer maps to ValueError / NotImplementedError.
"""
from __future__ import annotations

import ctypes
import os
import re
import subprocess

import numpy as np
import pytest

from sap_warp_amd import _native as nat

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
HEADER = os.path.join(ROOT, "include", "sapx.h")


def _declared_functions():
    text = open(HEADER).read()
    text = re.sub(r"/\*.*?\*/", "", text, flags=re.S)
    return sorted(set(re.findall(r"\b(sapx_[a-z_]+)\s*\(", text)))


def test_library_exports_every_declared_function():
    lib = nat.load()
    names = _declared_functions()
    assert len(names) >= 14
    missing = [n for n in names if not hasattr(lib, n)]
    assert not missing, missing
    assert set(names) <= set(nat.EXPORTS), set(names) - set(nat.EXPORTS)


def test_version_and_last_error():
    lib = nat.load()
    assert lib.sapx_version() >= 100
    assert isinstance(lib.sapx_last_error(), bytes)


STRUCTS = {
    "SapxModelDesc": nat.SapxModelDesc,
    "SapxSolverOpts": nat.SapxSolverOpts,
    "SapxStateView": nat.SapxStateView,
    "SapxControlView": nat.SapxControlView,
    "SapxContactsView": nat.SapxContactsView,
    "SapxSchedule": nat.SapxSchedule,
    "SapxLargeProblem": nat.SapxLargeProblem,
    "SapxLargeOpts": nat.SapxLargeOpts,
    "SapxLargeStats": nat.SapxLargeStats,
}


def test_ctypes_layout_matches_header(tmp_path):
    lines = ['#include <stdio.h>', '#include <stddef.h>', '#include "sapx.h"', "int main(void) {"]
    for cname, py in STRUCTS.items():
        lines.append(f'printf("{cname} size %zu\\n", sizeof({cname}));')
        for fname, _ in py._fields_:
            lines.append(f'printf("{cname}.{fname} %zu\\n", offsetof({cname}, {fname}));')
    lines.append("return 0; }")
    src = tmp_path / "layout.c"
    src.write_text("\n".join(lines))
    exe = tmp_path / "layout"
    subprocess.run(["gcc", "-std=c99", "-I", os.path.join(ROOT, "include"), str(src), "-o", str(exe)], check=True)
    out = subprocess.run([str(exe)], check=True, capture_output=True, text=True).stdout.split("\n")
    got = {}
    for line in out:
        if line:
            k, v = line.rsplit(" ", 1)
            got[k] = int(v)
    for cname, py in STRUCTS.items():
        assert got[f"{cname} size"] == ctypes.sizeof(py), cname
        for fname, _ in py._fields_:
            assert got[f"{cname}.{fname}"] == getattr(py, fname).offset, (cname, fname)


def _desc_from(model):
    keep = nat.HostArrays()
    return nat.model_desc(model, keep), keep


def _d6_model(angular, linear=()):
    from sap_warp_amd.runtime import GEO_SPHERE, SAP_JOINT_D6
    from sap_warp_amd.scenes import DofCfg, SceneBuilder, ShapeCfg, tf

    b = SceneBuilder()
    b.current_world = 0
    root = b.add_body(tf((0, 0, 1)))
    b.add_shape(root, GEO_SPHERE, tf(), (0.1, 0, 0), ShapeCfg())
    j0 = b.add_free_joint(root)
    child = b.add_body(tf((0, 0, 0.5)))
    b.add_shape(child, GEO_SPHERE, tf(), (0.1, 0, 0), ShapeCfg())
    j1 = b.add_joint(root, child, SAP_JOINT_D6, angular_axes=[DofCfg(axis=a) for a in angular],
                     linear_axes=[DofCfg(axis=a) for a in linear], parent_xform=tf((0, 0, -0.5)))
    b.add_articulation([j0, j1])
    return b.finalize()


@pytest.mark.parametrize("case", ["four_angles", "dof_dim_mismatch", "no_axes"])
def test_model_create_rejects_malformed_d6_joints(case):
    """D6 joints take 0-3 linear and 0-3 angular axes with one coordinate per dof and a
    joint_dof_dim that matches the dof count; anything else is rejected before any GPU work."""
    axes = ((1, 0, 0), (0, 1, 0), (0, 0, 1))
    if case == "four_angles":
        model = _d6_model(axes + ((1, 1, 0),))
    elif case == "no_axes":
        model = _d6_model(())
    else:
        model = _d6_model(axes[:2])
        model.joint_dof_dim[1] = (1, 2)
    lib = nat.load()
    desc, _keep = _desc_from(model)
    h = ctypes.c_void_p()
    err = NotImplementedError if case == "no_axes" else ValueError
    with pytest.raises(err, match="D6"):
        nat.check(lib.sapx_model_create(ctypes.byref(desc), ctypes.byref(h)), "sapx_model_create")


def test_model_create_rejects_unequal_worlds():
    from sap_warp_amd.scenes import build_ant_scene

    model = build_ant_scene(num_worlds=2, seed=1)
    jt = model.joint_type.copy()
    jt[model.joint_count // 2 + 1] = 0  # world 1's first hinge becomes prismatic
    from dataclasses import replace

    bad = replace(model, joint_type=jt)
    lib = nat.load()
    desc, _keep = _desc_from(bad)
    h = ctypes.c_void_p()
    with pytest.raises(NotImplementedError, match="topology"):
        nat.check(lib.sapx_model_create(ctypes.byref(desc), ctypes.byref(h)), "sapx_model_create")


def test_null_arguments_are_invalid():
    lib = nat.load()
    with pytest.raises(ValueError):
        nat.check(lib.sapx_model_create(None, None), "sapx_model_create")
    with pytest.raises(ValueError):
        nat.check(lib.sapx_solver_read(None, 0, None, 0, None), "sapx_solver_read")
    with pytest.raises(ValueError):
        nat.check(lib.sapx_collide(None, None, None, None), "sapx_collide")


def test_library_has_no_cpu_fallback_symbols():
    """The shipped library is the HIP path only: no oracle symbols linked in."""
    lib = nat.load()
    assert not hasattr(lib, "sapx_oracle_collide")
    assert np.dtype(np.int32).itemsize == 4


def test_autotune_candidates_cover_the_schedule_space():
    from sap_warp_amd.autotune import candidate_schedules

    small = candidate_schedules(16)
    assert {(c["kernel"], c["small_capacity"]) for c in small} == {("fused", 64), ("wave", 16)}
    full = candidate_schedules(64)
    keys = {(c["kernel"], c["small_capacity"], c["fork"], c["world_order"]) for c in full}
    assert len(keys) == len(full) == 11
    assert ("wave", 16, True, True) in keys and ("wave", 32, False, False) in keys and ("wave", 64, False, True) in keys
    assert not any(c["small_capacity"] == 64 and c["fork"] for c in full)

"""The C ABI: libfsgh.so loads, exports every entry point include/fsgh.h
declares, and rejects bad input with FSGH_E_* codes (host-side checks; no
GPU compute is issued here)."""
import ctypes as C
import os
import re

import numpy as np
import pytest

from fierro_amd import _abi, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def header_functions():
    txt = open(os.path.join(ROOT, "include", "fsgh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsgh_\w+)\s*\(", txt)))


def test_header_and_exports_agree():
    names = header_functions()
    assert len(names) >= 30
    assert sorted(_abi.EXPORTS) == names
    lib = _abi.load()
    for n in names:
        assert hasattr(lib, n), n


def test_exports_via_nm():
    import subprocess

    out = subprocess.run(["nm", "-D", "--defined-only", _abi.LIB_PATH], capture_output=True,
                         text=True, check=True).stdout
    syms = set(re.findall(r" T (fsgh_\w+)", out))
    assert set(header_functions()) <= syms


def test_library_contains_gfx950_code():
    """The offload bundle built here carries a gfx950 code object."""
    data = open(_abi.LIB_PATH, "rb").read()
    assert b"amdgcn-amd-amdhsa--gfx950" in data


def _p(a):
    return a.ctypes.data_as(C.c_void_p)


def test_create_rejects_bad_input():
    L = _abi.load()
    coords, conn = P.box_mesh(2, 2, 2)
    ctx = C.c_void_p()
    desc = _abi.MeshDesc(coords.shape[0], conn.shape[0], 1, 1, None, None)
    assert L.fsgh_create(None, _p(conn), _p(coords), -1, C.byref(ctx)) == -1
    bad = conn.copy()
    bad[3, 2] = coords.shape[0]                     # node id out of range
    assert L.fsgh_create(C.byref(desc), _p(bad), _p(coords), -1, C.byref(ctx)) == -1
    desc0 = _abi.MeshDesc(coords.shape[0], 0, 1, 1, None, None)
    assert L.fsgh_create(C.byref(desc0), _p(conn), _p(coords), -1, C.byref(ctx)) == -1
    # batched offsets that do not cover the mesh
    noff = np.array([0, 10, coords.shape[0] - 1], np.int64)
    eoff = np.array([0, 4, 8], np.int64)
    desc2 = _abi.MeshDesc(coords.shape[0], conn.shape[0], 2, 0, _p(noff).value, _p(eoff).value)
    assert L.fsgh_create(C.byref(desc2), _p(conn), _p(coords), -1, C.byref(ctx)) == -1
    # elements referencing nodes of another instance
    noff = np.array([0, 9, coords.shape[0]], np.int64)
    desc3 = _abi.MeshDesc(coords.shape[0], conn.shape[0], 2, 0, _p(noff).value, _p(eoff).value)
    assert L.fsgh_create(C.byref(desc3), _p(conn), _p(coords), -1, C.byref(ctx)) == -1


def test_host_only_context_error_paths():
    L = _abi.load()
    coords, conn = P.box_mesh(3, 3, 3)
    ctx = C.c_void_p()
    desc = _abi.MeshDesc(coords.shape[0], conn.shape[0], 1, 1, None, None)
    assert L.fsgh_create(C.byref(desc), _p(conn), _p(coords), -1, C.byref(ctx)) == 0
    try:
        # operators need a device context
        assert L.fsgh_get_force_sgh(ctx, 1e-16, 1e-8, 1e-3, 1.0) == -5
        assert L.fsgh_run(ctx, 1, 0) == -5
        buf = np.zeros(27)
        assert L.fsgh_download(ctx, _abi.ELEM_DEN, _p(buf), buf.nbytes) == -5
        # reflected condition on a sphere is rejected (reference indexes out of range)
        b = (_abi.Boundary * 1)(_abi.Boundary(4, 1, 1.0))
        assert L.fsgh_set_bcs(ctx, 0, 1, b) == -6
        b = (_abi.Boundary * 1)(_abi.Boundary(7, 1, 1.0))
        assert L.fsgh_set_bcs(ctx, 0, 1, b) == -1
        assert L.fsgh_set_bcs(ctx, 3, 0, None) == -1
        msg = L.fsgh_last_error(ctx)
        assert msg is not None
    finally:
        assert L.fsgh_destroy(ctx) == 0
    assert L.fsgh_destroy(None) == -1


def test_device_count_without_gpu_is_safe():
    n = _abi.load().fsgh_device_count()
    assert n >= 0


def test_tuning_entry_points_on_host_context():
    """Launch tuning needs a device context; the defaults read back on any."""
    L = _abi.load()
    coords, conn = P.box_mesh(2, 2, 2)
    ctx = C.c_void_p()
    desc = _abi.MeshDesc(coords.shape[0], conn.shape[0], 1, 1, None, None)
    assert L.fsgh_create(C.byref(desc), _p(conn), _p(coords), -1, C.byref(ctx)) == 0
    try:
        out = (_abi.Tuning * _abi.NUM_TUNED_KERNELS)()
        assert L.fsgh_get_tuning(ctx, out, _abi.NUM_TUNED_KERNELS) == 0
        assert [(t.kernel, t.block, t.min_waves, t.xcd_order) for t in out] == \
            [(k, 256, 1, 16) for k in range(_abi.NUM_TUNED_KERNELS)]   # chunks of 16 per XCD
        assert L.fsgh_autotune(ctx, 1, out) == -5
        cfg = (_abi.Tuning * 1)(_abi.Tuning(_abi.KERNEL_FORCE, 128, 1, 1, 0.0, 0.0))
        assert L.fsgh_set_tuning(ctx, cfg, 1) == -5
        assert L.fsgh_get_tuning(ctx, None, 1) == -1
    finally:
        assert L.fsgh_destroy(ctx) == 0


def test_persistent_flag_is_switched_off():
    """FSGH_RUN_PERSISTENT stays off until the abort of its first cooperative
    launch is explained: fsgh_run returns FSGH_E_UNSUPPORTED with the reason
    before anything runs (here on a host-only context, which no other run
    flag gets past)."""
    from fierro_amd.sgh import MeshMaps

    mm = MeshMaps(P.sedov(2, cycles=1))
    L = _abi.load()
    assert L.fsgh_run(mm.ctx, 1, _abi.RUN_PERSISTENT) == -6
    assert b"switched off" in L.fsgh_last_error(mm.ctx)
    assert L.fsgh_run(mm.ctx, 1, _abi.RUN_GRAPH) == -5

"""fsgh_energy_noalloc: the energy tally of caller-held device arrays with no
allocation, against fsgh_energy on the same state (bitwise: same kernels,
same partial-sum count, same element order), on two calls in a row with the
second input smaller, in 3D and in 2D RZ; plus the host-side rejections.

Runs last (and alone in a fresh process): the tally's partial sums live in a
fixed per-GPU array that every call rewrites.
"""
import ctypes as C

import numpy as np
import pytest

from fierro_amd import _abi, problems as P

pytestmark = pytest.mark.gpu


class Hip:
    """The HIP runtime libfsgh.so itself links (the test allocates the
    caller's arrays with it; the entry point under test allocates nothing)."""

    def __init__(self):
        self.L = C.CDLL("libamdhip64.so.7")
        for f in ("hipMalloc", "hipFree", "hipMemcpy", "hipDeviceSynchronize", "hipSetDevice",
                  "hipStreamCreate", "hipStreamDestroy", "hipStreamSynchronize"):
            getattr(self.L, f).restype = C.c_int
        self.L.hipMemcpy.argtypes = [C.c_void_p, C.c_void_p, C.c_size_t, C.c_int]
        self.L.hipFree.argtypes = [C.c_void_p]
        self.L.hipStreamDestroy.argtypes = [C.c_void_p]
        self.L.hipStreamSynchronize.argtypes = [C.c_void_p]
        assert self.L.hipSetDevice(0) == 0
        self.live = []

    def put(self, a):
        a = np.ascontiguousarray(a, np.float64)
        p = C.c_void_p()
        assert self.L.hipMalloc(C.byref(p), max(a.nbytes, 8)) == 0
        self.live.append(p)
        assert self.L.hipMemcpy(p, a.ctypes.data_as(C.c_void_p), a.nbytes, 1) == 0  # H2D
        return p

    def get(self, p, n):
        out = np.empty(n, np.float64)
        assert self.L.hipMemcpy(out.ctypes.data_as(C.c_void_p), p, out.nbytes, 2) == 0  # D2H
        return out

    def free(self):
        for p in self.live:
            self.L.hipFree(p)
        self.live = []


def tally_noalloc(L, dims, fields, stream=None):
    """Copy the fields to device arrays (the caller's arrays) and tally them."""
    hip = Hip()
    try:
        vel = fields["node_vel"][1]
        emass = hip.put(fields["elem_mass"])
        sie = hip.put(fields["elem_sie"][1])
        nmass = hip.put(fields["node_mass"])
        u, v = hip.put(vel[:, 0]), hip.put(vel[:, 1])
        w = hip.put(vel[:, 2]) if dims == 3 else None
        radius = hip.put(fields["node_coords"][1][:, 1]) if dims == 2 else None
        out = hip.put(np.full(2, np.nan))
        E, N = len(fields["elem_mass"]), len(fields["node_mass"])
        rc = L.fsgh_energy_noalloc(0, dims, E, emass, sie, N, nmass, u, v, w, radius, out, stream)
        assert rc == _abi.FSGH_OK, rc
        assert hip.L.hipDeviceSynchronize() == 0
        return hip.get(out, 2)
    finally:
        hip.free()


def run_case(spec, cycles):
    from fierro_amd.sgh import SGHSolver

    g = SGHSolver(spec, renumber=False)
    g.sgh_solve(cycles=cycles)
    ie, ke = g.energy(0)
    fields = {k: g.download(k) for k in ("elem_mass", "elem_sie", "node_mass", "node_vel",
                                         "node_coords")}
    dims = g.dims
    g.close()
    return dims, fields, np.array([ie, ke])


def test_noalloc_tally_equals_fsgh_energy_twice():
    L = _abi.load()
    # the second call has the smaller input: the partial slots past its
    # blocks must not keep the first call's sums
    for tag, spec, cycles in (("sedov 14^3", P.sedov(14, cycles=30), 30),
                              ("sedov 5^3", P.sedov(5, cycles=12), 12),
                              ("sod_rz 60x6", P.sod_rz(60, 6, cycles=25), 25),
                              ("noh_rz 8x8", P.noh_rz(8, cycles=10), 10)):
        print(f"[noalloc] {tag}", flush=True)
        dims, fields, ref = run_case(spec, cycles)
        got = tally_noalloc(L, dims, fields)
        assert ref[1] > 0
        assert np.array_equal(got, ref), (tag, got, ref)


def test_noalloc_on_a_stream():
    """Asynchronous on the caller's stream (a default-priority stream)."""
    L = _abi.load()
    dims, fields, ref = run_case(P.sedov(6, cycles=8), 8)
    hip = Hip()
    st = C.c_void_p()
    assert hip.L.hipStreamCreate(C.byref(st)) == 0
    try:
        got = tally_noalloc(L, dims, fields, stream=st)
    finally:
        hip.L.hipStreamSynchronize(st)
        hip.L.hipStreamDestroy(st)
    assert np.array_equal(got, ref)


def test_noalloc_rejects_bad_arguments():
    """All of these are refused on the host before any GPU work."""
    L = _abi.load()
    dummy = C.c_void_p(16)   # never dereferenced: every call below is rejected
    big = _abi.NOALLOC_MAX_ENTRIES + 1
    args = lambda dims, E, N, out=dummy: (0, dims, E, dummy, dummy, N, dummy, dummy, dummy,  # noqa: E731
                                          dummy, dummy, out, None)
    assert L.fsgh_energy_noalloc(*args(3, big, 10)) == -1
    assert L.fsgh_energy_noalloc(*args(3, 10, big)) == -1
    assert L.fsgh_energy_noalloc(*args(4, 10, 10)) == -1
    assert L.fsgh_energy_noalloc(*args(3, -1, 10)) == -1
    assert L.fsgh_energy_noalloc(*args(3, 10, 10, out=None)) == -1
    assert L.fsgh_energy_noalloc(0, 3, 10, None, dummy, 10, dummy, dummy, dummy, dummy, None,
                                 dummy, None) == -1
    assert L.fsgh_energy_noalloc(0, 2, 10, dummy, dummy, 10, dummy, dummy, dummy, None, None,
                                 dummy, None) == -1
    assert L.fsgh_energy_noalloc(99, 3, 10, dummy, dummy, 10, dummy, dummy, dummy, dummy, None,
                                 dummy, None) == -1

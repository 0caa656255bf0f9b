"""ctypes front end of the CPU oracle (test infrastructure only).

Only tests/, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of
``bench.py`` import this module; the product path (``fierro_amd``) never does.

``OracleProblem`` holds one problem in the reference's data layout and runs
the reference's setup sequence (K/driver.cpp:138-267 -> K/setup.cpp:10-424)
and operators through ``liboracle_sgh.so``.  Problems are described by the
plain ``ProblemSpec`` objects of ``fierro_amd.problems`` (duck-typed here).
"""
from __future__ import annotations

import ctypes as C
import os
import subprocess

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))

i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
f64p = np.ctypeslib.ndpointer(np.float64, flags="C_CONTIGUOUS")
P = C.c_void_p


class OrcMesh(C.Structure):
    _fields_ = [("num_nodes", C.c_int64), ("num_elems", C.c_int64),
                ("nodes_in_elem", P), ("corners_in_node_start", P),
                ("corners_in_node", P), ("elems_in_node", P),
                ("elems_in_elem_start", P), ("elems_in_elem", P),
                ("num_bdy_sets", C.c_int64), ("bdy_nodes_in_set_start", P),
                ("bdy_nodes_in_set", P)]


class OrcState(C.Structure):
    _fields_ = [("node_coords", P), ("node_vel", P), ("node_mass", P),
                ("elem_den", P), ("elem_pres", P), ("elem_sspd", P),
                ("elem_sie", P), ("elem_vol", P), ("elem_div", P),
                ("elem_mass", P), ("elem_mat_id", P), ("elem_statev", P),
                ("num_state_vars", C.c_int64), ("corner_force", P), ("mat_eos", P),
                ("elem_stress", P), ("mat_strength", P), ("mat_strength_params", P)]


class OrcMaterial(C.Structure):
    _fields_ = [("q1", C.c_double), ("q1ex", C.c_double),
                ("q2", C.c_double), ("q2ex", C.c_double)]


class OrcBoundary(C.Structure):
    _fields_ = [("surface", C.c_int64), ("value", C.c_double), ("hydro_bc", C.c_int64)]


class OrcFill(C.Structure):
    _fields_ = [("volume", C.c_int64), ("mat_id", C.c_int64),
                ("x1", C.c_double), ("x2", C.c_double), ("y1", C.c_double),
                ("y2", C.c_double), ("z1", C.c_double), ("z2", C.c_double),
                ("radius1", C.c_double), ("radius2", C.c_double),
                ("velocity", C.c_int64), ("u", C.c_double), ("v", C.c_double),
                ("w", C.c_double), ("speed", C.c_double), ("sie", C.c_double),
                ("den", C.c_double)]


class OrcRunParams(C.Structure):
    _fields_ = [("time_value", C.c_double), ("time_final", C.c_double),
                ("dt_start", C.c_double), ("dt_max", C.c_double),
                ("dt_min", C.c_double), ("dt_cfl", C.c_double),
                ("graphics_time", C.c_double), ("graphics_dt_ival", C.c_double),
                ("graphics_cyc_ival", C.c_int64), ("cycle_stop", C.c_int64),
                ("rk_num_stages", C.c_int64), ("fuzz", C.c_double),
                ("small", C.c_double), ("dt_replay", C.c_void_p), ("dt_record", C.c_void_p),
                ("write_record", C.c_void_p)]


class OrcRunResult(C.Structure):
    _fields_ = [("cycles", C.c_int64), ("time_value", C.c_double), ("dt", C.c_double),
                ("graphics_time", C.c_double), ("ie0", C.c_double), ("ke0", C.c_double),
                ("te0", C.c_double), ("ie1", C.c_double), ("ke1", C.c_double),
                ("te1", C.c_double), ("seconds", C.c_double)]


_LIBS: dict[str, C.CDLL] = {}


def _run_params(run, cycle_stop, dt_replay, record_dt, rk_num_stages, record_writes):
    """orc_run_params for one sgh_solve call (start time 0, first graphics
    time = graphics_dt_ival as K/driver.cpp:39 sets it).  The returned
    arrays must stay alive for the call."""
    r = dict(run)
    if cycle_stop is not None:
        r["cycle_stop"] = cycle_stop
    ncyc = int(r["cycle_stop"])
    rep = None if dt_replay is None else np.ascontiguousarray(dt_replay, np.float64)
    if rep is not None and rep.size < ncyc:
        raise ValueError("dt_replay shorter than cycle_stop")
    rec = np.zeros(ncyc) if record_dt else None
    wrec = np.zeros(ncyc, np.int64) if record_writes else None
    p = OrcRunParams(0.0, r["time_final"], r["dt_start"], r["dt_max"], r["dt_min"],
                     r["dt_cfl"], r["graphics_dt_ival"], r["graphics_dt_ival"],
                     int(r["graphics_cyc_ival"]), ncyc, int(rk_num_stages), r["fuzz"],
                     r["small"], _ptr(rep), _ptr(rec), _ptr(wrec))
    p._keep = (rep, rec, wrec)
    return p, rec, wrec


def _run_result(res, rec, wrec):
    out = {k: getattr(res, k) for k, _ in OrcRunResult._fields_}
    if rec is not None:
        out["dt_history"] = rec[:res.cycles]
    if wrec is not None:
        out["write_cycles"] = [int(c) + 1 for c in np.nonzero(wrec[:res.cycles])[0]]
    return out


def build() -> None:
    """Compile the oracle libraries (gcc, see oracle/Makefile)."""
    subprocess.run(["make", "-s", "-C", _HERE], check=True)


def load_library(path: str) -> C.CDLL:
    """Open an oracle build and declare the non-int signatures."""
    L = C.CDLL(path)
    L.orc_get_vol_hex.restype = C.c_double
    L.orc_get_vol_hex.argtypes = [f64p, f64p, f64p]
    L.orc_get_bmatrix.argtypes = [f64p, f64p, f64p, f64p]
    L.orc_build_elem_elem_connectivity.restype = C.c_int64
    L.orc_build_patch_connectivity.restype = C.c_int64
    L.orc_build_bdy_sets.restype = C.c_int64
    L.orc_read_ensight.restype = C.c_int
    L.orc_rk_alpha.restype = C.c_double
    L.orc_rk_alpha.argtypes = [C.c_int64, C.c_int64]
    return L


def lib(fast: bool = False, strength_model: bool = False) -> C.CDLL:
    """The parity build, the timing build (fast) or the parity build whose
    user strength hook runs the hypoelastic test model (strength_model)."""
    name = ("liboracle_sgh_strength.so" if strength_model else
            "liboracle_sgh_fast.so" if fast else "liboracle_sgh.so")
    if name not in _LIBS:
        path = os.path.join(_HERE, name)
        if not os.path.exists(path):
            build()
        _LIBS[name] = load_library(path)
    return _LIBS[name]


def _ptr(a: np.ndarray | None):
    return None if a is None else a.ctypes.data_as(C.c_void_p)


def read_ensight(path: str):
    L = lib()
    nn, ne = C.c_int64(), C.c_int64()
    rc = L.orc_read_ensight(path.encode(), C.byref(nn), C.byref(ne), None, None)
    if rc != 0:
        raise IOError(f"orc_read_ensight({path}) -> {rc}")
    coords = np.zeros((nn.value, 3))
    conn = np.zeros((ne.value, 8), np.int64)
    rc = L.orc_read_ensight(path.encode(), C.byref(nn), C.byref(ne), _ptr(coords), _ptr(conn))
    if rc != 0:
        raise IOError(f"orc_read_ensight({path}) -> {rc}")
    return coords, conn


def gen_box(nx, ny, nz, lx_num, ly_num, lz_num, l_den):
    N = (nx + 1) * (ny + 1) * (nz + 1)
    E = nx * ny * nz
    coords = np.zeros((N, 3))
    conn = np.zeros((E, 8), np.int64)
    lib().orc_gen_box(C.c_int64(nx), C.c_int64(ny), C.c_int64(nz), C.c_double(lx_num),
                      C.c_double(ly_num), C.c_double(lz_num), C.c_double(l_den),
                      _ptr(coords), _ptr(conn))
    return coords, conn


def connectivity(coords: np.ndarray, conn: np.ndarray):
    """All reference connectivity (K/mesh.h:180-649) as a dict of int64 arrays."""
    L = lib()
    N, E = coords.shape[0], conn.shape[0]
    conn = np.ascontiguousarray(conn, np.int64)
    cin_start = np.zeros(N + 1, np.int64)
    cin = np.zeros(8 * E, np.int64)
    ein = np.zeros(8 * E, np.int64)
    cie = np.zeros((E, 8), np.int64)
    L.orc_build_corner_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(cin_start),
                                    _ptr(cin), _ptr(ein), _ptr(cie))
    ee_start = np.zeros(E + 1, np.int64)
    nb = L.orc_build_elem_elem_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn),
                                            _ptr(cin_start), _ptr(ein), _ptr(ee_start), None)
    ee = np.zeros(max(nb, 1), np.int64)
    L.orc_build_elem_elem_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(cin_start),
                                       _ptr(ein), _ptr(ee_start), _ptr(ee))
    counts = np.zeros(3, np.int64)
    L.orc_build_patch_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(ee_start),
                                   _ptr(ee), None, None, None, None, _ptr(counts))
    P_, BP, BN = (int(c) for c in counts)
    pie = np.zeros((E, 6), np.int64)
    nip = np.zeros((max(P_, 1), 4), np.int64)
    bp = np.zeros(max(BP, 1), np.int64)
    bn = np.zeros(max(BN, 1), np.int64)
    L.orc_build_patch_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(ee_start),
                                   _ptr(ee), _ptr(pie), _ptr(nip), _ptr(bp), _ptr(bn),
                                   _ptr(counts))
    return dict(corners_in_node_start=cin_start, corners_in_node=cin, elems_in_node=ein,
                corners_in_elem=cie, elems_in_elem_start=ee_start, elems_in_elem=ee[:nb],
                patches_in_elem=pie, nodes_in_patch=nip[:P_], bdy_patches=bp[:BP],
                bdy_nodes=bn[:BN], num_patches=P_)


def bdy_sets(coords, cx, bcs):
    L = lib()
    ns = len(bcs)
    arr = (OrcBoundary * max(ns, 1))(*[OrcBoundary(int(b[0]), float(b[1]), int(b[2])) for b in bcs])
    start = np.zeros(ns + 1, np.int64)
    cap = max(ns * max(len(cx["bdy_patches"]), 1) * 4, 1)
    out = np.zeros(cap, np.int64)
    xyz = np.ascontiguousarray(coords, np.float64)
    tot = L.orc_build_bdy_sets(_ptr(xyz), C.c_int64(len(cx["bdy_patches"])),
                               _ptr(cx["bdy_patches"]), _ptr(cx["nodes_in_patch"]),
                               C.c_int64(ns), arr, C.c_int64(xyz.shape[0]), _ptr(start),
                               _ptr(out))
    return start, out[:tot].copy(), arr


class OracleProblem:
    """One SGH problem in the reference layout, set up as K/driver.cpp does."""

    def __init__(self, spec, fast: bool = False, coords=None, conn=None,
                 strength_model: bool = False):
        self.L = lib(fast, strength_model)
        self.spec = spec
        if coords is None:
            coords, conn = spec.mesh_arrays(oracle_gen=gen_box, oracle_read=read_ensight)
        self.coords0 = np.ascontiguousarray(coords, np.float64)
        self.conn = np.ascontiguousarray(conn, np.int64)
        N, E = self.coords0.shape[0], self.conn.shape[0]
        self.N, self.E = N, E
        self.cx = connectivity(self.coords0, self.conn)
        self.bset_start, self.bset_nodes, self._bc = bdy_sets(self.coords0, self.cx, spec.bcs)
        nsv = spec.num_state_vars
        self.node_coords = np.zeros((2, N, 3))
        self.node_coords[0] = self.coords0
        self.node_coords[1] = self.coords0
        self.node_vel = np.zeros((2, N, 3))
        self.node_mass = np.zeros(N)
        z = lambda: np.zeros(E)  # noqa: E731
        self.elem_den, self.elem_pres, self.elem_sspd = z(), z(), z()
        self.elem_vol, self.elem_div, self.elem_mass = z(), z(), z()
        self.elem_sie = np.zeros((2, E))
        self.elem_mat_id = np.zeros(E, np.int64)
        self.elem_statev = np.zeros((E, nsv))
        self.corner_force = np.zeros((8 * E, 3))
        self._mesh = OrcMesh(N, E, _ptr(self.conn), _ptr(self.cx["corners_in_node_start"]),
                             _ptr(self.cx["corners_in_node"]), _ptr(self.cx["elems_in_node"]),
                             _ptr(self.cx["elems_in_elem_start"]), _ptr(self.cx["elems_in_elem"]),
                             len(spec.bcs), _ptr(self.bset_start), _ptr(self.bset_nodes))
        self._state = OrcState(_ptr(self.node_coords), _ptr(self.node_vel), _ptr(self.node_mass),
                               _ptr(self.elem_den), _ptr(self.elem_pres), _ptr(self.elem_sspd),
                               _ptr(self.elem_sie), _ptr(self.elem_vol), _ptr(self.elem_div),
                               _ptr(self.elem_mass), _ptr(self.elem_mat_id),
                               _ptr(self.elem_statev), nsv, _ptr(self.corner_force), None, None,
                               None, None)
        self._set_strength_params(spec)
        mats = spec.materials
        if spec.eos:
            self.mat_eos = np.ascontiguousarray(spec.eos, np.int64)
            self._state.mat_eos = _ptr(self.mat_eos)
        self.elem_stress = None
        if getattr(spec, "strength", None):
            self.enable_stress()
        self._mat = (OrcMaterial * len(mats))(*[OrcMaterial(*m) for m in mats])
        self.setup()

    def _set_strength_params(self, spec):
        """Per-material strength parameters (ProblemSpec.strength_params)."""
        sp = getattr(spec, "strength_params", None)
        if sp:
            a = np.zeros((len(spec.materials), 8))
            for k, row in enumerate(sp):
                a[k, :len(row)] = row
            self.mat_strength_params = a
            self._state.mat_strength_params = _ptr(a)

    def enable_stress(self):
        """Carry a stress field [2][E][3][3] (zero, as K/setup.cpp:214-220
        sets it) and the materials' strength models."""
        if self.elem_stress is None:
            self.elem_stress = np.zeros((2, self.E, 3, 3))
            self._state.elem_stress = _ptr(self.elem_stress)
        st = getattr(self.spec, "strength", None) or [0] * len(self.spec.materials)
        self.mat_strength = np.ascontiguousarray(st, np.int64)
        self._state.mat_strength = _ptr(self.mat_strength)

    # -- setup: K/driver.cpp:239 get_vol, then K/setup.cpp --
    def setup(self):
        L, m, s = self.L, C.byref(self._mesh), C.byref(self._state)
        spec = self.spec
        L.orc_get_vol(m, s)
        fills = (OrcFill * len(spec.fills))(*[OrcFill(**f) for f in spec.fills])
        sv = np.ascontiguousarray(spec.state_vars, np.float64)
        nsv = np.ascontiguousarray(spec.mat_num_state_vars, np.int64)
        L.orc_setup_fills(m, s, C.c_int64(len(spec.fills)), fills, _ptr(sv), _ptr(nsv))
        L.orc_boundary_velocity(m, s, self._bc)
        L.orc_node_mass(m, s)

    # -- operators (names follow the reference) --
    def get_vol(self):
        self.L.orc_get_vol(C.byref(self._mesh), C.byref(self._state))

    def get_divergence(self):
        self.L.orc_get_divergence(C.byref(self._mesh), C.byref(self._state))

    def get_force_sgh(self, dt, rk_alpha):
        r = self.spec.run
        self.L.orc_get_force_sgh(C.byref(self._mesh), C.byref(self._state), self._mat,
                                 C.c_double(r["fuzz"]), C.c_double(r["small"]),
                                 C.c_double(dt), C.c_double(rk_alpha))

    def update_velocity_sgh(self, rk_alpha, dt):
        self.L.orc_update_velocity_sgh(C.byref(self._mesh), C.byref(self._state),
                                       C.c_double(rk_alpha), C.c_double(dt))

    def boundary_velocity(self):
        self.L.orc_boundary_velocity(C.byref(self._mesh), C.byref(self._state), self._bc)

    def update_energy_sgh(self, rk_alpha, dt):
        self.L.orc_update_energy_sgh(C.byref(self._mesh), C.byref(self._state),
                                     C.c_double(rk_alpha), C.c_double(dt))

    def update_position_sgh(self, rk_alpha, dt):
        self.L.orc_update_position_sgh(C.byref(self._mesh), C.byref(self._state),
                                       C.c_double(rk_alpha), C.c_double(dt))

    def update_state(self, dt=0.0, rk_alpha=1.0):
        self.L.orc_update_state(C.byref(self._mesh), C.byref(self._state), C.c_double(dt),
                                C.c_double(rk_alpha))

    def rk_init(self):
        self.L.orc_rk_init(C.byref(self._mesh), C.byref(self._state))

    def get_timestep(self, time_value, graphics_time, dt):
        r = self.spec.run
        d = C.c_double(dt)
        self.L.orc_get_timestep(C.byref(self._mesh), C.byref(self._state),
                                C.c_double(time_value), C.c_double(graphics_time),
                                C.c_double(r["time_final"]), C.c_double(r["dt_max"]),
                                C.c_double(r["dt_min"]), C.c_double(r["dt_cfl"]), C.byref(d),
                                C.c_double(r["fuzz"]))
        return d.value

    def energy(self):
        ie, ke = C.c_double(), C.c_double()
        self.L.orc_energy_tallies(C.byref(self._mesh), C.byref(self._state), C.byref(ie),
                                  C.byref(ke))
        return ie.value, ke.value

    def solve(self, cycle_stop=None, dt_replay=None, record_dt=False, rk_num_stages=2,
              record_writes=False):
        """sgh_solve (K/sgh_solve.cpp:125-634).  ``record_dt`` returns the dt of
        every cycle as ``dt_history``; ``record_writes`` returns the cycles
        (1-based count of cycles done) after which graphics output is written
        as ``write_cycles``."""
        p, rec, wrec = _run_params(self.spec.run, cycle_stop, dt_replay, record_dt,
                                   rk_num_stages, record_writes)
        res = OrcRunResult()
        self.L.orc_sgh_solve(C.byref(self._mesh), C.byref(self._state), self._mat, self._bc,
                             C.byref(p), C.byref(res))
        return _run_result(res, rec, wrec)

    def stage(self, dt, rk_alpha):
        """One RK stage in the reference order (K/sgh_solve.cpp:355-486)."""
        self.get_divergence()
        self.get_force_sgh(dt, rk_alpha)
        self.update_velocity_sgh(rk_alpha, dt)
        self.boundary_velocity()
        self.update_energy_sgh(rk_alpha, dt)
        self.update_position_sgh(rk_alpha, dt)
        self.get_vol()
        self.update_state(dt, rk_alpha)


def bmatrix(x, y, z):
    B = np.zeros(24)
    lib().orc_get_bmatrix(B, np.ascontiguousarray(x, np.float64),
                          np.ascontiguousarray(y, np.float64), np.ascontiguousarray(z, np.float64))
    return B.reshape(8, 3)


def vol_hex(x, y, z):
    return lib().orc_get_vol_hex(np.ascontiguousarray(x, np.float64),
                                 np.ascontiguousarray(y, np.float64),
                                 np.ascontiguousarray(z, np.float64))


# ---------------------------------------------------------------------------
# 2D axisymmetric (RZ) path: rz_oracle.c
# ---------------------------------------------------------------------------
class OrzMesh(C.Structure):
    _fields_ = [("num_nodes", C.c_int64), ("num_elems", C.c_int64),
                ("nodes_in_elem", P), ("corners_in_node_start", P), ("corners_in_node", P),
                ("elems_in_elem_start", P), ("elems_in_elem", P),
                ("num_bdy_sets", C.c_int64), ("bdy_nodes_in_set_start", P),
                ("bdy_nodes_in_set", P), ("num_bdy_nodes", C.c_int64), ("bdy_nodes", P),
                ("nodes_in_node_start", P), ("nodes_in_node", P)]


def _rz_lib(fast=False, strength_model=False):
    L = lib(fast, strength_model)
    if not getattr(L, "_rz_ready", False):
        L.orz_get_vol_quad.restype = C.c_double
        L.orz_get_area_quad.restype = C.c_double
        L.orz_heron.restype = C.c_double
        for f in ("orz_get_vol_quad", "orz_get_area_quad"):
            getattr(L, f).argtypes = [f64p, f64p]
        L.orz_heron.argtypes = [C.c_double] * 6
        L.orz_get_bmatrix2D.argtypes = [f64p, f64p, f64p]
        L.orz_get_area_weights2D.argtypes = [f64p, f64p, f64p]
        L.orz_build_elem_elem_connectivity.restype = C.c_int64
        L.orz_build_patch_connectivity.restype = C.c_int64
        L.orz_build_bdy_sets.restype = C.c_int64
        L.orz_read_ensight.restype = C.c_int
        L._rz_ready = True
    return L


def gen_rect(nx, ny, lx_num, ly_num, l_den):
    N, E = (nx + 1) * (ny + 1), nx * ny
    coords = np.zeros((N, 2))
    conn = np.zeros((E, 4), np.int64)
    _rz_lib().orz_gen_rect(C.c_int64(nx), C.c_int64(ny), C.c_double(lx_num), C.c_double(ly_num),
                           C.c_double(l_den), _ptr(coords), _ptr(conn))
    return coords, conn


def read_ensight_2d(path: str):
    L = _rz_lib()
    nn, ne = C.c_int64(), C.c_int64()
    rc = L.orz_read_ensight(path.encode(), C.byref(nn), C.byref(ne), None, None)
    if rc != 0:
        raise IOError(f"orz_read_ensight({path}) -> {rc}")
    coords = np.zeros((nn.value, 2))
    conn = np.zeros((ne.value, 4), np.int64)
    rc = L.orz_read_ensight(path.encode(), C.byref(nn), C.byref(ne), _ptr(coords), _ptr(conn))
    if rc != 0:
        raise IOError(f"orz_read_ensight({path}) -> {rc}")
    return coords, conn


def connectivity_rz(coords: np.ndarray, conn: np.ndarray):
    """The reference's 2D connectivity (K/mesh.h:180-796, num_dims = 2)."""
    L = _rz_lib()
    N, E = coords.shape[0], conn.shape[0]
    conn = np.ascontiguousarray(conn, np.int64)
    cin_start = np.zeros(N + 1, np.int64)
    cin = np.zeros(4 * E, np.int64)
    ein = np.zeros(4 * E, np.int64)
    L.orz_build_corner_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(cin_start),
                                    _ptr(cin), _ptr(ein))
    ee_start = np.zeros(E + 1, np.int64)
    nb = L.orz_build_elem_elem_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn),
                                            _ptr(cin_start), _ptr(ein), _ptr(ee_start), None)
    ee = np.zeros(max(nb, 1), np.int64)
    L.orz_build_elem_elem_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(cin_start),
                                       _ptr(ein), _ptr(ee_start), _ptr(ee))
    counts = np.zeros(3, np.int64)
    L.orz_build_patch_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(ee_start),
                                   _ptr(ee), None, None, None, _ptr(counts))
    P_, BP, BN = (int(c) for c in counts)
    nip = np.zeros((max(P_, 1), 2), np.int64)
    bp = np.zeros(max(BP, 1), np.int64)
    bn = np.zeros(max(BN, 1), np.int64)
    L.orz_build_patch_connectivity(C.c_int64(N), C.c_int64(E), _ptr(conn), _ptr(ee_start),
                                   _ptr(ee), _ptr(nip), _ptr(bp), _ptr(bn), _ptr(counts))
    nn_start = np.zeros(N + 1, np.int64)
    nn = np.zeros(max(2 * P_, 1), np.int64)
    L.orz_build_node_node_connectivity(C.c_int64(N), C.c_int64(P_), _ptr(nip), _ptr(nn_start),
                                       _ptr(nn))
    return dict(corners_in_node_start=cin_start, corners_in_node=cin, elems_in_node=ein,
                elems_in_elem_start=ee_start, elems_in_elem=ee[:nb], nodes_in_patch=nip[:P_],
                bdy_patches=bp[:BP], bdy_nodes=bn[:BN], num_patches=P_,
                nodes_in_node_start=nn_start, nodes_in_node=nn[:2 * P_])


def bdy_sets_rz(coords, cx, bcs):
    L = _rz_lib()
    ns = len(bcs)
    arr = (OrcBoundary * max(ns, 1))(*[OrcBoundary(int(b[0]), float(b[1]), int(b[2])) for b in bcs])
    start = np.zeros(ns + 1, np.int64)
    cap = max(ns * max(len(cx["bdy_patches"]), 1) * 2, 1)
    out = np.zeros(cap, np.int64)
    xy = np.ascontiguousarray(coords, np.float64)
    tot = L.orz_build_bdy_sets(_ptr(xy), C.c_int64(len(cx["bdy_patches"])),
                               _ptr(cx["bdy_patches"]), _ptr(cx["nodes_in_patch"]),
                               C.c_int64(ns), arr, C.c_int64(xy.shape[0]), _ptr(start), _ptr(out))
    return start, out[:tot].copy(), arr


class OracleRZ:
    """One 2D RZ problem in the reference layout (num_dims = 2), set up and
    run as K/driver.cpp and K/sgh_solve.cpp do on that branch."""

    def __init__(self, spec, fast: bool = False, coords=None, conn=None,
                 strength_model: bool = False):
        self.L = _rz_lib(fast, strength_model)
        self.spec = spec
        if coords is None:
            coords, conn = spec.mesh_arrays(oracle_gen=gen_rect, oracle_read=read_ensight_2d)
        self.coords0 = np.ascontiguousarray(coords, np.float64)
        self.conn = np.ascontiguousarray(conn, np.int64)
        N, E = self.coords0.shape[0], self.conn.shape[0]
        self.N, self.E = N, E
        self.cx = connectivity_rz(self.coords0, self.conn)
        self.bset_start, self.bset_nodes, self._bc = bdy_sets_rz(self.coords0, self.cx, spec.bcs)
        nsv = spec.num_state_vars
        self.node_coords = np.zeros((2, N, 2))
        self.node_coords[0] = self.coords0
        self.node_coords[1] = self.coords0
        self.node_vel = np.zeros((2, N, 2))
        self.node_mass = np.zeros(N)
        self.node_ext = np.zeros(N)
        z = lambda: np.zeros(E)  # noqa: E731
        self.elem_den, self.elem_pres, self.elem_sspd = z(), z(), z()
        self.elem_vol, self.elem_div, self.elem_mass = z(), z(), z()
        self.elem_sie = np.zeros((2, E))
        self.elem_mat_id = np.zeros(E, np.int64)
        self.elem_statev = np.zeros((E, nsv))
        self.corner_force = np.zeros((4 * E, 2))
        cx = self.cx
        self._mesh = OrzMesh(N, E, _ptr(self.conn), _ptr(cx["corners_in_node_start"]),
                             _ptr(cx["corners_in_node"]), _ptr(cx["elems_in_elem_start"]),
                             _ptr(cx["elems_in_elem"]), len(spec.bcs), _ptr(self.bset_start),
                             _ptr(self.bset_nodes), len(cx["bdy_nodes"]), _ptr(cx["bdy_nodes"]),
                             _ptr(cx["nodes_in_node_start"]), _ptr(cx["nodes_in_node"]))
        self._state = OrcState(_ptr(self.node_coords), _ptr(self.node_vel), _ptr(self.node_mass),
                               _ptr(self.elem_den), _ptr(self.elem_pres), _ptr(self.elem_sspd),
                               _ptr(self.elem_sie), _ptr(self.elem_vol), _ptr(self.elem_div),
                               _ptr(self.elem_mass), _ptr(self.elem_mat_id),
                               _ptr(self.elem_statev), nsv, _ptr(self.corner_force), None, None,
                               None, None)
        OracleProblem._set_strength_params(self, spec)
        if spec.eos:
            self.mat_eos = np.ascontiguousarray(spec.eos, np.int64)
            self._state.mat_eos = _ptr(self.mat_eos)
        self.elem_stress = None
        if getattr(spec, "strength", None):
            OracleProblem.enable_stress(self)
        mats = spec.materials
        self._mat = (OrcMaterial * len(mats))(*[OrcMaterial(*m) for m in mats])
        self.setup()

    def _ms(self):
        return C.byref(self._mesh), C.byref(self._state)

    enable_stress = OracleProblem.enable_stress

    def setup(self):
        L, spec = self.L, self.spec
        m, s = self._ms()
        L.orz_get_vol(m, s)
        fills = (OrcFill * len(spec.fills))(*[OrcFill(**f) for f in spec.fills])
        sv = np.ascontiguousarray(spec.state_vars, np.float64)
        nsv = np.ascontiguousarray(spec.mat_num_state_vars, np.int64)
        L.orz_setup_fills(m, s, C.c_int64(len(spec.fills)), fills, _ptr(sv), _ptr(nsv))
        L.orz_boundary_velocity(m, s, self._bc)
        L.orz_node_mass(m, s)

    def get_vol(self):
        self.L.orz_get_vol(*self._ms())

    def get_divergence(self):
        self.L.orz_get_divergence(*self._ms())

    def get_force_sgh(self, dt, rk_alpha):
        r = self.spec.run
        self.L.orz_get_force(*self._ms(), self._mat, C.c_double(r["fuzz"]), C.c_double(r["small"]),
                             C.c_double(dt), C.c_double(rk_alpha))

    def update_velocity_sgh(self, rk_alpha, dt):
        self.L.orz_update_velocity(*self._ms(), C.c_double(rk_alpha), C.c_double(dt))

    def boundary_velocity(self):
        self.L.orz_boundary_velocity(*self._ms(), self._bc)

    def update_energy_sgh(self, rk_alpha, dt):
        self.L.orz_update_energy(*self._ms(), C.c_double(rk_alpha), C.c_double(dt))

    def update_position_sgh(self, rk_alpha, dt):
        self.L.orz_update_position(*self._ms(), C.c_double(rk_alpha), C.c_double(dt))

    def update_state(self, dt=0.0, rk_alpha=1.0):
        self.L.orz_update_state(*self._ms())

    def rk_init(self):
        self.L.orz_rk_init(*self._ms())

    def extensive_mass(self):
        m, s = self._ms()
        self.L.orz_extensive_mass(m, s, _ptr(self.node_ext))

    def update_node_mass(self, tiny=None):
        tiny = self.spec.run["tiny"] if tiny is None else tiny
        m, s = self._ms()
        self.L.orz_update_node_mass(m, s, _ptr(self.node_ext), C.c_double(tiny))

    def get_timestep(self, time_value, graphics_time, dt):
        r = self.spec.run
        d = C.c_double(dt)
        self.L.orz_get_timestep(*self._ms(), C.c_double(time_value), C.c_double(graphics_time),
                                C.c_double(r["time_final"]), C.c_double(r["dt_max"]),
                                C.c_double(r["dt_min"]), C.c_double(r["dt_cfl"]), C.byref(d),
                                C.c_double(r["fuzz"]))
        return d.value

    def energy(self):
        ie, ke = C.c_double(), C.c_double()
        self.L.orz_energy_tallies(*self._ms(), C.byref(ie), C.byref(ke))
        return ie.value, ke.value

    def solve(self, cycle_stop=None, dt_replay=None, record_dt=False, rk_num_stages=2,
              record_writes=False):
        """The 2D branch of sgh_solve; arguments as ``OracleProblem.solve``."""
        p, rec, wrec = _run_params(self.spec.run, cycle_stop, dt_replay, record_dt,
                                   rk_num_stages, record_writes)
        res = OrcRunResult()
        self.L.orz_sgh_solve(*self._ms(), self._mat, self._bc, C.byref(p),
                             C.c_double(self.spec.run["tiny"]), C.byref(res))
        return _run_result(res, rec, wrec)

    def stage(self, dt, rk_alpha):
        """One RK stage of the 2D branch (K/sgh_solve.cpp:355-581)."""
        self.get_divergence()
        self.get_force_sgh(dt, rk_alpha)
        self.update_velocity_sgh(rk_alpha, dt)
        self.boundary_velocity()
        self.update_energy_sgh(rk_alpha, dt)
        self.update_position_sgh(rk_alpha, dt)
        self.get_vol()
        self.update_state()
        self.update_node_mass()


def rz_geometry(x, y):
    """(B [4][2], area, RZ volume, corner areas [4]) of one quad."""
    L = _rz_lib()
    x = np.ascontiguousarray(x, np.float64)
    y = np.ascontiguousarray(y, np.float64)
    B = np.zeros(8)
    A = np.zeros(4)
    L.orz_get_bmatrix2D(B, x, y)
    L.orz_get_area_weights2D(A, x, y)
    return B.reshape(4, 2), L.orz_get_area_quad(x, y), L.orz_get_vol_quad(x, y), A


def velgrad_3d(pr):
    """Velocity gradient of every element of a 3D OracleProblem (the
    reference's get_velgrad with the outward B), E x 3 x 3."""
    L = lib()
    out = np.zeros((pr.E, 3, 3))
    X, V = pr.node_coords[1], pr.node_vel[1]
    for e in range(pr.E):
        nd = pr.conn[e]
        B = bmatrix(X[nd, 0], X[nd, 1], X[nd, 2]).ravel()
        u, v, w = (np.ascontiguousarray(V[nd, d]) for d in range(3))
        o = np.zeros(9)
        L.orc_get_velgrad(_ptr(o), _ptr(u), _ptr(v), _ptr(w), _ptr(np.ascontiguousarray(B)),
                          C.c_double(pr.elem_vol[e]))
        out[e] = o.reshape(3, 3)
    return out


def velgrad_rz(pr):
    """Velocity gradient of every element of an OracleRZ (get_velgrad2D)."""
    L = _rz_lib()
    out = np.zeros((pr.E, 3, 3))
    X, V = pr.node_coords[1], pr.node_vel[1]
    for e in range(pr.E):
        nd = pr.conn[e]
        x, y = np.ascontiguousarray(X[nd, 0]), np.ascontiguousarray(X[nd, 1])
        B, area, _, _ = rz_geometry(x, y)
        u, v = np.ascontiguousarray(V[nd, 0]), np.ascontiguousarray(V[nd, 1])
        o = np.zeros(9)
        L.orz_get_velgrad2D(_ptr(o), _ptr(u), _ptr(v), _ptr(np.ascontiguousarray(B.ravel())),
                            C.c_double(pr.elem_vol[e]), C.c_double(area))
        out[e] = o.reshape(3, 3)
    return out

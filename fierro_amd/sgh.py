"""Host-side mirror of Fierro's SGH operator interface over the HIP library.

``SGHSolver`` plays the role of the reference driver (K/driver.cpp:94-313):
it builds the mesh, boundary sets and initial conditions through the C ABI
(include/fsgh.h) and exposes the per-stage operators under the reference's
names (K/mesh.h:938-1406) plus the fused cycle loop ``sgh_solve``.
Every method runs on the GPU through libfsgh.so; there is no CPU path.

A list of ProblemSpecs makes one batched context: the instances are packed as
a ragged mesh (nodes and elements concatenated in instance order), each with
its own boundary sets, materials, dt and stop state.  Specs with ``dims == 2``
make a 2D axisymmetric (RZ) context: quads, node fields with 2 components
(axial, radius), corner forces [4E][2].
"""
from __future__ import annotations

import ctypes as C

import numpy as np

from . import _abi
from ._abi import check

FIELD_SHAPES = {
    _abi.NODE_COORDS: ("node", (2, None, "D"), np.float64),
    _abi.NODE_VEL: ("node", (2, None, "D"), np.float64),
    _abi.NODE_MASS: ("node", (None,), np.float64),
    _abi.ELEM_DEN: ("elem", (None,), np.float64),
    _abi.ELEM_PRES: ("elem", (None,), np.float64),
    _abi.ELEM_SSPD: ("elem", (None,), np.float64),
    _abi.ELEM_SIE: ("elem", (2, None), np.float64),
    _abi.ELEM_VOL: ("elem", (None,), np.float64),
    _abi.ELEM_DIV: ("elem", (None,), np.float64),
    _abi.ELEM_MASS: ("elem", (None,), np.float64),
    _abi.ELEM_MAT_ID: ("elem", (None,), np.int64),
    _abi.ELEM_GAMMA: ("elem", (None,), np.float64),
    _abi.ELEM_CSMIN: ("elem", (None,), np.float64),
    _abi.CORNER_FORCE: ("corner", (None, "D"), np.float64),
    _abi.ELEM_STRESS: ("elem", (2, None, 3, 3), np.float64),
}
FIELDS = {"node_coords": _abi.NODE_COORDS, "node_vel": _abi.NODE_VEL,
          "node_mass": _abi.NODE_MASS, "elem_den": _abi.ELEM_DEN, "elem_pres": _abi.ELEM_PRES,
          "elem_sspd": _abi.ELEM_SSPD, "elem_sie": _abi.ELEM_SIE, "elem_vol": _abi.ELEM_VOL,
          "elem_div": _abi.ELEM_DIV, "elem_mass": _abi.ELEM_MASS,
          "elem_mat_id": _abi.ELEM_MAT_ID, "elem_gamma": _abi.ELEM_GAMMA,
          "elem_csmin": _abi.ELEM_CSMIN, "corner_force": _abi.CORNER_FORCE,
          "elem_stress": _abi.ELEM_STRESS}


def _p(a: np.ndarray):
    return a.ctypes.data_as(C.c_void_p)


def _dims(specs):
    dims = {getattr(s, "dims", 3) for s in specs}
    if len(dims) != 1:
        raise ValueError("a batch holds problems of one dimension")
    return dims.pop()


def pack(meshes):
    """Concatenate (coords, conn) pairs into one ragged mesh with offsets."""
    node_off, elem_off = [0], [0]
    xs, cs = [], []
    for coords, conn in meshes:
        cs.append(conn + node_off[-1])
        xs.append(coords)
        node_off.append(node_off[-1] + coords.shape[0])
        elem_off.append(elem_off[-1] + conn.shape[0])
    return (np.ascontiguousarray(np.concatenate(xs)), np.ascontiguousarray(np.concatenate(cs)),
            np.array(node_off, np.int64), np.array(elem_off, np.int64))


class SGHSolver:
    def __init__(self, specs, device: int = 0, renumber: bool = True, meshes=None,
                 library: str | None = None):
        # library: another build of the ABI (tests load the strength-model test variant)
        self.L = _abi.load(library)
        if self.L.fsgh_device_count() < 1:
            raise RuntimeError("no HIP device visible to libfsgh.so")
        if not isinstance(specs, (list, tuple)):
            specs = [specs]
        self.specs = list(specs)
        self.dims = _dims(self.specs)
        if meshes is None:
            meshes = [s.mesh_arrays() for s in self.specs]
        coords, conn, noff, eoff = pack(meshes)
        self.coords0, self.conn = coords, conn
        self.node_off, self.elem_off = noff, eoff
        self.N, self.E = coords.shape[0], conn.shape[0]
        self.ninst = len(self.specs)
        desc = _abi.MeshDesc(self.N, self.E, self.ninst, 1 if renumber else 0,
                             _p(noff).value if self.ninst > 1 else None,
                             _p(eoff).value if self.ninst > 1 else None, self.dims, 0)
        ctx = C.c_void_p()
        rc = self.L.fsgh_create(C.byref(desc), _p(conn), _p(coords), device, C.byref(ctx))
        self.ctx = ctx
        check(rc, ctx if ctx.value else None, "fsgh_create")
        self._setup()

    # -- K/driver.cpp:245 setup() per instance --
    def _setup(self):
        mats, index = [], []
        for s in self.specs:
            index.append(len(mats))
            mats.extend(s.materials)
        arr = (_abi.Material * len(mats))(*[_abi.Material(*m) for m in mats])
        check(self.L.fsgh_set_materials(self.ctx, len(mats), arr), self.ctx, "set_materials")
        for i, s in enumerate(self.specs):
            eos = list(s.eos or [])
            strength = list(getattr(s, "strength", None) or [])
            for k in range(len(s.materials)):
                kind = eos[k] if k < len(eos) else _abi.EOS_IDEAL_GAS
                st = strength[k] if k < len(strength) else _abi.STRENGTH_NONE
                if kind != _abi.EOS_IDEAL_GAS or st != _abi.STRENGTH_NONE:
                    check(self.L.fsgh_set_material_models(self.ctx, index[i] + k, int(kind),
                                                          int(st)),
                          self.ctx, "set_material_models")
            for k, params in enumerate(getattr(s, "strength_params", None) or []):
                if len(params):
                    self.set_strength_params(index[i] + k, params)
        for i, s in enumerate(self.specs):
            bcs = (_abi.Boundary * max(len(s.bcs), 1))(
                *[_abi.Boundary(int(b[0]), int(b[2]), float(b[1])) for b in s.bcs])
            check(self.L.fsgh_set_bcs(self.ctx, i, len(s.bcs), bcs), self.ctx, "set_bcs")
            fl = []
            for f in s.fills:
                d = dict(f)
                d["mat_id"] = int(d["mat_id"]) + index[i]
                fl.append(_abi.Fill(**{k: d[k] for k, _ in _abi.Fill._fields_ if k != "pad_"}))
            farr = (_abi.Fill * max(len(fl), 1))(*fl)
            nmat = len(mats)
            nsv = s.num_state_vars
            sv = np.zeros((nmat, nsv))
            msv = np.zeros(nmat, np.int32)
            sv[index[i]:index[i] + len(s.materials)] = s.state_vars[:, :nsv]
            msv[index[i]:index[i] + len(s.materials)] = s.mat_num_state_vars
            check(self.L.fsgh_setup_fills(self.ctx, i, len(fl), farr, nmat, nsv, _p(sv), _p(msv)),
                  self.ctx, "setup_fills")

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.fsgh_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    # -- fields --
    def _shape(self, fid):
        kind, shp, dt = FIELD_SHAPES[fid]
        npe = 4 if self.dims == 2 else 8
        n = {"node": self.N, "elem": self.E, "corner": npe * self.E}[kind]
        return tuple(n if s is None else (self.dims if s == "D" else s) for s in shp), dt

    def download(self, name):
        fid = FIELDS[name]
        shp, dt = self._shape(fid)
        out = np.zeros(shp, dt)
        check(self.L.fsgh_download(self.ctx, fid, _p(out), out.nbytes), self.ctx, name)
        return out

    def upload(self, name, arr):
        fid = FIELDS[name]
        shp, dt = self._shape(fid)
        a = np.ascontiguousarray(arr, dt)
        if a.shape != shp:
            raise ValueError(f"{name}: shape {a.shape} != {shp}")
        check(self.L.fsgh_upload(self.ctx, fid, _p(a), a.nbytes), self.ctx, name)

    def corners_in_node(self):
        start = np.zeros(self.N + 1, np.int64)
        corners = np.zeros((4 if self.dims == 2 else 8) * self.E, np.int64)
        check(self.L.fsgh_get_corners_in_node(self.ctx, _p(start), _p(corners)), self.ctx, "cin")
        return start, corners

    def renumbering(self):
        e = np.zeros(self.E, np.int64)
        n = np.zeros(self.N, np.int64)
        check(self.L.fsgh_get_renumbering(self.ctx, _p(e), _p(n)), self.ctx, "renumbering")
        return e, n

    def stage_info(self, elem_block=256):
        """(list entries, largest block) of the element kernels' per-block
        node lists for blocks of elem_block elements (zeros if not built)."""
        n, mx = C.c_int64(), C.c_int32()
        check(self.L.fsgh_get_stage_info(self.ctx, int(elem_block), C.byref(n), C.byref(mx)),
              self.ctx, "stage_info")
        return n.value, mx.value

    def device_csr(self):
        """The device's corner CSR in internal numbering (diagnostic)."""
        start = np.zeros(self.N + 1, np.int32)
        corners = np.zeros((4 if self.dims == 2 else 8) * self.E, np.int32)
        check(self.L.fsgh_get_device_csr(self.ctx, _p(start), _p(corners)), self.ctx, "device_csr")
        return start, corners

    def bc_set(self, inst, s):
        c = C.c_int64()
        check(self.L.fsgh_bc_set_size(self.ctx, inst, s, C.byref(c)), self.ctx, "bc size")
        out = np.zeros(c.value, np.int64)
        check(self.L.fsgh_get_bc_set(self.ctx, inst, s, _p(out)), self.ctx, "bc set")
        return out

    def bc_set_sizes(self, inst=0):
        out = []
        for s in range(len(self.specs[inst].bcs)):
            c = C.c_int64()
            check(self.L.fsgh_bc_set_size(self.ctx, inst, s, C.byref(c)), self.ctx, "bc size")
            out.append(c.value)
        return out

    # -- per-stage operators, reference names --
    def rk_init(self):
        check(self.L.fsgh_rk_init(self.ctx), self.ctx, "rk_init")

    def get_timestep(self, time_value, graphics_time, time_final, dt_max, dt_min, dt_cfl, dt,
                     fuzz):
        d = C.c_double(dt)
        check(self.L.fsgh_get_timestep(self.ctx, time_value, graphics_time, time_final, dt_max,
                                       dt_min, dt_cfl, fuzz, C.byref(d)), self.ctx, "get_timestep")
        return d.value

    def get_divergence(self):
        check(self.L.fsgh_get_divergence(self.ctx), self.ctx, "get_divergence")

    def get_force_sgh(self, fuzz, small, dt, rk_alpha):
        check(self.L.fsgh_get_force_sgh(self.ctx, fuzz, small, dt, rk_alpha), self.ctx, "force")

    def update_velocity_sgh(self, rk_alpha, dt):
        check(self.L.fsgh_update_velocity_sgh(self.ctx, rk_alpha, dt), self.ctx, "velocity")

    def boundary_velocity(self):
        check(self.L.fsgh_boundary_velocity(self.ctx), self.ctx, "boundary_velocity")

    def update_energy_sgh(self, rk_alpha, dt):
        check(self.L.fsgh_update_energy_sgh(self.ctx, rk_alpha, dt), self.ctx, "energy")

    def update_position_sgh(self, rk_alpha, dt):
        check(self.L.fsgh_update_position_sgh(self.ctx, rk_alpha, dt), self.ctx, "position")

    def get_vol(self):
        check(self.L.fsgh_get_vol(self.ctx), self.ctx, "get_vol")

    def update_state(self, dt=0.0, rk_alpha=1.0):
        check(self.L.fsgh_update_state(self.ctx, dt, rk_alpha), self.ctx, "update_state")

    def energy(self, inst=0):
        ie, ke = C.c_double(), C.c_double()
        check(self.L.fsgh_energy(self.ctx, inst, C.byref(ie), C.byref(ke)), self.ctx, "energy")
        return ie.value, ke.value

    def check_state(self):
        return self.L.fsgh_check(self.ctx)

    def stage(self, dt, rk_alpha):
        """One RK stage through the operator API, in the reference's order
        (K/sgh_solve.cpp:355-486)."""
        r = self.specs[0].run
        self.get_divergence()
        self.get_force_sgh(r["fuzz"], r["small"], dt, rk_alpha)
        self.update_velocity_sgh(rk_alpha, dt)
        self.boundary_velocity()
        self.update_energy_sgh(rk_alpha, dt)
        self.update_position_sgh(rk_alpha, dt)
        self.get_vol()
        self.update_state(dt, rk_alpha)
        if self.dims == 2:
            self.rz_update_node_mass(r["tiny"])

    def get_velgrad(self):
        """Velocity gradient of every element, E x 3 x 3 (caller numbering)."""
        out = np.zeros((self.E, 3, 3))
        check(self.L.fsgh_get_velgrad(self.ctx, _p(out)), self.ctx, "get_velgrad")
        return out

    def set_strength_params(self, material, params):
        p = np.ascontiguousarray(params, np.float64)
        check(self.L.fsgh_set_strength_params(self.ctx, int(material), p.size, _p(p)), self.ctx,
              "set_strength_params")

    # -- 2D RZ node mass (K/sgh_solve.cpp:107-116, 495-581) --
    def rz_extensive_mass(self):
        check(self.L.fsgh_rz_extensive_mass(self.ctx), self.ctx, "rz_extensive_mass")

    def rz_update_node_mass(self, tiny=1.0e-12):
        check(self.L.fsgh_rz_update_node_mass(self.ctx, float(tiny)), self.ctx, "rz_node_mass")

    # -- fused cycle loop --
    def _run_params(self):
        ps = []
        for s in self.specs:
            r = s.run
            ps.append(_abi.RunParams(r["time_final"], r["dt_start"], r["dt_max"], r["dt_min"],
                                     r["dt_cfl"], r["graphics_dt_ival"],
                                     int(r["graphics_cyc_ival"]), int(r["cycle_stop"]),
                                     r["fuzz"], r["small"], r.get("tiny", 1.0e-12)))
        return (_abi.RunParams * len(ps))(*ps)

    def run_begin(self):
        check(self.L.fsgh_run_begin(self.ctx, self._run_params()), self.ctx, "run_begin")

    def run(self, ncycles, use_graph=True, stop_on_output=False, persistent=False):
        """fsgh_run; persistent: one cooperative launch per cycle
        (FSGH_RUN_PERSISTENT), which ignores use_graph."""
        flags = (_abi.RUN_GRAPH if use_graph else 0) | \
            (_abi.RUN_STOP_ON_OUTPUT if stop_on_output else 0) | \
            (_abi.RUN_PERSISTENT if persistent else 0)
        check(self.L.fsgh_run(self.ctx, int(ncycles), flags), self.ctx, "run")

    def status(self, inst=0):
        st = _abi.InstanceStatus()
        check(self.L.fsgh_run_status(self.ctx, inst, C.byref(st)), self.ctx, "status")
        return {k: getattr(st, k) for k, _ in _abi.InstanceStatus._fields_}

    def sgh_solve(self, cycles=None, use_graph=True, output=None, persistent=False):
        """The cycle loop of K/sgh_solve.cpp:125-634 (fused kernels).  With
        ``output(solver, time)`` (single instance) the dumps are written at
        t=0 and at every cycle meeting the output condition, as
        write_outputs is called at K/sgh_solve.cpp:45,609-628."""
        ie0 = [self.energy(i) for i in range(self.ninst)]
        self.run_begin()
        n = cycles if cycles is not None else max(int(s.run["cycle_stop"]) for s in self.specs)
        if output is None:
            self.run(n, use_graph, persistent=persistent)
        else:
            output(self, 0.0)
            done = 0
            while done < n:
                before = self.status(0)["cycles"]
                self.run(n - done, use_graph, stop_on_output=True, persistent=persistent)
                st = self.status(0)
                done += st["cycles"] - before
                if st["write_pending"]:
                    output(self, st["time_value"])
                if st["stopped"] or st["cycles"] == before:
                    break
        ie1 = [self.energy(i) for i in range(self.ninst)]
        return [dict(ie0=a[0], ke0=a[1], te0=a[0] + a[1], ie1=b[0], ke1=b[1], te1=b[0] + b[1],
                     **self.status(i)) for i, (a, b) in enumerate(zip(ie0, ie1))]

    # -- discrete adjoint / tangent-linear model (frozen time steps) --
    def forward(self, ncycles, checkpoint_every=20, dt_replay=None):
        """Runs ``ncycles`` cycles of every instance from the set-up state,
        recording dt and checkpointing the state every ``checkpoint_every``
        cycles; ``dt_replay`` ([ncycles][instances]) replays given steps.
        Raises FsghError if the state was advanced since the set-up or the
        last uploads (set it again first)."""
        rep = None
        if dt_replay is not None:
            rep = np.ascontiguousarray(dt_replay, np.float64).reshape(-1)
            if rep.size != ncycles * self.ninst:
                raise ValueError("dt_replay must hold ncycles x instances steps")
        check(self.L.fsgh_forward(self.ctx, self._run_params(), int(ncycles),
                                  int(checkpoint_every), None if rep is None else _p(rep)),
              self.ctx, "forward")
        self._adj_ncyc = int(ncycles)

    def dt_history(self):
        out = np.zeros((self._adj_ncyc, self.ninst))
        check(self.L.fsgh_dt_history(self.ctx, _p(out), out.size), self.ctx, "dt_history")
        return out

    def backward(self, seed_x=None, seed_v=None, seed_e=None, seed_den=None):
        """Gradients of <seed, final state> (coordinates, velocities,
        energy, density M/V) with respect to the initial density, specific
        internal energy, gamma and node velocity."""
        def arr(a, shape):
            return None if a is None else np.ascontiguousarray(a, np.float64).reshape(shape)
        sx, sv = arr(seed_x, (self.N, 3)), arr(seed_v, (self.N, 3))
        se, sd = arr(seed_e, (self.E,)), arr(seed_den, (self.E,))
        g = dict(den0=np.zeros(self.E), sie0=np.zeros(self.E), gamma=np.zeros(self.E),
                 vel0=np.zeros((self.N, 3)))
        ptr = [None if a is None else _p(a) for a in (sx, sv, se, sd)]
        check(self.L.fsgh_backward(self.ctx, *ptr, _p(g["den0"]), _p(g["sie0"]),
                                   _p(g["gamma"]), _p(g["vel0"])), self.ctx, "backward")
        return g

    def tangent(self, d_den0=None, d_sie0=None, d_gamma=None, d_vel0=None):
        """Directional derivative of the final (coords, vel, sie) along the
        given perturbation of the initial state."""
        def arr(a, shape):
            return None if a is None else np.ascontiguousarray(a, np.float64).reshape(shape)
        ins = [arr(d_den0, (self.E,)), arr(d_sie0, (self.E,)), arr(d_gamma, (self.E,)),
               arr(d_vel0, (self.N, 3))]
        t = dict(x=np.zeros((self.N, 3)), v=np.zeros((self.N, 3)), e=np.zeros(self.E))
        check(self.L.fsgh_tangent(self.ctx, *[None if a is None else _p(a) for a in ins],
                                  _p(t["x"]), _p(t["v"]), _p(t["e"])), self.ctx, "tangent")
        return t

    # -- launch configuration --
    @staticmethod
    def _tuning_dicts(arr):
        return {_abi.TUNED_KERNEL_NAMES[t.kernel]: {k: getattr(t, k) for k, _ in _abi.Tuning._fields_
                                                   if k != "kernel"} for t in arr}

    def autotune(self, reps=10):
        """Time every compiled launch variant of the four fused-cycle kernels on
        the current state and keep the fastest (the state is left unchanged).
        Returns {kernel: {block, min_waves, xcd_order, ms, ms_default}}."""
        arr = (_abi.Tuning * _abi.NUM_TUNED_KERNELS)()
        check(self.L.fsgh_autotune(self.ctx, int(reps), arr), self.ctx, "autotune")
        return self._tuning_dicts(arr)

    def set_tuning(self, cfg):
        """cfg: {kernel name: {block, min_waves, xcd_order}} (missing keys keep
        the current value)."""
        cur = self.get_tuning()
        items = []
        for name, v in cfg.items():
            k = _abi.TUNED_KERNEL_NAMES.index(name)
            c = dict(cur[name])
            c.update(v)
            items.append(_abi.Tuning(k, int(c["block"]), int(c["min_waves"]),
                                     int(c["xcd_order"]), 0.0, 0.0))
        arr = (_abi.Tuning * len(items))(*items)
        check(self.L.fsgh_set_tuning(self.ctx, arr, len(items)), self.ctx, "set_tuning")

    def get_tuning(self):
        arr = (_abi.Tuning * _abi.NUM_TUNED_KERNELS)()
        check(self.L.fsgh_get_tuning(self.ctx, arr, _abi.NUM_TUNED_KERNELS), self.ctx,
              "get_tuning")
        return self._tuning_dicts(arr)

    # -- profiling --
    def profile(self, enable=True):
        check(self.L.fsgh_profile_enable(self.ctx, 1 if enable else 0), self.ctx, "profile")

    def profile_read(self):
        names = C.create_string_buffer(4096)
        ms = np.zeros(64)
        calls = np.zeros(64, np.int64)
        cnt = C.c_int32()
        check(self.L.fsgh_profile_read(self.ctx, 64, names, 4096, _p(ms), _p(calls),
                                       C.byref(cnt)), self.ctx, "profile_read")
        nm = names.value.decode().split("\n")[:cnt.value]
        return {k: (float(ms[i]), int(calls[i])) for i, k in enumerate(nm)}


class MeshMaps:
    """Host-only context (device = -1): the library's connectivity, Morton
    renumbering and boundary sets, without touching a GPU."""

    def __init__(self, specs, renumber: bool = True, meshes=None):
        self.L = _abi.load()
        if not isinstance(specs, (list, tuple)):
            specs = [specs]
        self.specs = list(specs)
        self.dims = _dims(self.specs)
        if meshes is None:
            meshes = [s.mesh_arrays() for s in self.specs]
        coords, conn, noff, eoff = pack(meshes)
        self.coords0, self.conn, self.node_off, self.elem_off = coords, conn, noff, eoff
        self.N, self.E = coords.shape[0], conn.shape[0]
        n = len(self.specs)
        desc = _abi.MeshDesc(self.N, self.E, n, 1 if renumber else 0,
                             _p(noff).value if n > 1 else None, _p(eoff).value if n > 1 else None,
                             self.dims, 0)
        ctx = C.c_void_p()
        rc = self.L.fsgh_create(C.byref(desc), _p(conn), _p(coords), -1, C.byref(ctx))
        self.ctx = ctx
        check(rc, ctx if ctx.value else None, "fsgh_create(host-only)")
        for i, s in enumerate(self.specs):
            bcs = (_abi.Boundary * max(len(s.bcs), 1))(
                *[_abi.Boundary(int(b[0]), int(b[2]), float(b[1])) for b in s.bcs])
            check(self.L.fsgh_set_bcs(self.ctx, i, len(s.bcs), bcs), self.ctx, "set_bcs")

    corners_in_node = SGHSolver.corners_in_node
    renumbering = SGHSolver.renumbering

    def bc_set(self, inst, s):
        c = C.c_int64()
        check(self.L.fsgh_bc_set_size(self.ctx, inst, s, C.byref(c)), self.ctx, "bc size")
        out = np.zeros(c.value, np.int64)
        check(self.L.fsgh_get_bc_set(self.ctx, inst, s, _p(out)), self.ctx, "bc set")
        return out

    def close(self):
        if getattr(self, "ctx", None) and self.ctx.value:
            self.L.fsgh_destroy(self.ctx)
            self.ctx = C.c_void_p()

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

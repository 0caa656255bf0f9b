"""Problem definitions: the compiled-in ``input()`` of the reference as data.

The reference selects its problem by editing K/input.cpp:50-651 (materials,
fills, boundary sets, run controls) and passing a mesh file (K/driver.cpp:138).
Here a problem is a ``ProblemSpec``: plain data that both the HIP solver
(``fierro_amd.sgh``) and the CPU oracle consume, so both run the same inputs.

2D axisymmetric (RZ) problems have ``dims == 2``: quads in the (axial x,
radius y) half plane from ``rect_mesh`` (node (i,j) = i + (nx+1) j, element
nodes (i,j),(i+1,j),(i+1,j+1),(i,j+1), counter-clockwise as the reference's
2D meshes, K/mesh.h:366-373) or a 2D EnSight file (``read_ensight_2d``).

Mesh generator: ``box_mesh`` numbers nodes i-fastest, then j, then k, and
elements with nodes (i,j,k),(i+1,j,k),(i+1,j+1,k),(i,j+1,k) followed by the
k+1 layer -- the numbering of the reference's EnSight meshes (e.g.
Explicit-Lagrange/meshes/mesh_2x2x2.geo:91-100).  Coordinates are the
correctly rounded k*L/n, computed as (k*Lnum)/(n*Lden) in FP64, which is what
reading a mesh written with exact decimal coordinates gives.
"""
from __future__ import annotations

import dataclasses
from dataclasses import dataclass, field

import numpy as np

# enums of K/mesh.h:820-918
GLOBAL, BOX, CYLINDER, SPHERE = 0, 1, 2, 3
CARTESIAN, RADIAL, SPHERICAL = 0, 1, 2
X_PLANE, Y_PLANE, Z_PLANE = 0, 1, 2
FIXED, REFLECTED = 0, 1
EOS_IDEAL_GAS, EOS_USER = 0, 1      # the material models of K/state.h:140-170
STRENGTH_NONE, STRENGTH_HYPO, STRENGTH_HYPER = 0, 1, 2   # strength_tag, K/state.h:107-112

# K/driver.cpp:35-62 defaults
DEFAULT_RUN = dict(time_final=1.0e16, dt_start=1.0e-8, dt_max=1.0e-2, dt_min=1.0e-8,
                   dt_cfl=0.4, graphics_dt_ival=1.0e8, graphics_cyc_ival=50,
                   cycle_stop=1000000000, fuzz=1.0e-16, tiny=1.0e-12, small=1.0e-8)


def fill(volume=GLOBAL, mat_id=0, den=1.0, sie=0.0, velocity=CARTESIAN, u=0.0, v=0.0, w=0.0,
         speed=0.0, x1=0.0, x2=0.0, y1=0.0, y2=0.0, z1=0.0, z2=0.0, radius1=0.0, radius2=0.0):
    """One mat_fill_t (K/mesh.h:858-890)."""
    return dict(volume=volume, mat_id=mat_id, x1=x1, x2=x2, y1=y1, y2=y2, z1=z1, z2=z2,
                radius1=radius1, radius2=radius2, velocity=velocity, u=u, v=v, w=w,
                speed=speed, sie=sie, den=den)


@dataclass
class ProblemSpec:
    name: str
    mesh: dict                      # {"kind": "box", nx, ny, nz, lx, ly, lz, lden} or {"kind": "geo", "path"}
    materials: list                 # [(q1, q1ex, q2, q2ex)] per material
    state_vars: np.ndarray          # num_materials x num_state_vars (gamma, csmin, ...)
    mat_num_state_vars: list
    fills: list
    bcs: list                       # [(surface, value, hydro_bc)]
    run: dict = field(default_factory=lambda: dict(DEFAULT_RUN))
    num_state_vars: int = 6
    eos: list = None                # per material: EOS_IDEAL_GAS (default) or EOS_USER
    dims: int = 3                   # 2: axisymmetric RZ (mesh kinds "rect", "geo2")
    strength: list = None           # per material: STRENGTH_NONE (default), _HYPO or _HYPER
    strength_params: list = None    # per material: up to 8 parameters of its strength model

    def mesh_arrays(self, oracle_gen=None, oracle_read=None):
        """(coords N x dims, nodes_in_elem E x 2^dims).  The oracle passes its
        own generator/reader so that its inputs come from its own code."""
        m = self.mesh
        if m["kind"] == "box":
            gen = oracle_gen or box_mesh
            return gen(m["nx"], m["ny"], m["nz"], m["lx"], m["ly"], m["lz"], m["lden"])
        if m["kind"] == "geo":
            rd = oracle_read or read_ensight
            return rd(m["path"])
        if m["kind"] == "rect":
            gen = oracle_gen or rect_mesh
            return gen(m["nx"], m["ny"], m["lx"], m["ly"], m["lden"])
        if m["kind"] == "geo2":
            rd = oracle_read or read_ensight_2d
            return rd(m["path"])
        raise ValueError(m["kind"])

    def with_run(self, **kw):
        r = dict(self.run)
        r.update(kw)
        return dataclasses.replace(self, run=r)

    @property
    def num_elems(self):
        m = self.mesh
        if m["kind"] == "box":
            return m["nx"] * m["ny"] * m["nz"]
        return m["nx"] * m["ny"] if m["kind"] == "rect" else None


def box_mesh(nx, ny, nz, lx_num=1.0, ly_num=1.0, lz_num=1.0, l_den=1.0):
    """Structured hex box on [0,Lx]x[0,Ly]x[0,Lz], L = L_num/L_den."""
    nx1, ny1, nz1 = nx + 1, ny + 1, nz + 1
    k, j, i = np.meshgrid(np.arange(nz1), np.arange(ny1), np.arange(nx1), indexing="ij")
    coords = np.empty((nx1 * ny1 * nz1, 3))
    coords[:, 0] = (i.ravel().astype(np.float64) * lx_num) / (float(nx) * l_den)
    coords[:, 1] = (j.ravel().astype(np.float64) * ly_num) / (float(ny) * l_den)
    coords[:, 2] = (k.ravel().astype(np.float64) * lz_num) / (float(nz) * l_den)
    ke, je, ie = np.meshgrid(np.arange(nz), np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (ie + nx1 * (je + ny1 * ke)).ravel().astype(np.int64)
    conn = np.empty((nx * ny * nz, 8), np.int64)
    conn[:, 0] = n0
    conn[:, 1] = n0 + 1
    conn[:, 2] = n0 + 1 + nx1
    conn[:, 3] = n0 + nx1
    conn[:, 4:] = conn[:, :4] + nx1 * ny1
    return coords, conn


def rect_mesh(nx, ny, lx_num=1.0, ly_num=1.0, l_den=1.0):
    """Structured quads on [0,Lx]x[0,Ly] (x axial, y radius), L = L_num/L_den."""
    nx1, ny1 = nx + 1, ny + 1
    j, i = np.meshgrid(np.arange(ny1), np.arange(nx1), indexing="ij")
    coords = np.empty((nx1 * ny1, 2))
    coords[:, 0] = (i.ravel().astype(np.float64) * lx_num) / (float(nx) * l_den)
    coords[:, 1] = (j.ravel().astype(np.float64) * ly_num) / (float(ny) * l_den)
    je, ie = np.meshgrid(np.arange(ny), np.arange(nx), indexing="ij")
    n0 = (ie + nx1 * je).ravel().astype(np.int64)
    conn = np.empty((nx * ny, 4), np.int64)
    conn[:, 0] = n0
    conn[:, 1] = n0 + 1
    conn[:, 2] = n0 + 1 + nx1
    conn[:, 3] = n0 + nx1
    return coords, conn


def read_ensight_2d(path):
    """EnSight Gold ASCII 2D reader following K/read_mesh.cpp:12-147 with
    num_dims = 2: x and y blocks kept, the z block read and dropped, 4 node
    ids per element."""
    with open(path) as f:
        lines = f.read().split("\n")
    toks = " ".join(lines[8:]).split()
    n = int(toks[0])
    xyz = np.array(toks[1:1 + 3 * n], dtype=np.float64).reshape(3, n).T
    rest = toks[1 + 3 * n:]
    ne = int(rest[1])
    conn = np.array(rest[2:2 + 4 * ne], dtype=np.int64).reshape(ne, 4) - 1
    return np.ascontiguousarray(xyz[:, :2]), conn


def read_ensight(path):
    """EnSight Gold ASCII hexa8 reader following K/read_mesh.cpp:12-147."""
    with open(path) as f:
        lines = f.read().split("\n")
    pos = 8
    toks = " ".join(lines[pos:]).split()
    n = int(toks[0])
    xyz = np.array(toks[1:1 + 3 * n], dtype=np.float64).reshape(3, n).T.copy()
    rest = toks[1 + 3 * n:]
    assert rest[0] == "hexa8", rest[0]
    ne = int(rest[1])
    conn = np.array(rest[2:2 + 8 * ne], dtype=np.int64).reshape(ne, 8) - 1
    return xyz, conn


def _gas(gamma, csmin=1.0e-14, nmat=1):
    sv = np.zeros((nmat, 6))
    sv[:, 0] = gamma
    sv[:, 1] = csmin
    sv[:, 2] = 1.0
    return sv


def sedov(n=32, q2=1.333, dt_cfl=0.5, cycles=200, gamma=5.0 / 3.0):
    """Point-energy blast on [0,1.2]^3 with n^3 hexes: rho=1, u=0, e=1e-10
    except the origin element, e=0.25/V (IE0 = 0.25); symmetry (reflected)
    planes x=y=z=0; q1=1, q2 as given, q1ex=1, q2ex=0; RK2, CFL dt_cfl,
    dt_start 1e-5, a fixed number of cycles.

    This is the benchmark variant of the blast (the plain-mesh workloads at
    32^3 and 256^3 are quoted on it), not the reference's compiled-in
    Sedov3D input, which ``reference_sedov`` reproduces.  It differs from
    K/input.cpp:105-202 in four ways: the energy goes into the origin
    element only (sphere fill of radius h, e = 0.25/V) instead of the sphere
    of radius 1.2/8 with sie = 963.652344 (1.2/30)^3 / r^3; q2 = 1.333 instead
    of 1.3333; three reflected planes instead of six (the blast never reaches
    x, y, z = 1.2 within the cycles run); and time_final 1e16 with a cycle
    limit instead of time_final 1.0.  The material, EOS, q1 / q1ex / q2ex and
    the time-step controls are the reference's."""
    h = 1.2 / n
    vol0 = h * h * h
    spec = ProblemSpec(
        name=f"sedov{n}",
        mesh=dict(kind="box", nx=n, ny=n, nz=n, lx=12.0, ly=12.0, lz=12.0, lden=10.0),
        materials=[(1.0, 1.0, q2, 0.0)],
        state_vars=_gas(gamma),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=1.0e-10),
               fill(SPHERE, den=1.0, sie=0.25 / vol0, radius1=0.0, radius2=h)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED), (Z_PLANE, 0.0, REFLECTED)],
    )
    spec.run.update(time_final=1.0e16, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec



def blast(n, E0, gamma, rho0=1.0, cycles=150, q2=1.333, dt_cfl=0.5):
    """Point blast of the calibration sweep: the Sedov geometry of ``sedov``
    (n^3 hexes on [0,1.2]^3, reflected x=y=z=0) with density rho0, the
    energy E0 deposited in the origin element as e = E0 / (rho0 V), e=1e-10
    elsewhere, and the given gamma."""
    h = 1.2 / n
    spec = sedov(n, q2=q2, dt_cfl=dt_cfl, cycles=cycles, gamma=gamma)
    spec.name = f"blast{n}"
    spec.fills = [fill(GLOBAL, den=rho0, sie=1.0e-10),
                  fill(SPHERE, den=rho0, sie=E0 / (rho0 * h * h * h), radius1=0.0, radius2=h)]
    return spec


def blast_ensemble(count=64, seed=0, cycles=150):
    """The calibration sweep: ``count`` blasts with n ~ U{48..96},
    E0 ~ U(0.1, 1.0), gamma ~ U(1.3, 1.8), rho0 = 1, drawn per instance in
    that order from numpy's PCG64(seed); one more draw from the same stream
    is the target.  Returns (specs, params, target_spec, target_params),
    params = [(n, E0, gamma, rho0)]."""
    rng = np.random.Generator(np.random.PCG64(seed))
    params = []
    for _ in range(count + 1):
        n = int(rng.integers(48, 97))
        E0 = float(rng.uniform(0.1, 1.0))
        gamma = float(rng.uniform(1.3, 1.8))
        params.append((n, E0, gamma, 1.0))
    specs = [blast(n, E0, g, r, cycles=cycles) for n, E0, g, r in params]
    return specs[:count], params[:count], specs[count], params[count]


def sedov_user_region(n=8, cycles=100):
    """``sedov`` with a second material in the box [0.6,1.2] x [0,1.2] x
    [0,1.2] whose equation of state is the user model (material plugin hook,
    K/state.h:140-170, K/user_mat.cpp:14-40); same q coefficients."""
    spec = sedov(n, cycles=cycles)
    spec.name = f"sedov{n}_user"
    spec.materials = [spec.materials[0], spec.materials[0]]
    spec.state_vars = _gas(5.0 / 3.0, nmat=2)
    spec.mat_num_state_vars = [3, 3]
    spec.fills = spec.fills + [fill(BOX, mat_id=1, den=2.0, sie=1.0e-3,
                                    x1=0.6, x2=1.2, y1=0.0, y2=1.2, z1=0.0, z2=1.2)]
    spec.eos = [EOS_IDEAL_GAS, EOS_USER]
    return spec

PI = 3.141592653589793   # K/mesh.h:8


def _rz(spec):
    spec.dims = 2
    return spec


def sedov_rz(n=32, cycles=200, q2=1.3333, dt_cfl=0.4, dt_min=1.0e-8):
    """2D RZ Sedov blast (the reference's SedovRZ test input,
    tests-for-Explicit-Lagrange-Kokkos/input_for_SedovRZ_test.cpp:110-190):
    n x n quads on [0,1.2]^2 in (axial, radius), rho=1, e=1e-10, energy
    0.5 * 0.49339 / (rho V_s) in the sphere of radius 1.2/n, with
    V_s = PI (r2^3 - r1^3); reflected planes x=0, y=0 (the axis), x=1.2,
    y=1.2; gamma 5/3, q1=1, q2=1.3333, CFL 0.4 (K/driver.cpp:49).

    The deposit's specific energy grows as n^3 and its CFL step as n^-2.5:
    at n = 4096 the step (~2.6e-9) is below the reference's floor dt_min =
    1e-8 (K/driver.cpp:25-62), which then clamps the step up past the
    stable one and the run goes non-finite in its second cycle (the oracle
    and the GPU alike).  Meshes that fine need a lower dt_min."""
    import math
    r2 = 1.2 / n
    vol = PI * (math.pow(r2, 3) - math.pow(0.0, 3))
    spec = _rz(ProblemSpec(
        name=f"sedov_rz{n}",
        mesh=dict(kind="rect", nx=n, ny=n, lx=12.0, ly=12.0, lden=10.0),
        materials=[(1.0, 1.0, q2, 0.0)],
        state_vars=_gas(5.0 / 3.0),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=1.0e-10),
               fill(SPHERE, den=1.0, sie=(0.5 * 0.49339 / (1.0 * vol)), radius1=0.0,
                    radius2=r2)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED),
             (X_PLANE, 1.2, REFLECTED), (Y_PLANE, 1.2, REFLECTED)],
    ))
    spec.run.update(time_final=1.0, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=dt_min,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def noh_rz(n=32, cycles=100, dt_cfl=0.4):
    """2D RZ Noh implosion (input_for_NohRZ_test.cpp:110-160): n x n quads on
    [0,1]^2, rho=1, e=1e-9, radial inward velocity of speed 1 in the
    (axial, radius) plane, reflected x=0 and y=0, gamma 5/3, q2 = 1.3333."""
    spec = _rz(ProblemSpec(
        name=f"noh_rz{n}",
        mesh=dict(kind="rect", nx=n, ny=n, lx=1.0, ly=1.0, lden=1.0),
        materials=[(1.0, 1.0, 1.3333, 0.0)],
        state_vars=_gas(5.0 / 3.0),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=1e-9, velocity=RADIAL, speed=-1.0)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED)],
    ))
    spec.run.update(time_final=0.6, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def sod_rz(nx=100, ny=10, cycles=200, dt_cfl=0.4):
    """Sod along the axis in RZ coordinates (K/input.cpp:393-476, SodZ):
    [0,1] x [0,0.1], rho=1, e=2.5 (p=1) and, in x >= 0.5, rho=0.125, e=2.5
    (p=0.125), gamma 1.4; reflected x=0, y=0, x=1, y=0.1."""
    spec = _rz(ProblemSpec(
        name=f"sod_rz{nx}x{ny}",
        mesh=dict(kind="rect", nx=nx, ny=ny, lx=1.0, ly=0.1, lden=1.0),
        materials=[(1.0, 1.0, 1.3333, 0.0)],
        state_vars=_gas(1.4),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=2.5),
               fill(BOX, den=0.125, sie=2.5, x1=0.5, x2=1.0, y1=0.0, y2=1.0, z1=0.0, z2=1.0)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED),
             (X_PLANE, 1.0, REFLECTED), (Y_PLANE, 0.1, REFLECTED)],
    ))
    spec.run.update(time_final=0.2, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def reference_sedov(mesh_path=None, n=None, cycles=100000):
    """The reference's default compiled-in problem, Sedov3D (K/input.cpp:69-202
    with the K/driver.cpp:35-62 defaults): 6 reflected planes of the
    [0,1.2]^3 box, sphere fill of radius 1.2/8 with
    sie = 963.652344 (1.2/30)^3 / (1.2/8)^3, q1 = q1ex = 1, q2 = 1.3333,
    q2ex = 0, gamma 5/3, csmin 1e-14, CFL 0.4, dt_start 1e-5, time_final 1.0,
    graphics every 0.25, cycle_stop 100000.  The mesh is a .geo file
    (``mesh_path``, e.g. the reference's mesh_Sedov_32.geo) or the generated
    n^3 box with the same numbering (``n``); ``cycles`` lowers cycle_stop."""
    r2 = 1.2 / 8.0
    sie = (963.652344 * (1.2 / 30.0) ** 3) / (r2 ** 3)
    if mesh_path is not None:
        mesh = dict(kind="geo", path=mesh_path)
    else:
        mesh = dict(kind="box", nx=n, ny=n, nz=n, lx=12.0, ly=12.0, lz=12.0, lden=10.0)
    spec = ProblemSpec(
        name="reference_sedov3d" if n is None else f"reference_sedov3d_{n}",
        mesh=mesh,
        materials=[(1.0, 1.0, 1.3333, 0.0)],
        state_vars=_gas(5.0 / 3.0),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=1.0e-10),
               fill(SPHERE, den=1.0, sie=sie, radius1=0.0, radius2=r2)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED), (Z_PLANE, 0.0, REFLECTED),
             (X_PLANE, 1.2, REFLECTED), (Y_PLANE, 1.2, REFLECTED), (Z_PLANE, 1.2, REFLECTED)],
    )
    spec.run.update(time_final=1.0, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    graphics_dt_ival=0.25, graphics_cyc_ival=1000000, cycle_stop=cycles)
    return spec


def sod_tube(nx, rho_r, p_r, gamma, cycles=500, ny=4, nz=4, dt_cfl=0.5):
    """One Sod-type shock tube: nx x ny x nz cubic hexes, h = 1/nx, on
    [0,1]x[0,ny h]x[0,nz h]; left rho=1, p=1, right (rho_r, p_r); interface at
    x = 0.5 by element centroid; reflected walls on all six faces."""
    e_l = 1.0 / ((gamma - 1.0) * 1.0)
    e_r = p_r / ((gamma - 1.0) * rho_r)
    spec = ProblemSpec(
        name=f"sod{nx}",
        mesh=dict(kind="box", nx=nx, ny=ny, nz=nz, lx=float(nx), ly=float(ny), lz=float(nz),
                  lden=float(nx)),
        materials=[(1.0, 1.0, 1.333, 0.0)],
        state_vars=_gas(gamma),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=e_l),
               fill(BOX, den=rho_r, sie=e_r, x1=0.5, x2=2.0, y1=-1.0, y2=1.0e9, z1=-1.0,
                    z2=1.0e9)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (X_PLANE, 1.0, REFLECTED),
             (Y_PLANE, 0.0, REFLECTED), (Y_PLANE, ny / nx, REFLECTED),
             (Z_PLANE, 0.0, REFLECTED), (Z_PLANE, nz / nx, REFLECTED)],
    )
    spec.run.update(time_final=1.0e16, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def sod_ensemble(count=512, seed=0, cycles=500):
    """The batched ensemble: `count` tubes with nx ~ U{64..1024},
    rho_R, p_R ~ U(0.05, 0.5), gamma ~ U(1.3, 1.7), drawn per instance in that
    order from numpy's PCG64(seed)."""
    rng = np.random.Generator(np.random.PCG64(seed))
    out = []
    for _ in range(count):
        nx = int(rng.integers(64, 1025))
        rho_r = float(rng.uniform(0.05, 0.5))
        p_r = float(rng.uniform(0.05, 0.5))
        gamma = float(rng.uniform(1.3, 1.7))
        out.append(sod_tube(nx, rho_r, p_r, gamma, cycles=cycles))
    return out


TG_VORTEX = 5   # velocity fill kind (K/mesh.h:851)


def taylor_green(n=16, nz=2, cycles=100, gamma=5.0 / 3.0, statev4=None, dims=3, dt_cfl=0.4):
    """Taylor-Green vortex on the unit square (K/setup.cpp:326-350): velocity
    (sin pi x cos pi y, -cos pi x sin pi y, 0), pressure
    0.25 (cos 2 pi x + cos 2 pi y) + 1 by element centroid and energy
    p / (rho (statev[4] - 1)); rho = 1.  The material carries 6 state
    variables with gamma in slot 0 and ``statev4`` in slot 4 (default gamma,
    the physical choice; 0 reproduces a material without that slot, for which
    the reference's energy comes out negative).  3D: n x n x nz hexes on
    [0,1] x [0,1] x [0, nz/n]; 2D: n x n quads.  Reflected on every side
    (the vortex's normal velocity vanishes on the walls)."""
    sv = _gas(gamma)
    sv[0, 4] = gamma if statev4 is None else statev4
    if dims == 2:
        mesh = dict(kind="rect", nx=n, ny=n, lx=1.0, ly=1.0, lden=1.0)
        bcs = [(X_PLANE, 0.0, REFLECTED), (X_PLANE, 1.0, REFLECTED),
               (Y_PLANE, 0.0, REFLECTED), (Y_PLANE, 1.0, REFLECTED)]
    else:
        mesh = dict(kind="box", nx=n, ny=n, nz=nz, lx=float(n), ly=float(n), lz=float(nz),
                    lden=float(n))
        bcs = [(X_PLANE, 0.0, REFLECTED), (X_PLANE, 1.0, REFLECTED),
               (Y_PLANE, 0.0, REFLECTED), (Y_PLANE, 1.0, REFLECTED),
               (Z_PLANE, 0.0, REFLECTED), (Z_PLANE, nz / n, REFLECTED)]
    spec = ProblemSpec(
        name=f"taylor_green{n}" + ("_rz" if dims == 2 else ""),
        mesh=mesh,
        materials=[(1.0, 1.0, 1.3333, 0.0)],
        state_vars=sv,
        mat_num_state_vars=[6],
        fills=[fill(GLOBAL, den=1.0, sie=1.0, velocity=TG_VORTEX)],
        bcs=bcs,
        dims=dims,
    )
    spec.run.update(time_final=1.0e16, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def noh(n=16, cycles=100, dt_cfl=0.4):
    """Noh implosion (K/input.cpp:297-340): rho=1, e=1e-9, spherical inward
    velocity of speed 1 on [0,1]^3 with n^3 hexes, reflected x=y=z=0 planes,
    q2 = 1.3333, gamma 5/3, csmin 1e-14."""
    spec = ProblemSpec(
        name=f"noh{n}",
        mesh=dict(kind="box", nx=n, ny=n, nz=n, lx=1.0, ly=1.0, lz=1.0, lden=1.0),
        materials=[(1.0, 1.0, 1.3333, 0.0)],
        state_vars=_gas(5.0 / 3.0),
        mat_num_state_vars=[6],
        fills=[fill(GLOBAL, den=1.0, sie=1e-9, velocity=SPHERICAL, speed=-1.0)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED), (Z_PLANE, 0.0, REFLECTED)],
    )
    spec.run.update(time_final=0.6, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def anvil(nx=16, ny=16, nz=96, cycles=400, speed=0.2, dt_cfl=0.5):
    """Taylor-anvil-style impact: a 0.32 x 0.32 x 2.0 bar of nx x ny x nz
    hexes, ideal gas gamma=1.4, rho=1, e=1e-3, uniform velocity (0,0,-speed)
    into a rigid reflected wall at z=0, with symmetry (reflected) planes x=0
    and y=0 as in the reference's anvil input (K/input.cpp:627-641)."""
    spec = ProblemSpec(
        name=f"anvil{nx}x{ny}x{nz}",
        mesh=dict(kind="box", nx=nx, ny=ny, nz=nz, lx=0.32, ly=0.32, lz=2.0, lden=1.0),
        materials=[(1.0, 1.0, 1.333, 0.0)],
        state_vars=_gas(1.4),
        mat_num_state_vars=[3],
        fills=[fill(GLOBAL, den=1.0, sie=1.0e-3, velocity=CARTESIAN, w=-speed)],
        bcs=[(X_PLANE, 0.0, REFLECTED), (Y_PLANE, 0.0, REFLECTED), (Z_PLANE, 0.0, REFLECTED)],
    )
    spec.run.update(time_final=1.0e16, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=dt_cfl, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def anvil_objective(spec, coords0):
    """Seed of the anvil objective J = final bar length + base lateral spread,
    a linear functional of the final node coordinates: the mean z of the top
    face (z0 = lz) plus the mean x of the base nodes on the outer face
    x0 = lx and the mean y of the base nodes on y0 = ly (the bar is a
    quarter, with symmetry planes x = 0 and y = 0).  Returns seed [N][3]."""
    m = spec.mesh
    x0 = np.asarray(coords0, np.float64)
    tol = 1e-12
    seed = np.zeros_like(x0)
    top = np.abs(x0[:, 2] - m["lz"]) < tol
    base = np.abs(x0[:, 2]) < tol
    bx = base & (np.abs(x0[:, 0] - m["lx"]) < tol)
    by = base & (np.abs(x0[:, 1] - m["ly"]) < tol)
    seed[top, 2] += 1.0 / top.sum()
    seed[bx, 0] += 1.0 / bx.sum()
    seed[by, 1] += 1.0 / by.sum()
    return seed

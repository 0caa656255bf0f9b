"""Set-up paths no benchmark problem takes, on the GPU against the oracle:
two materials (different viscosity coefficients and gamma) painted by four
overlapping fills -- global, cylinder, sphere and box, in that order, the
last match winning (K/setup.cpp:117-232) -- with cartesian, radial and
spherical initial velocities (K/setup.cpp:247-330); a fixed wall
(hydro_bc 0, every velocity component held at zero, K/boundary.cpp:36-44)
beside reflected planes; and cylinder and sphere boundary surfaces
(K/setup.cpp:537-556), whose sets on a box mesh are empty or hold only the
patches at that radius.  Set-up fields and boundary sets must be bitwise the
oracle's; 60 fused cycles meet the usual bars.  3D and the 2D RZ path."""
import numpy as np
import pytest

from fierro_amd import problems as P

pytestmark = pytest.mark.gpu

STATE = ["node_coords", "node_vel", "node_mass", "elem_den", "elem_pres", "elem_sspd",
         "elem_sie", "elem_vol", "elem_mass", "elem_mat_id"]


def rel(a, b):
    d = float(np.max(np.abs(np.asarray(a, float) - np.asarray(b, float))))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


def fields(pr):
    return dict(node_coords=pr.node_coords, node_vel=pr.node_vel, node_mass=pr.node_mass,
                elem_den=pr.elem_den, elem_pres=pr.elem_pres, elem_sspd=pr.elem_sspd,
                elem_sie=pr.elem_sie, elem_vol=pr.elem_vol, elem_mass=pr.elem_mass,
                elem_mat_id=pr.elem_mat_id)


@pytest.fixture(scope="module")
def Solver():
    from fierro_amd.sgh import SGHSolver
    return SGHSolver


def two_materials():
    sv = np.zeros((2, 6))
    sv[0, :3] = (1.4, 1.0e-14, 1.0)
    sv[1, :3] = (5.0 / 3.0, 1.0e-12, 1.0)
    return sv


def layered_3d(n=10, cycles=60):
    spec = P.ProblemSpec(
        name=f"layered{n}",
        mesh=dict(kind="box", nx=n, ny=n, nz=n, lx=1.0, ly=1.0, lz=1.0, lden=1.0),
        materials=[(1.0, 1.0, 1.3333, 0.0), (0.5, 0.8, 1.0, 0.2)],
        state_vars=two_materials(), mat_num_state_vars=[3, 3],
        fills=[P.fill(P.GLOBAL, mat_id=0, den=1.0, sie=1.0e-3, u=0.1, v=-0.05, w=0.02),
               P.fill(P.CYLINDER, mat_id=1, den=2.0, sie=0.5, velocity=P.RADIAL, speed=0.3,
                      radius1=0.2, radius2=0.55),
               P.fill(P.SPHERE, mat_id=0, den=1.0, sie=2.0, velocity=P.SPHERICAL, speed=-0.2,
                      radius1=0.0, radius2=0.25),
               P.fill(P.BOX, mat_id=1, den=0.5, sie=0.1, x1=0.6, x2=1.0, y1=0.0, y2=0.3,
                      z1=0.0, z2=1.0)],
        bcs=[(P.Z_PLANE, 0.0, P.FIXED), (P.X_PLANE, 0.0, P.REFLECTED),
             (P.Y_PLANE, 0.0, P.REFLECTED), (3, 0.5, P.FIXED), (4, 1.0, P.FIXED)])
    spec.run.update(time_final=1.0, dt_start=1.0e-5, dt_max=1.0e-2, dt_min=1.0e-8,
                    dt_cfl=0.4, graphics_dt_ival=1.0e16, graphics_cyc_ival=1000000000,
                    cycle_stop=cycles)
    return spec


def layered_rz(n=12, cycles=60):
    spec = P.sedov_rz(n, cycles=cycles)
    spec.materials = [(1.0, 1.0, 1.3333, 0.0), (0.5, 0.8, 1.0, 0.2)]
    spec.state_vars = two_materials()
    spec.mat_num_state_vars = [3, 3]
    spec.fills = [P.fill(P.GLOBAL, mat_id=0, den=1.0, sie=1.0e-3, u=0.1, v=0.0),
                  P.fill(P.CYLINDER, mat_id=1, den=2.0, sie=0.5, velocity=P.RADIAL, speed=0.3,
                         radius1=0.3, radius2=0.7),
                  P.fill(P.BOX, mat_id=1, den=0.5, sie=0.1, x1=0.8, x2=1.2, y1=0.0, y2=0.4),
                  P.fill(P.SPHERE, mat_id=0, den=1.0, sie=2.0, radius1=0.0, radius2=0.25)]
    spec.bcs = [(P.X_PLANE, 0.0, P.FIXED), (P.Y_PLANE, 0.0, P.REFLECTED),
                (P.X_PLANE, 1.2, P.REFLECTED), (P.Y_PLANE, 1.2, P.REFLECTED),
                (3, 1.2, P.FIXED)]
    return spec


@pytest.mark.parametrize("dims", [3, 2])
def test_setup_paths_match_oracle(oracle, Solver, dims):
    spec = layered_3d() if dims == 3 else layered_rz()
    print(f"[setup paths] dims={dims}", flush=True)
    pr = (oracle.OracleProblem if dims == 3 else oracle.OracleRZ)(spec)
    g = Solver(spec, renumber=True)
    for k, v in fields(pr).items():
        got = g.download(k)
        if k in ("node_vel", "elem_sie"):
            got, v = got[1], v[1]
        assert np.array_equal(got, v), k
    mats = g.download("elem_mat_id")
    assert set(np.unique(mats)) == {0, 1}
    sizes = np.diff(pr.bset_start)
    assert g.bc_set_sizes() == list(sizes)
    for i in range(len(spec.bcs)):
        assert np.array_equal(g.bc_set(0, i), pr.bset_nodes[pr.bset_start[i]:pr.bset_start[i + 1]])
    assert sizes[0] > 0
    # the fixed wall holds every component of its nodes at zero from the start
    wall = g.bc_set(0, 0)
    assert np.all(g.download("node_vel")[1][wall] == 0.0)
    ro = pr.solve()
    rg = g.sgh_solve(use_graph=True)[0]
    assert rg["cycles"] == ro["cycles"] == 60
    assert abs(rg["time_value"] - ro["time_value"]) <= 1e-12 * ro["time_value"]
    assert abs(rg["te1"] - ro["te1"]) <= 1e-12 * abs(ro["te1"])
    for k in STATE:
        assert rel(g.download(k), fields(pr)[k]) <= 1e-9, k
    assert np.all(g.download("node_vel")[1][wall] == 0.0)

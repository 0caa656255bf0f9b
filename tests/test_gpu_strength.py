"""GPU parity of the strength-model plumbing against the oracle.

A non-zero Cauchy stress is uploaded (the shipped hypo / hyper models are
empty, as the reference's are, so the stress stays what it was set to) and
every operator of one cycle, the fused loop and the velocity gradient are
compared with the oracle's stress path: tau = sigma - p I in the corner
forces (with the hoop terms in RZ), p = -tr(sigma)/3 in the user EOS, and the
stress carried through rk_init.  Same bar as the strength-free path.
"""
import numpy as np
import pytest

from fierro_amd import problems as P

pytestmark = pytest.mark.gpu

TOL_STAGE = 1e-12
STATE = ["node_coords", "node_vel", "node_mass", "elem_den", "elem_pres", "elem_sspd",
         "elem_sie", "elem_vol", "elem_div", "elem_mass"]


def rel(a, b):
    d = float(np.max(np.abs(np.asarray(a, float) - np.asarray(b, float))))
    s = float(np.max(np.abs(b)))
    return d / s if s > 0 else d


def fields(pr):
    return dict(node_coords=pr.node_coords, node_vel=pr.node_vel, node_mass=pr.node_mass,
                elem_den=pr.elem_den, elem_pres=pr.elem_pres, elem_sspd=pr.elem_sspd,
                elem_sie=pr.elem_sie, elem_vol=pr.elem_vol, elem_div=pr.elem_div,
                elem_mass=pr.elem_mass, corner_force=pr.corner_force)


def random_stress(pres, frac, seed):
    """A symmetric stress of each element's own pressure scale (a stress far
    above the local pressure would tangle the cold cells of a blast)."""
    rng = np.random.default_rng(seed)
    a = rng.standard_normal((pres.size, 3, 3)) * (frac * np.abs(pres))[:, None, None]
    s = 0.5 * (a + np.swapaxes(a, 1, 2))
    return np.stack([s, s])


def with_strength(spec, kinds):
    spec.strength = list(kinds)
    return spec


CASES = {
    "sedov8_hypo": lambda: with_strength(P.sedov(8, cycles=60), [P.STRENGTH_HYPO]),
    "anvil4x4x12_hyper": lambda: with_strength(P.anvil(4, 4, 12, cycles=60), [P.STRENGTH_HYPER]),
    "sedov8_user_hypo": lambda: with_strength(P.sedov_user_region(8, cycles=60),
                                              [P.STRENGTH_NONE, P.STRENGTH_HYPO]),
    "sedov_rz8_hypo": lambda: with_strength(P.sedov_rz(8, cycles=60), [P.STRENGTH_HYPO]),
    "sod_rz40x4_hypo": lambda: with_strength(P.sod_rz(40, 4, cycles=60), [P.STRENGTH_HYPO]),
}


@pytest.fixture(scope="module")
def Solver():
    from fierro_amd.sgh import SGHSolver
    return SGHSolver


def make_oracle(oracle, spec):
    return oracle.OracleRZ(spec) if spec.dims == 2 else oracle.OracleProblem(spec)


def push(g, pr):
    for k, v in fields(pr).items():
        g.upload(k, v)
    g.upload("elem_gamma", pr.elem_statev[:, 0])
    g.upload("elem_csmin", pr.elem_statev[:, 1])
    g.upload("elem_mat_id", pr.elem_mat_id)
    g.upload("elem_stress", pr.elem_stress)


@pytest.mark.parametrize("renumber", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_stressed_stage_operators(oracle, Solver, case, renumber):
    spec = CASES[case]()
    pr = make_oracle(oracle, spec)
    pr.solve(cycle_stop=6)
    pr.elem_stress[:] = random_stress(pr.elem_pres, 0.2, 3)
    g = Solver(spec, renumber=renumber)
    push(g, pr)
    assert np.array_equal(g.download("elem_stress"), pr.elem_stress)
    r = spec.run
    rz = spec.dims == 2
    if rz:
        pr.extensive_mass()
        g.rz_extensive_mass()
    dt_o = pr.get_timestep(0.0, r["graphics_dt_ival"], r["dt_start"] * 10)
    dt_g = g.get_timestep(0.0, r["graphics_dt_ival"], r["time_final"], r["dt_max"], r["dt_min"],
                          r["dt_cfl"], r["dt_start"] * 10, r["fuzz"])
    assert dt_g == dt_o
    pr.elem_stress[0] = 0.0
    g.upload("elem_stress", pr.elem_stress)
    pr.rk_init()
    g.rk_init()
    assert np.array_equal(g.download("elem_stress"), pr.elem_stress)
    vg = oracle.velgrad_rz(pr) if rz else oracle.velgrad_3d(pr)
    assert rel(g.get_velgrad(), vg) <= TOL_STAGE
    for stage in range(2):
        alpha = 1.0 / (2.0 - stage)
        steps = [("get_divergence", "elem_div"), ("get_force_sgh", "corner_force"),
                 ("update_velocity_sgh", None), ("boundary_velocity", "node_vel"),
                 ("update_energy_sgh", "elem_sie"), ("update_position_sgh", "node_coords"),
                 ("get_vol", "elem_vol"), ("update_state", "elem_pres")]
        if rz:
            steps.append(("node_mass", "node_mass"))
        for name, field in steps:
            if name == "get_force_sgh":
                pr.get_force_sgh(dt_o, alpha)
                g.get_force_sgh(r["fuzz"], r["small"], dt_o, alpha)
            elif name == "update_state":
                pr.update_state(dt_o, alpha)
                g.update_state(dt_o, alpha)
            elif name == "node_mass":
                pr.update_node_mass(r["tiny"])
                g.rz_update_node_mass(r["tiny"])
            elif name in ("update_velocity_sgh", "update_energy_sgh", "update_position_sgh"):
                getattr(pr, name)(alpha, dt_o)
                getattr(g, name)(alpha, dt_o)
            else:
                getattr(pr, name)()
                getattr(g, name)()
            if field is not None:
                got = g.download(field)
                assert rel(got, fields(pr)[field]) <= TOL_STAGE, (stage, name)
    for k in STATE:
        assert rel(g.download(k), fields(pr)[k]) <= TOL_STAGE, k
    assert np.array_equal(g.download("elem_stress"), pr.elem_stress)


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_stressed_fused_cycles(oracle, Solver, case, use_graph):
    spec = CASES[case]()
    pr = make_oracle(oracle, spec)
    g = Solver(spec, renumber=True)
    pr.elem_stress[:] = random_stress(pr.elem_pres, 0.1, 9)
    g.upload("elem_stress", pr.elem_stress)
    ro = pr.solve()
    rg = g.sgh_solve(use_graph=use_graph)[0]
    assert rg["cycles"] == ro["cycles"] == 60
    assert abs(rg["time_value"] - ro["time_value"]) <= 1e-14 * ro["time_value"]
    assert abs(rg["te1"] - ro["te1"]) <= 1e-12 * abs(ro["te1"])
    den = g.download("elem_den")
    l1 = float(np.sum(np.abs(den - pr.elem_den)) / np.sum(np.abs(pr.elem_den)))
    assert l1 <= 1e-9, l1
    for k in STATE:
        assert rel(g.download(k), fields(pr)[k]) <= 1e-9, k
    assert np.array_equal(g.download("elem_stress"), pr.elem_stress)


def test_zero_stress_matches_strength_free_run(Solver):
    """With the stress left at zero a hypo material runs the same cycle as
    the strength-free build, bit for bit."""
    a = Solver(with_strength(P.sedov(10, cycles=50), [P.STRENGTH_HYPO]), renumber=True)
    b = Solver(P.sedov(10, cycles=50), renumber=True)
    a.sgh_solve(use_graph=True)
    b.sgh_solve(use_graph=True)
    for k in STATE:
        assert np.array_equal(a.download(k), b.download(k)), k


def test_user_eos_reads_stress(Solver):
    spec = with_strength(P.sedov_user_region(6, cycles=1), [P.STRENGTH_NONE, P.STRENGTH_HYPO])
    g = Solver(spec)
    E = g.E
    sig = np.array([[1.0, 0.2, 0.0], [0.2, -2.0, 0.1], [0.0, 0.1, 4.0]])
    st = np.zeros((2, E, 3, 3))
    st[1] = sig
    g.upload("elem_stress", st)
    g.update_state(1e-3, 1.0)
    mat = g.download("elem_mat_id")
    p = g.download("elem_pres")
    assert np.all(p[mat == 1] == (0.0 - 1.0 - (-2.0) - 4.0) * (1.0 / 3.0))
    assert np.all(g.download("elem_sspd")[mat == 1] == 1.0e-8)


def test_strength_params_and_errors(Solver):
    from fierro_amd import _abi
    from fierro_amd._abi import FsghError
    g = Solver(with_strength(P.sedov(4, cycles=1), [P.STRENGTH_HYPO]))
    g.set_strength_params(0, [1.0, 2.0, 3.0])
    with pytest.raises(FsghError):
        g.set_strength_params(0, np.zeros(_abi.STRENGTH_PARAMS + 1))
    with pytest.raises(FsghError):
        g.set_strength_params(3, [1.0])
    # the launch-variant tuner and the adjoint cover the strength-free cycle only
    with pytest.raises(FsghError):
        g.autotune()
    with pytest.raises(FsghError):
        g.forward(2)
    assert g.L.fsgh_set_material_models(g.ctx, 0, _abi.EOS_IDEAL_GAS, 3) == -1
    rz = Solver(P.sedov_rz(4, cycles=1))
    assert rz.L.fsgh_set_material_models(rz.ctx, 0, _abi.EOS_IDEAL_GAS,
                                         _abi.STRENGTH_HYPER) == -6
    assert rz.L.fsgh_set_material_models(rz.ctx, 0, _abi.EOS_IDEAL_GAS,
                                         _abi.STRENGTH_HYPO) == 0

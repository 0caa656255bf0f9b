"""A strength model that acts, run through every hook call site against the
oracle.  The library's test variant (libfsgh_strength_test.so) and the
oracle's (liboracle_sgh_strength.so) replace the empty user strength body by
the same hypoelastic update, sigma = sigma_n + alpha dt 2 G dev((L + L^T)/2),
so the stress evolves: the HYPO call site after the corner forces (3D and
RZ, K/force_sgh.cpp:369-391, :774-801) and the HYPER call site before the
EOS (K/properties.cpp:46-95), each operator of one cycle and 60 fused
cycles (with and without the graph).  Same bars as the strength-free path;
the stress itself within the same bars."""
import numpy as np
import pytest

from fierro_amd import _abi, problems as P

from test_gpu_strength import fields, make_oracle, push, random_stress, rel, STATE, TOL_STAGE

pytestmark = pytest.mark.gpu


def model(spec, kinds, G):
    spec.strength = list(kinds)
    spec.strength_params = [[G] if k != P.STRENGTH_NONE else [] for k in kinds]
    return spec


CASES = {
    "sedov8_hypo": lambda: model(P.sedov(8, cycles=60), [P.STRENGTH_HYPO], 0.05),
    "anvil4x4x12_hyper": lambda: model(P.anvil(4, 4, 12, cycles=60), [P.STRENGTH_HYPER], 0.02),
    "sedov8_user_hypo": lambda: model(P.sedov_user_region(8, cycles=60),
                                      [P.STRENGTH_NONE, P.STRENGTH_HYPO], 0.05),
    "sedov_rz8_hypo": lambda: model(P.sedov_rz(8, cycles=60), [P.STRENGTH_HYPO], 0.05),
    "sod_rz40x4_hypo": lambda: model(P.sod_rz(40, 4, cycles=60), [P.STRENGTH_HYPO], 0.05),
}


@pytest.fixture(scope="module")
def Solver():
    from fierro_amd.sgh import SGHSolver

    def make(spec, **kw):
        return SGHSolver(spec, library=_abi.STRENGTH_TEST_LIB_PATH, **kw)
    return make


def oracle_model(oracle, spec):
    return (oracle.OracleRZ(spec, strength_model=True) if spec.dims == 2 else
            oracle.OracleProblem(spec, strength_model=True))


@pytest.mark.parametrize("case", sorted(CASES))
def test_model_stage_operators(oracle, Solver, case):
    spec = CASES[case]()
    pr = oracle_model(oracle, spec)
    pr.solve(cycle_stop=6)
    pr.elem_stress[:] = random_stress(pr.elem_pres, 0.2, 3)
    g = Solver(spec, renumber=True)
    push(g, pr)
    r = spec.run
    rz = spec.dims == 2
    if rz:
        pr.extensive_mass()
        g.rz_extensive_mass()
    dt = pr.get_timestep(0.0, r["graphics_dt_ival"], r["dt_start"] * 10)
    pr.rk_init()
    g.rk_init()
    s0 = pr.elem_stress[0].copy()
    for stage in range(2):
        alpha = 1.0 / (2.0 - stage)
        pr.get_divergence()
        g.get_divergence()
        pr.get_force_sgh(dt, alpha)
        g.get_force_sgh(r["fuzz"], r["small"], dt, alpha)
        assert rel(g.download("corner_force"), pr.corner_force) <= TOL_STAGE, (stage, "force")
        assert rel(g.download("elem_stress"), pr.elem_stress) <= TOL_STAGE, (stage, "stress")
        for name in ("update_velocity_sgh", "boundary_velocity", "update_energy_sgh",
                     "update_position_sgh", "get_vol"):
            args = (alpha, dt) if name.startswith("update") else ()
            getattr(pr, name)(*args)
            getattr(g, name)(*args)
        pr.update_state(dt, alpha)
        g.update_state(dt, alpha)
        if rz:
            pr.update_node_mass(r["tiny"])
            g.rz_update_node_mass(r["tiny"])
        assert rel(g.download("elem_stress"), pr.elem_stress) <= TOL_STAGE, (stage, "state")
        for k in STATE:
            assert rel(g.download(k), fields(pr)[k]) <= TOL_STAGE, (stage, k)
    # the model acted on the stage's stress (its size over one early cycle is
    # small; the fused runs below show it over 60)
    assert np.abs(pr.elem_stress[1] - s0).max() > 0.0


@pytest.mark.parametrize("use_graph", [False, True])
@pytest.mark.parametrize("case", sorted(CASES))
def test_model_fused_cycles(oracle, Solver, case, use_graph):
    spec = CASES[case]()
    pr = oracle_model(oracle, spec)
    g = Solver(spec, renumber=True)
    pr.elem_stress[:] = random_stress(pr.elem_pres, 0.1, 9)
    g.upload("elem_stress", pr.elem_stress)
    s_start = pr.elem_stress[1].copy()
    ro = pr.solve()
    rg = g.sgh_solve(use_graph=use_graph)[0]
    assert rg["cycles"] == ro["cycles"] == 60
    # the stress evolved (0.6 % to 1000 % of its starting scale; the
    # user-EOS case's hypo material sits in the still gas outside the blast)
    moved = np.abs(pr.elem_stress[1] - s_start).max() / np.abs(s_start).max()
    assert moved > (1e-12 if case == "sedov8_user_hypo" else 1e-3), moved
    assert abs(rg["te1"] - ro["te1"]) <= 1e-12 * abs(ro["te1"])
    den = g.download("elem_den")
    l1 = float(np.sum(np.abs(den - pr.elem_den)) / np.sum(np.abs(pr.elem_den)))
    assert l1 <= 1e-9, l1
    for k in STATE:
        assert rel(g.download(k), fields(pr)[k]) <= 1e-9, k
    assert rel(g.download("elem_stress"), pr.elem_stress) <= 1e-9

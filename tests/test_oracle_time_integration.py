"""Pins of the oracle's time integration (sgh_solve, K/sgh_solve.cpp:349-353,
K/time_integration.cpp:9-154) by what the mathematics fixes, not by a
restatement:

* temporal order: with the time steps frozen (dt_replay) the two-stage cycle
  with rk_alpha = 1/(2 - stage) is a second-order integrator, so halving dt
  divides the error at a fixed final time by 4; a one-stage cycle, a cycle
  whose stages all take full steps or one with wrong stage weights converges
  at first order or not at all.  The planted-mistake builds below show that
  these tests reject such integrators.
* the Sod shock tube against the exact Riemann solution (tests/riemann.py,
  Toro's method), L1 density error falling at first order with resolution.
* the output cadence (K/sgh_solve.cpp:594-628) on a run whose every time step
  is exactly 2^-10, so the output cycles are fixed by hand.
"""
import os
import subprocess

import numpy as np
import pytest

from fierro_amd import problems as P

import riemann as R

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE_DIR = os.path.join(os.path.dirname(HERE), "oracle")


def acoustic_spec(n=6, q=(0.0, 0.0, 0.0, 0.0)):
    """Closed box [0,1]^3 of ideal gas (gamma 1.4, rho 1, e 2.5), reflected on
    all six faces; the smooth initial fields are set by ``perturb``."""
    return P.ProblemSpec(
        name=f"acoustic{n}",
        mesh=dict(kind="box", nx=n, ny=n, nz=n, lx=1.0, ly=1.0, lz=1.0, lden=1.0),
        materials=[q], state_vars=P._gas(1.4), mat_num_state_vars=[3],
        fills=[P.fill(P.GLOBAL, den=1.0, sie=2.5)],
        bcs=[(P.X_PLANE, 0.0, P.REFLECTED), (P.X_PLANE, 1.0, P.REFLECTED),
             (P.Y_PLANE, 0.0, P.REFLECTED), (P.Y_PLANE, 1.0, P.REFLECTED),
             (P.Z_PLANE, 0.0, P.REFLECTED), (P.Z_PLANE, 1.0, P.REFLECTED)])


def perturb(pr, amp=0.05):
    """Smooth velocity field with zero normal component on every wall and a
    smooth energy field; pressure and sound speed follow from the EOS."""
    x = pr.node_coords[1]
    v = np.zeros_like(x)
    v[:, 0] = amp * np.sin(np.pi * x[:, 0]) * (1 + 0.3 * x[:, 1])
    v[:, 1] = amp * np.sin(np.pi * x[:, 1]) * (1 - 0.2 * x[:, 2])
    v[:, 2] = amp * np.sin(np.pi * x[:, 2]) * (1 + 0.4 * x[:, 0])
    pr.node_vel[0] = v
    pr.node_vel[1] = v
    xc = x[pr.conn].mean(axis=1)
    e = 2.5 * (1 + 0.1 * np.cos(np.pi * xc[:, 0]) * np.cos(np.pi * xc[:, 1])
               + 0.05 * np.cos(np.pi * xc[:, 2]))
    pr.elem_sie[0] = e
    pr.elem_sie[1] = e
    pr.update_state()


def run_fixed_dt(O, spec, dt, T, stages=2, fast=False):
    pr = O.OracleProblem(spec, fast=fast)
    perturb(pr)
    K = int(round(T / dt))
    r = pr.solve(cycle_stop=K, dt_replay=np.full(K, dt), rk_num_stages=stages)
    assert r["cycles"] == K
    return pr


def observed_orders(O, spec, stages=2, T=0.25, steps0=10, levels=4, ref_factor=64):
    """Observed order log2(err(dt)/err(dt/2)) for position, velocity and
    energy over ``levels`` halvings, errors against a run with dt/ref_factor."""
    dt0 = T / steps0
    ref = run_fixed_dt(O, spec, dt0 / ref_factor, T, stages)
    errs = []
    for k in range(levels):
        pr = run_fixed_dt(O, spec, dt0 / 2 ** k, T, stages)
        errs.append([np.abs(pr.node_coords[1] - ref.node_coords[1]).max(),
                     np.abs(pr.node_vel[1] - ref.node_vel[1]).max(),
                     np.abs(pr.elem_sie[1] - ref.elem_sie[1]).max()])
    errs = np.array(errs)
    assert np.all(errs > 0)
    return np.log2(errs[:-1] / errs[1:]), errs


def test_rk2_is_second_order(oracle):
    """Pressure-only cycle (no viscosity, so the semi-discrete system is
    smooth): every halving of dt divides the error by ~4."""
    assert order_test_passes(oracle, acoustic_spec())


def test_rk2_is_second_order_with_viscosity(oracle):
    """With the reference's viscosity coefficients (q1 = q1ex = 1,
    q2 = 1.333): the switches of the shock detector make the velocity's
    order uneven from level to level, position and energy stay second order
    and the velocity's mean order too."""
    orders, _ = observed_orders(oracle, acoustic_spec(q=(1.0, 1.0, 1.333, 0.0)))
    assert np.all(orders[:, [0, 2]] > 1.85), orders
    assert orders[:, 1].mean() > 1.8, orders


def test_one_stage_cycle_is_first_order(oracle):
    """The same check on a one-stage cycle (rk_num_stages = 1, alpha = 1)
    sees first order: the order test discriminates."""
    orders, _ = observed_orders(oracle, acoustic_spec(), stages=1)
    assert np.all(orders[:, 1:] < 1.3), orders


def _mutant(tmp_path_factory, which):
    """The oracle compiled with a planted mistake in the RK coefficient."""
    d = tmp_path_factory.mktemp(f"mutant{which}")
    so = str(d / "liboracle_mutant.so")
    src = [os.path.join(ORACLE_DIR, f) for f in ("sgh_oracle.c", "rz_oracle.c")]
    subprocess.run(["gcc", "-std=c11", "-fPIC", "-shared", "-O2", "-ffp-contract=off",
                    "-fopenmp", f"-DORC_PLANTED_MISTAKE={which}", "-o", so, *src, "-lm"],
                   check=True)
    return so


def order_test_passes(O, spec, levels=4):
    orders, _ = observed_orders(O, spec, levels=levels)
    return bool(np.all((orders > 1.85) & (orders < 2.2)))


@pytest.mark.parametrize("which,caught_by", [(1, "order"), (2, "order"), (3, "sod")])
def test_planted_rk_mistakes_are_caught(oracle, tmp_path_factory, which, caught_by):
    """Mutant 1 takes a full step in every stage (alpha = 1): consistent but
    first order.  Mutant 2 uses the weights 1/2, 1/3: inconsistent and first
    order.  Mutant 3 uses 1/4, 1/2: a second-order integrator of the wrong
    equation (y' = f/2), which a self-convergence test cannot see but the
    exact Riemann solution does.  All three still conserve total energy to
    round-off (the discrete energy identity holds for any alpha), so only
    the order and Sod tests tell them from the correct oracle, which passes
    both."""
    L = oracle.load_library(_mutant(tmp_path_factory, which))
    assert [oracle.lib().orc_rk_alpha(2, s) for s in (0, 1)] == [0.5, 1.0]
    assert [L.orc_rk_alpha(2, s) for s in (0, 1)] != [0.5, 1.0]
    key = "liboracle_sgh.so"
    saved = oracle._LIBS[key]
    try:
        oracle._LIBS[key] = L
        spec = acoustic_spec()
        pr = oracle.OracleProblem(spec)
        perturb(pr)
        r = pr.solve(cycle_stop=5, dt_replay=np.full(5, 0.01))
        assert abs(r["te1"] - r["te0"]) <= 1e-13 * abs(r["te0"])    # still conservative
        order_ok = order_test_passes(oracle, spec, levels=3)
        sod_ok, _ = sod_test_passes(oracle)
    finally:
        oracle._LIBS[key] = saved
    assert not (order_ok and sod_ok)
    if caught_by == "order":
        assert not order_ok
    else:
        assert order_ok and not sod_ok


def acoustic_rz_spec(n=8, q=(0.0, 0.0, 0.0, 0.0)):
    """Closed RZ box (axial x, radius y) [0,1] x [0,1] of ideal gas, reflected
    on the four sides (y = 0 is the axis)."""
    spec = P.ProblemSpec(
        name=f"acoustic_rz{n}", mesh=dict(kind="rect", nx=n, ny=n, lx=1.0, ly=1.0, lden=1.0),
        materials=[q], state_vars=P._gas(1.4), mat_num_state_vars=[3],
        fills=[P.fill(P.GLOBAL, den=1.0, sie=2.5)],
        bcs=[(P.X_PLANE, 0.0, P.REFLECTED), (P.X_PLANE, 1.0, P.REFLECTED),
             (P.Y_PLANE, 0.0, P.REFLECTED), (P.Y_PLANE, 1.0, P.REFLECTED)])
    spec.dims = 2
    return spec


def test_rz_rk2_is_second_order(oracle):
    """The 2D RZ cycle (with its per-stage areal node-mass update) under the
    same frozen-step refinement: second order, and first order with one
    stage."""
    def run(dt, T, stages):
        pr = oracle.OracleRZ(acoustic_rz_spec())
        x = pr.node_coords[1]
        v = np.zeros_like(x)
        v[:, 0] = 0.05 * np.sin(np.pi * x[:, 0]) * (1 + 0.3 * x[:, 1])
        v[:, 1] = 0.05 * np.sin(np.pi * x[:, 1]) * (1 - 0.2 * x[:, 0])
        pr.node_vel[0] = v
        pr.node_vel[1] = v
        K = int(round(T / dt))
        pr.solve(cycle_stop=K, dt_replay=np.full(K, dt), rk_num_stages=stages)
        return np.concatenate([pr.node_coords[1].ravel(), pr.node_vel[1].ravel(),
                               pr.elem_sie[1]])

    for stages, lo, hi in ((2, 1.85, 2.2), (1, 0.7, 1.3)):
        T, dt0 = 0.25, 0.025
        ref = run(dt0 / 64, T, stages)
        errs = np.array([np.abs(run(dt0 / 2 ** k, T, stages) - ref).max() for k in range(4)])
        orders = np.log2(errs[:-1] / errs[1:])
        assert np.all((orders > lo) & (orders < hi)), (stages, orders)


def test_riemann_helper_matches_published_star_states():
    """Toro's Test 1 (Sod) star state, Table 4.3 of the reference text:
    p* = 0.30313, u* = 0.92745, rho*_L = 0.42632, rho*_R = 0.26557; and
    the symmetric double-rarefaction (Test 2): p* = 0.00189, u* = 0."""
    ps, us = R.star_state(1.0, 0.0, 1.0, 0.125, 0.0, 0.1, 1.4)
    assert ps == pytest.approx(0.30313, abs=5e-6)
    assert us == pytest.approx(0.92745, abs=5e-6)
    rho, _, _ = R.sample([us * 0.2 + 0.5 - 1e-9, us * 0.2 + 0.5 + 1e-9], 0.2, 0.5,
                         (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    assert rho[0] == pytest.approx(0.42632, abs=5e-6)
    assert rho[1] == pytest.approx(0.26557, abs=5e-6)
    ps, us = R.star_state(1.0, -2.0, 0.4, 1.0, 2.0, 0.4, 1.4)
    assert ps == pytest.approx(0.00189, abs=5e-6)
    assert us == pytest.approx(0.0, abs=1e-12)


def sod_l1(oracle, nx):
    spec = P.sod_tube(nx, 0.125, 0.1, 1.4, cycles=100000, ny=1, nz=1).with_run(time_final=0.2)
    pr = oracle.OracleProblem(spec)
    r = pr.solve()
    assert r["time_value"] == pytest.approx(0.2, abs=1e-15)
    x = pr.node_coords[1][pr.conn]
    xc = x[:, :, 0].mean(axis=1)
    dx = x[:, [1, 2, 5, 6], 0].mean(axis=1) - x[:, [0, 3, 4, 7], 0].mean(axis=1)
    rho, _, p = R.sample(xc, r["time_value"], 0.5, (1.0, 0.0, 1.0), (0.125, 0.0, 0.1), 1.4)
    return np.sum(np.abs(pr.elem_den - rho) * dx), np.sum(np.abs(pr.elem_pres - p) * dx)


def sod_test_passes(oracle):
    errs = np.array([sod_l1(oracle, nx) for nx in (100, 200, 400)])
    ratios = errs[:-1] / errs[1:]
    return bool(np.all(ratios > 1.7) and errs[-1, 0] < 5e-3 and errs[-1, 1] < 5e-3), errs


def test_sod_converges_to_exact_riemann_solution(oracle):
    """Sod tube (nx x 1 x 1 hexes, reflected walls, t = 0.2) against the
    exact solution at the moving element centroids: L1 errors of density
    and pressure fall at first order (shock-capturing) with resolution
    (measured 0.0136, 0.0067, 0.0035 in density at nx = 100, 200, 400)."""
    ok, errs = sod_test_passes(oracle)
    assert ok, errs


def test_output_cadence_by_hand(oracle):
    """Every step is exactly dt_max = 2^-10 (dt_start = dt_max, the CFL
    limit is far larger and the graphics clamp (gt - t) + fuzz never binds),
    so times are exact binary fractions.  With graphics_dt_ival = 5 * 2^-10
    and graphics_cyc_ival = 13 the write rule of K/sgh_solve.cpp:594-628
    gives, by hand, outputs after cycles 5, 10, 13 (cycle rule), 18, 23, 26
    (cycle rule), 31, 36, 39 (cycle rule)."""
    h = 2.0 ** -10
    spec = P.noh(6, cycles=40).with_run(dt_start=h, dt_max=h, graphics_dt_ival=5 * h,
                                        graphics_cyc_ival=13)
    pr = oracle.OracleProblem(spec)
    r = pr.solve(record_dt=True, record_writes=True)
    assert r["cycles"] == 40
    assert np.all(r["dt_history"] == h)
    assert r["time_value"] == 40 * h
    assert r["write_cycles"] == [5, 10, 13, 18, 23, 26, 31, 36, 39]

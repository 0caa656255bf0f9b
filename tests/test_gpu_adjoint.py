"""Discrete adjoint and tangent-linear model of the fused SGH cycle.

The model is the cycle with the time steps frozen to those of the forward
run.  Checks:
  * the checkpointed forward pass is the fused cycle loop, bit for bit;
  * dot-product test <J v, w> = <v, J^T w> between the tangent-linear
    model and the reverse pass, to 1e-10 relative;
  * gradients independent of the checkpoint interval (bitwise);
  * central finite differences (h = 1e-6) of the oracle with the same
    frozen time steps against the adjoint gradient;
  * the anvil gradient (16x16x96, 400 cycles, checkpoint every 20) of
    bar length + base spread with respect to the 24,576 initial densities
    and gamma against central differences on 32 random directions.

Finite differences need a point where the forward map is differentiable.
The anvil starts in uniform motion: in every element ahead of the shock
|v - v_bar| = 0 and div = 0, where the viscosity has kinks (|.|, the
compression switch q2 -> q2ex, |div| in the limiter, and the shock detector
ratio, whose scale there is `small` = 1e-8).  Perturbations of size 1e-6
cross those kinks, so central differences at that point carry an error of
order 1e-3 that shrinks only slowly with h, while the derivative (the limit
h -> 0, which the adjoint computes) exists.  The tight checks therefore use
a generic state (the same problem with a small random initial velocity
field that respects the boundary conditions), where central differences
converge; the uniform-motion anvil is checked to the accuracy that its
kinks allow.
"""
import numpy as np
import pytest

from fierro_amd import problems as P

pytestmark = pytest.mark.gpu

INIT = ["node_coords", "node_vel", "node_mass", "elem_den", "elem_pres", "elem_sspd",
        "elem_sie", "elem_vol", "elem_div", "elem_mass", "elem_gamma", "elem_csmin"]


@pytest.fixture(scope="module")
def Solver():
    from fierro_amd.sgh import SGHSolver
    return SGHSolver


def snapshot(g):
    return {k: g.download(k).copy() for k in INIT}


def controls_state(base, conn, rho, sie, gamma, vel):
    """Initial state from the controls, as the set-up builds it: M = rho V0,
    node mass sum M/8, ideal-gas pressure and sound speed (csmin floor)."""
    st = {k: v.copy() for k, v in base.items()}
    vol = base["elem_vol"]
    M = rho * vol
    m = np.zeros(base["node_mass"].shape)
    np.add.at(m, conn.reshape(-1), np.repeat(0.125 * M, 8))
    st["elem_den"] = rho.copy()
    st["elem_mass"] = M
    st["node_mass"] = m
    st["elem_sie"] = np.stack([sie, sie])
    st["elem_gamma"] = gamma.copy()
    st["elem_pres"] = (gamma - 1.0) * sie * rho
    st["elem_sspd"] = np.maximum(np.sqrt(gamma * (gamma - 1.0) * sie), base["elem_csmin"])
    st["node_vel"] = np.stack([vel, vel])
    return st


def load(g, st):
    for k in INIT:
        g.upload(k, st[k])


def generic_velocity(spec, coords, vel, amp, seed):
    """vel + a random field of amplitude amp whose components normal to the
    reflected planes x = 0, y = 0, z = 0 are zero on those planes."""
    rng = np.random.default_rng(seed)
    noise = amp * rng.standard_normal(vel.shape)
    for d in range(3):
        noise[np.abs(coords[:, d]) < 1e-12, d] = 0.0
    return vel + noise


def final_state(g):
    return (g.download("node_coords")[1], g.download("node_vel")[1], g.download("elem_sie")[1])


def random_dir(rng, N, E):
    return dict(den0=rng.standard_normal(E), sie0=1e-3 * rng.standard_normal(E),
                gamma=0.1 * rng.standard_normal(E), vel0=0.01 * rng.standard_normal((N, 3)))


def dot_grad(g, d):
    return sum(float(np.sum(g[k] * d[k])) for k in ("den0", "sie0", "gamma", "vel0"))


def test_forward_is_fused_loop(Solver):
    spec = P.anvil(4, 4, 12, cycles=30)
    a = Solver(spec)
    a.forward(30, checkpoint_every=7)
    b = Solver(spec)
    b.sgh_solve(cycles=30, use_graph=False)
    for k in ("node_coords", "node_vel", "elem_sie", "elem_den", "elem_pres", "elem_div"):
        assert np.array_equal(a.download(k), b.download(k)), k
    assert a.dt_history()[-1, 0] == b.status()["dt"]
    # replaying the recorded steps reproduces the run
    dt = a.dt_history()
    c = Solver(spec)
    c.forward(30, checkpoint_every=30, dt_replay=dt)
    assert np.array_equal(c.download("node_coords"), a.download("node_coords"))


def test_dot_product_sparse_seed(Solver):
    """A seed on a few nodes and elements near the origin: the adjoint
    spreads about two element layers per reversed stage, so over 3 cycles
    of a 32^3 mesh most waves and node blocks are never reached and the
    reverse pass skips them; the tangent model skips nothing, and
    <Jv,w> = <v,J^T w> holds as with a dense seed."""
    spec = P.sedov(32, cycles=3)
    g = Solver(spec, renumber=True)
    g.forward(3, checkpoint_every=2)
    rng = np.random.default_rng(11)
    v = random_dir(rng, g.N, g.E)
    x1 = g.download("node_coords")[1]
    near = np.argsort(np.linalg.norm(x1, axis=1))[:6]
    w = dict(x=np.zeros((g.N, 3)), v=np.zeros((g.N, 3)), e=np.zeros(g.E))
    w["v"][near] = rng.standard_normal((6, 3))
    w["x"][near[3:]] = rng.standard_normal((3, 3))
    elems = np.where(np.isin(g.conn, near).any(axis=1))[0]
    w["e"][elems] = rng.standard_normal(elems.size)
    t = g.tangent(v["den0"], v["sie0"], v["gamma"], v["vel0"])
    grad = g.backward(w["x"], w["v"], w["e"])
    lhs = sum(float(np.sum(t[k] * w[k])) for k in ("x", "v", "e"))
    rhs = dot_grad(grad, v)
    print(f"sparse seed: <Jv,w> = {lhs:.17e}  <v,J^T w> = {rhs:.17e}, "
          f"{np.count_nonzero(grad['den0'])} of {g.E} elements reached")
    assert rhs != 0.0
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))
    # most of the mesh is not reached (its gradient is zero)
    assert np.count_nonzero(grad["den0"]) < g.E // 2


def test_forward_refuses_advanced_state(Solver):
    """fsgh_forward differentiates from the set-up state: on a state that a
    run has advanced it fails (FSGH_E_STATE) until the state is set again."""
    from fierro_amd._abi import FsghError

    spec = P.anvil(4, 4, 12, cycles=10)
    g = Solver(spec)
    st = {k: g.download(k) for k in INIT}
    g.forward(10, checkpoint_every=5)
    x1 = g.download("node_coords")
    with pytest.raises(FsghError, match="advanced"):
        g.forward(10, checkpoint_every=5)
    # setting gamma, velocities and energies again is not enough: the
    # coordinates are still the advanced ones
    for k in ("elem_gamma", "node_vel", "elem_sie"):
        g.upload(k, st[k])
    with pytest.raises(FsghError, match="advanced"):
        g.forward(10, checkpoint_every=5)
    load(g, st)
    g.forward(10, checkpoint_every=5)
    assert np.array_equal(g.download("node_coords"), x1)
    h = Solver(spec)
    h.sgh_solve(cycles=3, use_graph=False)
    with pytest.raises(FsghError, match="advanced"):
        h.forward(10, checkpoint_every=5)


def test_dt_history_matches_oracle(oracle, Solver):
    spec = P.anvil(4, 4, 12, cycles=30)
    g = Solver(spec, renumber=True)
    g.forward(30, checkpoint_every=10)
    ro = oracle.OracleProblem(spec).solve(record_dt=True)
    assert np.allclose(g.dt_history()[:, 0], ro["dt_history"], rtol=1e-12, atol=0)


@pytest.mark.parametrize("case", ["anvil", "sedov", "sedov_user"])
def test_dot_product(Solver, case):
    spec = {"anvil": lambda: P.anvil(4, 4, 12, cycles=20),
            "sedov": lambda: P.sedov(6, cycles=20),
            "sedov_user": lambda: P.sedov_user_region(6, cycles=20)}[case]()
    g = Solver(spec)
    g.forward(20, checkpoint_every=7)
    rng = np.random.default_rng(3)
    v = random_dir(rng, g.N, g.E)
    w = dict(x=rng.standard_normal((g.N, 3)), v=rng.standard_normal((g.N, 3)),
             e=rng.standard_normal(g.E))
    t = g.tangent(v["den0"], v["sie0"], v["gamma"], v["vel0"])
    grad = g.backward(w["x"], w["v"], w["e"])
    lhs = sum(float(np.sum(t[k] * w[k])) for k in ("x", "v", "e"))
    rhs = dot_grad(grad, v)
    print(f"{case}: <Jv,w> = {lhs:.17e}  <v,J^T w> = {rhs:.17e}")
    assert abs(lhs - rhs) <= 1e-10 * max(abs(lhs), abs(rhs))
    # the reverse pass left the forward result in place
    x1 = g.download("node_coords")[1]
    h = Solver(spec)
    h.forward(20, checkpoint_every=20)
    assert np.array_equal(x1, h.download("node_coords")[1])


def test_checkpoint_interval_invariance(Solver):
    spec = P.anvil(4, 4, 12, cycles=15)
    rng = np.random.default_rng(5)
    out = []
    for every in (1, 4, 15):
        g = Solver(spec)
        g.forward(15, checkpoint_every=every)
        rs = np.random.default_rng(9)
        out.append(g.backward(rs.standard_normal((g.N, 3)), rs.standard_normal((g.N, 3)),
                              rs.standard_normal(g.E)))
    for k in out[0]:
        assert np.array_equal(out[0][k], out[1][k]), k
        assert np.array_equal(out[0][k], out[2][k]), k
    del rng


def test_fd_against_oracle(oracle, Solver):
    """Central differences of the oracle (same frozen steps) against the
    adjoint gradient, for perturbations of rho0, e0, gamma and v0, at a
    generic state."""
    spec = P.anvil(4, 4, 16, cycles=30)
    n = 30
    g = Solver(spec, renumber=True)
    base = snapshot(g)
    v = generic_velocity(spec, g.coords0, base["node_vel"][1], 0.02, 77)
    base["node_vel"] = np.stack([v, v])
    load(g, base)
    g.forward(n, checkpoint_every=8)
    dt = g.dt_history()
    rng = np.random.default_rng(11)
    w = dict(x=rng.standard_normal((g.N, 3)), v=rng.standard_normal((g.N, 3)),
             e=rng.standard_normal(g.E))
    grad = g.backward(w["x"], w["v"], w["e"])
    rho0, sie0 = base["elem_den"], base["elem_sie"][1]
    gam0, vel0 = base["elem_gamma"], base["node_vel"][1]
    h = 1e-6
    for k in range(4):
        d = random_dir(rng, g.N, g.E)
        vals = []
        for s in (1.0, -1.0):
            pr = oracle.OracleProblem(spec)      # set up on construction
            b = dict(elem_vol=pr.elem_vol.copy(), node_mass=pr.node_mass.copy(),
                     elem_csmin=pr.elem_statev[:, 1].copy())
            st = controls_state(dict(base, **b), pr.conn, rho0 + s * h * d["den0"],
                                sie0 + s * h * d["sie0"], gam0 + s * h * d["gamma"],
                                vel0 + s * h * d["vel0"])
            pr.elem_den[:] = st["elem_den"]
            pr.elem_mass[:] = st["elem_mass"]
            pr.node_mass[:] = st["node_mass"]
            pr.elem_sie[:] = st["elem_sie"]
            pr.elem_pres[:] = st["elem_pres"]
            pr.elem_sspd[:] = st["elem_sspd"]
            pr.elem_statev[:, 0] = st["elem_gamma"]
            pr.node_vel[:] = st["node_vel"]
            pr.solve(cycle_stop=n, dt_replay=dt[:, 0])
            vals.append(float(np.sum(pr.node_coords[1] * w["x"]) + np.sum(pr.node_vel[1] * w["v"])
                              + np.sum(pr.elem_sie[1] * w["e"])))
        fd = (vals[0] - vals[1]) / (2 * h)
        ad = dot_grad(grad, d)
        print(f"direction {k}: fd {fd:.12e} adjoint {ad:.12e} rel {abs(fd - ad) / abs(ad):.2e}")
        assert abs(fd - ad) <= 1e-6 * abs(ad)


def anvil_fd(g, base, seed, dt, n, ndir, h, rng):
    """Central differences of the anvil objective along random directions in
    (rho0, gamma), against the adjoint; returns (fd, adjoint) pairs."""
    grad = g.backward(seed_x=seed)
    g_rho, g_gamma = grad["den0"], float(np.sum(grad["gamma"]))
    assert np.all(np.isfinite(g_rho)) and np.isfinite(g_gamma)
    rho0, sie0 = base["elem_den"], base["elem_sie"][1]
    gam0, vel0 = base["elem_gamma"], base["node_vel"][1]
    out = []
    for k in range(ndir):
        d_rho = rng.standard_normal(g.E)
        d_gam = rng.standard_normal()
        vals = []
        for s in (1.0, -1.0):
            st = controls_state(base, g.conn, rho0 + s * h * d_rho, sie0,
                                gam0 + s * h * d_gam, vel0)
            load(g, st)
            g.forward(n, checkpoint_every=n, dt_replay=dt)
            vals.append(float(np.sum(seed * g.download("node_coords")[1])))
        out.append(((vals[0] - vals[1]) / (2 * h), float(np.dot(g_rho, d_rho)) + g_gamma * d_gam))
    return out


def test_anvil_gradient_generic_state(Solver):
    """The anvil objective's gradient (16x16x96, 400 cycles, checkpoint
    every 20) at a generic state: central differences (h = 1e-6) agree to
    1e-3 relative.  Over 400 cycles of a strong shock a few of the limiter
    and switch kinks still fall within +-h of some element's state, which
    bounds the agreement of any finite difference at this size (the 30-cycle
    test against the oracle agrees to ~1e-7)."""
    spec = P.anvil(16, 16, 96, cycles=400)
    n = 400
    g = Solver(spec, renumber=True)
    base = snapshot(g)
    v = generic_velocity(spec, g.coords0, base["node_vel"][1], 0.01, 5)
    base["node_vel"] = np.stack([v, v])
    load(g, base)
    seed = P.anvil_objective(spec, g.coords0)
    g.forward(n, checkpoint_every=20)
    dt = g.dt_history()
    pairs = anvil_fd(g, base, seed, dt, n, 6, 1e-6, np.random.default_rng(7))
    for k, (fd, ad) in enumerate(pairs):
        print(f"generic dir {k}: fd {fd:+.10e} adjoint {ad:+.10e} rel {abs(fd - ad) / abs(ad):.2e}")
    assert max(abs(fd - ad) / abs(ad) for fd, ad in pairs) <= 1e-3


def test_anvil_gradient_fd_32_directions(Solver):
    """The anvil objective's gradient with respect to the 24,576 initial
    densities and gamma (one scalar for the bar), 400 cycles, checkpoint
    every 20, from the uniform-motion start, against central differences
    (h = 1e-6) on 32 random directions with the same frozen steps.  The
    start is a kink point of the viscosity (module docstring): the bound is
    on the error relative to the RMS directional derivative."""
    spec = P.anvil(16, 16, 96, cycles=400)
    n = 400
    g = Solver(spec, renumber=True)
    base = snapshot(g)
    seed = P.anvil_objective(spec, g.coords0)
    g.forward(n, checkpoint_every=20)
    dt = g.dt_history()
    J0 = float(np.sum(seed * g.download("node_coords")[1]))
    pairs = anvil_fd(g, base, seed, dt, n, 32, 1e-6, np.random.default_rng(2024))
    rms = float(np.sqrt(np.mean([ad * ad for _, ad in pairs])))
    worst = 0.0
    for k, (fd, ad) in enumerate(pairs):
        worst = max(worst, abs(fd - ad) / rms)
        print(f"J={J0:.10f} dir {k:2d}: fd {fd:+.10e} adjoint {ad:+.10e} "
              f"err/rms {abs(fd - ad) / rms:.2e}")
    assert worst <= 5e-3


def test_calibration_gradient_fd(Solver):
    """Batched calibration: gradients of each blast's radial-profile misfit
    with respect to (E0, gamma, rho0), from one batched forward + reverse
    pass, against central differences of batched runs (same frozen steps)
    in which every instance's parameter is perturbed at once."""
    from fierro_amd import calibration as CAL

    cycles, every = 30, 7
    params = [(8, 0.6, 1.4, 1.0), (10, 0.3, 1.7, 1.0), (12, 0.8, 1.5, 1.2)]
    target = CAL.target_profile(Solver, P.blast(10, 0.5, 1.55, 1.0, cycles=cycles), cycles)

    def run(ps, dt=None):
        g = Solver([P.blast(n, E0, gm, r, cycles=cycles) for n, E0, gm, r in ps])
        out = CAL.gradient(g, ps, target, cycles, every, dt_replay=dt)
        g.close()
        return out

    J, dE0, dgam, drho, dt = run(params)
    assert np.all(J > 0)
    h = 1e-6
    for which, grad in ((1, dE0), (2, dgam), (3, drho)):
        Jp = run([p[:which] + (p[which] + h,) + p[which + 1:] for p in params], dt)[0]
        Jm = run([p[:which] + (p[which] - h,) + p[which + 1:] for p in params], dt)[0]
        fd = (Jp - Jm) / (2 * h)
        for i in range(len(params)):
            print(f"param {which} inst {i}: fd {fd[i]:+.10e} adjoint {grad[i]:+.10e}")
        assert np.allclose(fd, grad, rtol=1e-5, atol=1e-9 * np.max(np.abs(grad)))

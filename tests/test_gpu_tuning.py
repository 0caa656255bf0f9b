"""Launch-configuration variants of the fused-cycle kernels.

Every compiled variant (workgroup size, occupancy floor, XCD dealing) must
give bit-identical results: the per-element and per-node arithmetic does not
depend on the launch shape.  The autotuner must leave the state unchanged,
also between cycles of a running loop, where a change of the energy
kernel's table moves the dt partial minima to new block ranges.
"""
import numpy as np
import pytest

from fierro_amd import _abi, problems as P

pytestmark = pytest.mark.gpu

FIELDS = ["node_coords", "node_vel", "elem_den", "elem_pres", "elem_sspd", "elem_sie",
          "elem_vol", "elem_div"]


def batch():
    return [P.sod_tube(37, 0.2, 0.3, 1.5, cycles=25), P.sedov(7, cycles=25),
            P.anvil(3, 3, 10, cycles=25)]


def run_fields(cfg=None, cycles=25, graph=True, make=batch):
    from fierro_amd.sgh import SGHSolver
    g = SGHSolver(make())
    try:
        if cfg:
            g.set_tuning(cfg)
        g.run_begin()
        g.run(cycles, use_graph=graph)
        out = {k: g.download(k) for k in FIELDS}
        out["status"] = [g.status(i) for i in range(g.ninst)]
        return out
    finally:
        g.close()


def same(a, b):
    for k in FIELDS:
        assert np.array_equal(a[k], b[k]), k
    assert a["status"] == b["status"]


HEAVY = [(bs, w, x) for bs in (64, 128, 256) for w in (1, 3) for x in (0, 1)]
LIGHT = [(bs, x) for bs in (64, 128, 256, 512) for x in (0, 1)]


@pytest.fixture(scope="module")
def reference():
    return run_fields()


@pytest.mark.parametrize("i", range(len(HEAVY)))
def test_variants_bitwise(reference, i):
    """Each heavy variant for force and energy, and a light variant for the
    node kernels, in one configuration; all 12 + 8 variants are covered."""
    bs, w, x = HEAVY[i]
    lb, lx = LIGHT[i % len(LIGHT)]
    fb, fw, fx = HEAVY[(i + 5) % len(HEAVY)]
    cfg = {"force": dict(block=fb, min_waves=fw, xcd_order=fx),
           "energy_state_div": dict(block=bs, min_waves=w, xcd_order=x),
           "vel_bc_pos": dict(block=lb, min_waves=1, xcd_order=lx),
           "div_extremes": dict(block=LIGHT[(i + 3) % len(LIGHT)][0], min_waves=1,
                                xcd_order=LIGHT[(i + 3) % len(LIGHT)][1])}
    same(run_fields(cfg), reference)


def big_batch():
    # 292,296 elements: 4,568 work items of 64 (tables of fewer than 4,096
    # keep one contiguous range per XCD; chunks are at most 1/32 of the table)
    return [P.sod_tube(300, 0.2, 0.3, 1.5, cycles=25), P.sedov(66, cycles=25)]


@pytest.fixture(scope="module")
def big_reference():
    return run_fields(make=big_batch)


@pytest.mark.parametrize("chunk", [2, 3, 16])
def test_chunk_dealing_bitwise(big_reference, chunk):
    """Work items dealt to the XCDs in chunks of consecutive blocks (a chunk
    count that does not divide the block count leaves uneven XCD queues; 16
    is cut to the table's limit): the same results bit for bit."""
    cfg = {"force": dict(block=64, min_waves=3, xcd_order=chunk),
           "energy_state_div": dict(block=64, min_waves=1, xcd_order=chunk),
           "vel_bc_pos": dict(block=64, min_waves=1, xcd_order=chunk),
           "div_extremes": dict(block=128, min_waves=1, xcd_order=chunk)}
    same(run_fields(cfg, make=big_batch), big_reference)
    same(run_fields(cfg, graph=False, make=big_batch), big_reference)


def test_operator_path_under_tuning():
    """The per-stage force operator (and its divergence-extremes prepass) run
    in the context's launch configuration: the reference's cycle through the
    operator API gives the same fields bit for bit under any configuration."""
    from fierro_amd.sgh import SGHSolver
    from fierro_amd.tuning import TABLE

    def ops(cfg):
        spec = P.sedov(40, cycles=6)
        r = spec.run
        g = SGHSolver(spec)
        try:
            if cfg:
                g.set_tuning(cfg)
            t, dt = 0.0, r["dt_start"]
            for _ in range(6):
                dt = g.get_timestep(t, r["graphics_dt_ival"], r["time_final"], r["dt_max"],
                                    r["dt_min"], r["dt_cfl"], dt, r["fuzz"])
                g.rk_init()
                for stage in range(2):
                    g.stage(dt, 1.0 / (2.0 - stage))
                t += dt
            return {k: g.download(k) for k in FIELDS}
        finally:
            g.close()

    ref = ops(None)
    for cfg in (TABLE, {"force": dict(block=64, min_waves=3, xcd_order=0),
                        "div_extremes": dict(block=512, min_waves=1, xcd_order=1)}):
        got = ops(cfg)
        for k in FIELDS:
            assert np.array_equal(got[k], ref[k]), k


def test_autotune_midrun_leaves_state(reference):
    from fierro_amd.sgh import SGHSolver
    g = SGHSolver(batch())
    try:
        g.run_begin()
        g.run(9)
        res = g.autotune(reps=2)
        assert set(res) == set(_abi.TUNED_KERNEL_NAMES)
        for v in res.values():
            assert v["ms"] > 0 and v["ms_default"] >= v["ms"]
        assert {k: (v["block"], v["min_waves"], v["xcd_order"]) for k, v in res.items()} == \
            {k: (v["block"], v["min_waves"], v["xcd_order"]) for k, v in g.get_tuning().items()}
        # an energy-table change between cycles moves the dt partials
        g.set_tuning({"energy_state_div": dict(block=64, min_waves=3, xcd_order=0)})
        g.run(7)
        g.set_tuning({"energy_state_div": dict(block=128, min_waves=1, xcd_order=1),
                      "force": dict(block=64, min_waves=1, xcd_order=0)})
        g.run(9)
        got = {k: g.download(k) for k in FIELDS}
        got["status"] = [g.status(i) for i in range(3)]
    finally:
        g.close()
    same(got, reference)


def test_autotune_before_run_and_eager_loop(reference):
    from fierro_amd.sgh import SGHSolver
    g = SGHSolver(batch())
    try:
        g.autotune(reps=1)
        g.run_begin()
        g.run(25, use_graph=False)
        got = {k: g.download(k) for k in FIELDS}
        got["status"] = [g.status(i) for i in range(3)]
    finally:
        g.close()
    same(got, reference)


def test_invalid_variants_rejected():
    from fierro_amd.sgh import SGHSolver
    from fierro_amd._abi import FsghError
    g = SGHSolver(batch())
    try:
        with pytest.raises(FsghError):
            g.set_tuning({"force": dict(block=96)})
        with pytest.raises(FsghError):
            g.set_tuning({"div_extremes": dict(min_waves=3)})
        with pytest.raises(FsghError):
            g.set_tuning({"force": dict(block=512)})
        with pytest.raises(FsghError):
            g.set_tuning({"vel_bc_pos": dict(xcd_order=-1)})
        with pytest.raises(FsghError):
            g.set_tuning({"energy_state_div": dict(xcd_order=5000)})
        assert all(v["block"] == 256 for v in g.get_tuning().values())
    finally:
        g.close()


def test_adjoint_under_tuning():
    """Forward recompute and dt history use the tuned kernels; gradients and
    time steps are bitwise those of the default configuration."""
    from fierro_amd.sgh import SGHSolver

    def grads(cfg):
        g = SGHSolver(P.anvil(4, 4, 12, cycles=12))
        try:
            if cfg:
                g.set_tuning(cfg)
            g.forward(12, checkpoint_every=5)
            rng = np.random.default_rng(3)
            sx = rng.standard_normal((g.N, 3))
            out = g.backward(seed_x=sx)
            return g.dt_history(), out
        finally:
            g.close()

    d0, g0 = grads(None)
    d1, g1 = grads({"force": dict(block=64, min_waves=3, xcd_order=0),
                    "energy_state_div": dict(block=128, min_waves=3, xcd_order=0),
                    "vel_bc_pos": dict(block=512, xcd_order=0),
                    "div_extremes": dict(block=64, xcd_order=0)})
    assert np.array_equal(d0, d1)
    for k in g0:
        assert np.array_equal(g0[k], g1[k]), k

"""Failure detection: fsgh_check finds non-finite or non-physical state (a
NaN or infinity in any field, a non-positive element volume or mass, a
non-positive node mass) and returns FSGH_E_NONFINITE with a message; a
clean state passes.  The bad values are injected from the host (uploads, a
tangled node); computing on them is ordinary arithmetic, so no kernel can
fault, and a fused run with a NaN is bounded by its cycle count."""
import numpy as np
import pytest

from fierro_amd import _abi, problems as P

pytestmark = pytest.mark.gpu

NONFINITE = -4


@pytest.fixture()
def solver():
    from fierro_amd.sgh import SGHSolver
    g = SGHSolver(P.sedov(6, cycles=10), renumber=True)
    yield g
    g.close()


def check(g):
    rc = g.check_state()
    msg = (g.L.fsgh_last_error(g.ctx) or b"").decode()
    return rc, msg


def test_clean_state_passes(solver):
    assert check(solver)[0] == _abi.FSGH_OK
    solver.sgh_solve()
    assert check(solver)[0] == _abi.FSGH_OK


@pytest.mark.parametrize("field,index,value", [
    ("elem_sie", (1, 17), np.nan),
    ("elem_pres", (40,), np.inf),
    ("elem_den", (3,), -np.inf),
    ("elem_sspd", (100,), np.nan),
    ("node_vel", (1, 25, 2), np.nan),
    ("node_coords", (1, 7, 0), np.inf),
    ("elem_vol", (50,), -1.0e-3),
    ("elem_vol", (51,), 0.0),
    ("elem_mass", (9,), 0.0),
    ("node_mass", (12,), -2.0),
])
def test_injected_fault_is_found(solver, field, index, value):
    a = solver.download(field)
    a[index] = value
    solver.upload(field, a)
    rc, msg = check(solver)
    assert rc == NONFINITE, (field, rc)
    assert "non-finite" in msg


def test_tangled_element_gives_negative_volume(solver):
    """A node pushed through the opposite face inverts its element: the
    volume operator gives V < 0 and the check reports it."""
    x = solver.download("node_coords")
    conn = solver.conn
    e = 0
    n_top = conn[e, 6]            # corner (1,1,1) of element 0 (reference numbering)
    x[:, n_top, :] = x[1, conn[e, 0], :] - 0.5 * (x[1, conn[e, 6], :] - x[1, conn[e, 0], :])
    solver.upload("node_coords", x)
    solver.get_vol()
    vol = solver.download("elem_vol")
    assert vol[e] < 0
    assert check(solver)[0] == NONFINITE


def test_nan_spreads_through_the_cycle_and_is_found(solver):
    """A NaN in one element's energy reaches its neighbours within a few
    cycles of the fused loop (which ends at its cycle count) and the check
    finds it; the same run from a clean state passes."""
    e = solver.download("elem_sie")
    e[1, 0] = np.nan
    solver.upload("elem_sie", e)
    solver.sgh_solve(cycles=3)
    den = solver.download("elem_den")
    assert np.isnan(den).sum() + np.isnan(solver.download("elem_pres")).sum() >= 1
    assert check(solver)[0] == NONFINITE

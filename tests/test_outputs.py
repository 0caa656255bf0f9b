"""EnSight Gold and state-dump writers (formats of K/outputs.cpp): the
geometry file reads back through the reference-format reader (oracle and
package readers), scalar/vector files have the reference's layout, and the
state dump's columns are the element centroid, radii and state."""
import os

import numpy as np

from fierro_amd import outputs, problems as P


def oracle_fields(pr):
    return dict(node_coords=pr.node_coords, node_vel=pr.node_vel, node_mass=pr.node_mass,
                elem_den=pr.elem_den, elem_pres=pr.elem_pres, elem_sspd=pr.elem_sspd,
                elem_sie=pr.elem_sie, elem_vol=pr.elem_vol, elem_mass=pr.elem_mass,
                elem_mat_id=pr.elem_mat_id, nodes_in_elem=pr.conn)


def test_ensight_round_trip(oracle, tmp_path):
    pr = oracle.OracleProblem(P.noh(4, cycles=5))
    pr.solve()
    f = oracle_fields(pr)
    w = outputs.EnsightWriter(str(tmp_path))
    base = w.write(f, 0.0)
    w.write(f, 0.125)
    x, c = oracle.read_ensight(base + ".geo")
    assert np.array_equal(c, pr.conn)
    assert np.allclose(x, pr.node_coords[1], rtol=1e-5, atol=1e-12)
    xp, cp = P.read_ensight(base + ".geo")
    assert np.array_equal(cp, c) and np.array_equal(xp, x)
    den = open(base + ".den").read().split("\n")
    assert den[:4] == ["Per_elem scalar values", "part", "         1", "hexa8"]
    assert np.allclose(np.array(den[4:4 + pr.E], float), pr.elem_den, rtol=1e-5)
    sw = np.array(open(base + ".elem_switch").read().split("\n")[4:4 + pr.E], float)
    assert np.array_equal(sw[:4], [1, -1, 1, -1])
    vel = open(base + ".vel").read().split("\n")
    assert vel[3] == "block"
    v = np.array(vel[4:4 + 3 * pr.N], float).reshape(3, pr.N).T
    assert np.allclose(v, pr.node_vel[1], rtol=1e-5, atol=1e-12)
    case = open(os.path.join(tmp_path, "ensight", "Outputs_SGH.case")).read()
    assert "model: data/Outputs_SGH.*****.geo" in case
    assert "number of steps:    2" in case
    assert case.strip().endswith("1.25000e-01")
    speed = np.array(open(base + ".speed").read().split("\n")[4:4 + pr.E], float)
    vmean = pr.node_vel[1][pr.conn].mean(1)
    assert np.allclose(speed, np.linalg.norm(vmean, axis=1), rtol=1e-5, atol=1e-12)


def test_state_dump(oracle, tmp_path):
    pr = oracle.OracleProblem(P.sedov(4, cycles=3))
    pr.solve()
    path = outputs.write_state(oracle_fields(pr), 0.0123, str(tmp_path))
    assert os.path.basename(path) == "elem_state_t1.23000e-02.txt"
    lines = open(path).read().split("\n")
    assert lines[0] == "# state dump file"
    rows = np.array([list(map(float, ln.split())) for ln in lines[2:2 + pr.E]])
    cen = pr.node_coords[1][pr.conn].mean(1)
    assert np.allclose(rows[:, :3], cen, atol=1e-6)
    assert np.allclose(rows[:, 4], np.linalg.norm(cen, axis=1), atol=1e-6)
    assert np.allclose(rows[:, 5], pr.elem_den, atol=1e-6)


def test_ensight_and_state_rz(oracle, tmp_path):
    """A 2D RZ dump: quad4 elements with 4 node ids, a zero third coordinate
    and velocity component, speed and centroid over the 4 nodes
    (K/outputs.cpp:131-157,184-199,232-267,459-476)."""
    pr = oracle.OracleRZ(P.sedov_rz(4, cycles=5))
    pr.solve()
    f = oracle_fields(pr)
    base = outputs.EnsightWriter(str(tmp_path)).write(f, 0.0)
    geo = open(base + ".geo").read().split("\n")
    N, E = pr.N, pr.E
    assert int(geo[8]) == N
    xyz = np.array(geo[9:9 + 3 * N], float).reshape(3, N)
    assert np.allclose(xyz[:2].T, pr.node_coords[1], rtol=1e-5, atol=1e-12)
    assert np.all(xyz[2] == 0.0)
    assert geo[9 + 3 * N] == "quad4" and int(geo[10 + 3 * N]) == E
    ids = np.array([list(map(int, ln.split())) for ln in geo[11 + 3 * N:11 + 3 * N + E]])
    assert np.array_equal(ids, pr.conn + 1)
    assert open(base + ".den").read().split("\n")[3] == "quad4"
    v = np.array(open(base + ".vel").read().split("\n")[4:4 + 3 * N], float).reshape(3, N)
    assert np.allclose(v[:2].T, pr.node_vel[1], rtol=1e-5, atol=1e-12) and np.all(v[2] == 0.0)
    speed = np.array(open(base + ".speed").read().split("\n")[4:4 + E], float)
    vmean = pr.node_vel[1][pr.conn].mean(1)
    assert np.allclose(speed, np.linalg.norm(vmean, axis=1), rtol=1e-5, atol=1e-12)
    assert speed.max() > 0

    path = outputs.write_state(f, 0.5, str(tmp_path))
    rows = np.array([list(map(float, ln.split()))
                     for ln in open(path).read().split("\n")[2:2 + E]])
    cen = pr.node_coords[1][pr.conn].mean(1)
    assert np.allclose(rows[:, :2], cen, atol=1e-6) and np.all(rows[:, 2] == 0.0)
    assert np.allclose(rows[:, 3], np.linalg.norm(cen, axis=1), atol=1e-6)
    assert np.allclose(rows[:, 4], rows[:, 3])

"""Mesh input and connectivity: oracle pinned by the reference's mesh fixtures
and closed forms, the library's host maps bit-exact against the oracle.

Fixtures: tests/golden/mesh_2x2x2.geo and mesh_Sedov_8.geo are copies of the
reference's Explicit-Lagrange/meshes files; geo_digests.json holds SHA-256
digests of the parsed arrays of larger reference meshes (written by
tests/golden/make_geo_digests.py).
"""
import hashlib
import json
import os

import numpy as np
import pytest

from fierro_amd import problems as P

from conftest import REFERENCE, have_reference

GOLD = os.path.join(os.path.dirname(__file__), "golden")


def sha(a):
    return hashlib.sha256(np.ascontiguousarray(a).tobytes()).hexdigest()


@pytest.mark.parametrize("name,n,L", [("mesh_2x2x2.geo", 2, (2.0, 1.0)),
                                      ("mesh_Sedov_8.geo", 8, (12.0, 10.0))])
def test_generators_reproduce_reference_meshes(oracle, name, n, L):
    x, c = oracle.read_ensight(os.path.join(GOLD, name))
    xg, cg = oracle.gen_box(n, n, n, L[0], L[0], L[0], L[1])
    assert np.array_equal(c, cg) and np.array_equal(x, xg)
    xp, cp = P.read_ensight(os.path.join(GOLD, name))
    assert np.array_equal(xp, x) and np.array_equal(cp, c)
    xb, cb = P.box_mesh(n, n, n, L[0], L[0], L[0], L[1])
    assert np.array_equal(xb, x) and np.array_equal(cb, c)


@pytest.mark.parametrize("name,n", [("mesh_Sedov_16.geo", 16), ("mesh_Sedov_32.geo", 32),
                                    ("mesh_12x12x12.geo", 12), ("mesh_4x4x4.geo", 4)])
def test_generator_matches_reference_digests(name, n):
    d = json.load(open(os.path.join(GOLD, "geo_digests.json")))[name]
    x, c = P.box_mesh(n, n, n, 12.0, 12.0, 12.0, 10.0)
    assert x.shape[0] == d["num_nodes"] and c.shape[0] == d["num_elems"]
    assert sha(c) == d["conn_sha256"]
    assert sha(x) == d["coords_sha256"]
    if have_reference():
        xr, cr = P.read_ensight(os.path.join(REFERENCE, "Explicit-Lagrange", "meshes", name))
        assert np.array_equal(xr, x) and np.array_equal(cr, c)


@pytest.mark.parametrize("dims", [(1, 1, 1), (2, 2, 2), (3, 2, 4), (5, 5, 5), (7, 3, 2)])
def test_connectivity_closed_forms(oracle, dims):
    nx, ny, nz = dims
    x, c = oracle.gen_box(nx, ny, nz, 1.0, 1.0, 1.0, 1.0)
    cx = oracle.connectivity(x, c)
    E = nx * ny * nz
    # every corner appears once, ascending element order per node (K/mesh.h:218-241)
    st, cin = cx["corners_in_node_start"], cx["corners_in_node"]
    assert np.array_equal(np.sort(cin), np.arange(8 * E))
    for n in range(x.shape[0]):
        lst = cin[st[n]:st[n + 1]]
        assert np.all(np.diff(lst) > 0)
        assert np.all(c.ravel()[lst] == n)
    # face counts of a box
    faces = nx * ny * (nz + 1) + nx * (ny + 1) * nz + (nx + 1) * ny * nz
    assert cx["num_patches"] == faces
    assert len(cx["bdy_patches"]) == 2 * (nx * ny + ny * nz + nx * nz)
    inner = max(nx - 1, 0) * max(ny - 1, 0) * max(nz - 1, 0)
    assert len(cx["bdy_nodes"]) == (nx + 1) * (ny + 1) * (nz + 1) - inner
    # elems_in_elem: brute force "shares a node" sets, first-seen order
    ees, ee = cx["elems_in_elem_start"], cx["elems_in_elem"]
    nodes = [set(r) for r in c.tolist()]
    for e in range(E):
        got = list(ee[ees[e]:ees[e + 1]])
        brute = {f for f in range(E) if f != e and nodes[e] & nodes[f]}
        assert set(got) == brute and len(got) == len(brute)


FACE_NODES = [(0, 4, 7, 3), (1, 2, 6, 5), (0, 1, 5, 4), (2, 3, 7, 6), (0, 3, 2, 1), (4, 5, 6, 7)]


def first_seen_sets(x, conn, bcs):
    """Boundary-node sets by brute force: a face is on the boundary when its
    node set occurs once in the mesh; walking the boundary faces by element,
    then face (K/mesh.h:353-358 face node order), a node joins a set the first
    time a face on that set's plane holds it."""
    from collections import Counter

    key = lambda e, f: frozenset(int(conn[e, a]) for a in FACE_NODES[f])  # noqa: E731
    count = Counter(key(e, f) for e in range(len(conn)) for f in range(6))
    out = []
    for surf, val, _ in bcs:
        seen, order = set(), []
        for e in range(len(conn)):
            for f in range(6):
                fn = [int(conn[e, a]) for a in FACE_NODES[f]]
                if count[key(e, f)] != 1 or not all(abs(x[n, surf] - val) <= 1e-7 for n in fn):
                    continue
                for n in fn:
                    if n not in seen:
                        seen.add(n)
                        order.append(n)
        out.append(np.array(order, np.int64))
    return out


@pytest.mark.parametrize("dims", [(2, 2, 2), (3, 2, 4), (5, 1, 2)])
def test_bdy_set_order_first_seen(oracle, dims):
    """The oracle's boundary sets (patch numbering through hash keys and
    element neighbours, K/mesh.h:395-513) hold their nodes in the order the
    brute-force walk meets them; the order is not ascending."""
    spec = P.sod_tube(dims[0], 0.125, 0.1, 1.4, ny=dims[1], nz=dims[2])
    pr = oracle.OracleProblem(spec)
    exp = first_seen_sets(pr.coords0, pr.conn, spec.bcs)
    unsorted = 0
    for s in range(len(spec.bcs)):
        got = pr.bset_nodes[pr.bset_start[s]:pr.bset_start[s + 1]]
        assert np.array_equal(got, exp[s]), s
        unsorted += int(np.any(np.diff(got) < 0))
    assert unsorted > 0


def test_bdy_sets_geometry(oracle):
    spec = P.reference_sedov(os.path.join(GOLD, "mesh_Sedov_8.geo"))
    pr = oracle.OracleProblem(spec)
    x = pr.coords0
    for s, (surf, val, _) in enumerate(spec.bcs):
        got = pr.bset_nodes[pr.bset_start[s]:pr.bset_start[s + 1]]
        exp = np.nonzero(np.abs(x[:, surf] - val) <= 1e-7)[0]
        assert np.array_equal(np.sort(got), exp)


@pytest.mark.parametrize("renumber", [False, True])
@pytest.mark.parametrize("dims", [(4, 4, 4), (6, 3, 5), (12, 4, 4)])
def test_library_maps_match_oracle(oracle, dims, renumber):
    from fierro_amd.sgh import MeshMaps

    nx, ny, nz = dims
    spec = P.sod_tube(nx, 0.125, 0.1, 1.4, ny=ny, nz=nz)
    mm = MeshMaps(spec, renumber=renumber)
    cx = oracle.connectivity(mm.coords0, mm.conn)
    st, cin = mm.corners_in_node()
    assert np.array_equal(st, cx["corners_in_node_start"])
    assert np.array_equal(cin, cx["corners_in_node"])
    bs, bn, _ = oracle.bdy_sets(mm.coords0, cx, spec.bcs)
    for s in range(len(spec.bcs)):
        # same nodes in the same (first-seen) order
        assert np.array_equal(mm.bc_set(0, s), bn[bs[s]:bs[s + 1]])
    e, n = mm.renumbering()
    assert np.array_equal(np.sort(e), np.arange(mm.E))
    assert np.array_equal(np.sort(n), np.arange(mm.N))
    if not renumber:
        assert np.array_equal(e, np.arange(mm.E)) and np.array_equal(n, np.arange(mm.N))


def test_morton_order_is_local():
    """Morton renumbering shortens the element-index window over which a node
    is reused (lexicographic order: two k-layers, n^2+n+1 elements)."""
    from fierro_amd.sgh import MeshMaps

    nn = 32
    mm = MeshMaps(P.sedov(nn), renumber=True)
    e, n = mm.renumbering()
    st, cin = mm.corners_in_node()
    el = e[cin >> 3]
    spans = np.maximum.reduceat(el, st[:-1]) - np.minimum.reduceat(el, st[:-1])
    assert np.median(spans) < (nn * nn + nn + 1) / 8
    # the first 8 elements of the new order form an aligned 2x2x2 block
    first8 = np.argsort(e)[:8]
    ijk = np.stack([first8 % nn, (first8 // nn) % nn, first8 // (nn * nn)], 1)
    assert np.all(ijk.max(0) - ijk.min(0) == 1)


def test_batched_maps(oracle):
    from fierro_amd.sgh import MeshMaps

    specs = [P.sod_tube(nx, 0.2, 0.2, 1.4) for nx in (8, 13, 5)]
    mm = MeshMaps(specs, renumber=True)
    e, n = mm.renumbering()
    for i in range(3):
        e0, e1 = mm.elem_off[i], mm.elem_off[i + 1]
        n0, n1 = mm.node_off[i], mm.node_off[i + 1]
        assert np.array_equal(np.sort(e[e0:e1]), np.arange(e0, e1))
        assert np.array_equal(np.sort(n[n0:n1]), np.arange(n0, n1))
        x = mm.coords0[n0:n1]
        conn_i = mm.conn[e0:e1] - n0
        cx = oracle.connectivity(x, conn_i)
        bs, bn, _ = oracle.bdy_sets(x, cx, specs[i].bcs)
        for s, (surf, val, _) in enumerate(specs[i].bcs):
            exp = np.nonzero(np.abs(x[:, surf] - val) <= 1e-7)[0]
            assert np.array_equal(np.sort(mm.bc_set(i, s)) - n0, exp)
            assert np.array_equal(mm.bc_set(i, s) - n0, bn[bs[s]:bs[s + 1]])

"""Graphics and state dumps in the reference's on-disk formats.

* EnSight Gold ASCII case/geometry/variable files (K/outputs.cpp:85-409):
  ``ensight/Outputs_SGH.case`` and ``ensight/data/Outputs_SGH.NNNNN.{geo,den,
  pres,sie,vol,mass,sspd,speed,mat_id,elem_switch,pos,vel}``.
* Element state text dump ``state/elem_state_t<time>.txt`` (K/outputs.cpp:413-513).

The fields come from ``SGHSolver.download`` (reference layout, caller
numbering), so a run on the GPU produces the same files as the reference's
``write_outputs`` at the same state.  A 2D RZ state (coordinates [2][N][2],
4-node connectivity) is written as the reference writes it: quad4 elements,
a zero third coordinate and velocity component, and element means over the
element's own 4 nodes (K/outputs.cpp:131-157,184-199,232-267,459-476).
"""
from __future__ import annotations

import os

import numpy as np

SCALARS = ["den", "pres", "sie", "vol", "mass", "sspd", "speed", "mat_id", "elem_switch"]
VECTORS = ["pos", "vel"]
NAME = "Outputs_SGH"


def gather_fields(solver):
    """Everything write_outputs reads (K/outputs.cpp:29-40), downloaded once."""
    f = {k: solver.download(k) for k in ("node_coords", "node_vel", "node_mass", "elem_den",
                                         "elem_pres", "elem_sspd", "elem_sie", "elem_vol",
                                         "elem_mass", "elem_mat_id")}
    f["nodes_in_elem"] = solver.conn
    return f


def _e(x):
    return "%12.5e" % x


def _lines(values):
    return "".join("%12.5e\n" % v for v in values)


def _pad3(x):
    """N x dims -> N x 3 with a zero third column for 2D (K/outputs.cpp:186-199)."""
    x = np.asarray(x, np.float64)
    if x.shape[1] == 3:
        return x
    return np.concatenate([x, np.zeros((x.shape[0], 3 - x.shape[1]))], axis=1)


def elem_mean(node_field1, conn):
    """Mean of a node field over each element's nodes, summed in ascending
    local order (K/outputs.cpp:131-150, :459-476)."""
    v = _pad3(node_field1)[conn]             # E x nodes x 3
    acc = np.zeros((conn.shape[0], 3))
    for a in range(conn.shape[1]):
        acc = acc + v[:, a, :]
    return acc / conn.shape[1]


def elem_speed(node_vel1, conn):
    """|mean of the element's node velocities| over the mesh's num_dims
    components (K/outputs.cpp:131-157)."""
    dims = np.asarray(node_vel1).shape[1]
    acc = elem_mean(node_vel1, conn)
    s = np.zeros(conn.shape[0])
    for d in range(dims):
        s = s + acc[:, d] * acc[:, d]
    return np.sqrt(s)


class EnsightWriter:
    """Mirror of ``ensight()``: keeps graphics_id and graphics_times."""

    def __init__(self, root="."):
        self.root = root
        self.graphics_id = 0
        self.graphics_times: list[float] = []

    def write(self, fields, time_value):
        root = self.root
        data = os.path.join(root, "ensight", "data")
        os.makedirs(data, exist_ok=True)
        gid = self.graphics_id
        conn = np.asarray(fields["nodes_in_elem"], np.int64)
        speed = elem_speed(fields["node_vel"][1], conn)
        x1 = _pad3(fields["node_coords"][1])
        v1 = _pad3(fields["node_vel"][1])
        E, N = conn.shape[0], x1.shape[0]
        etype = "hexa8\n" if conn.shape[1] == 8 else "quad4\n"
        switch = np.where(np.arange(E) % 2 == 0, 1.0, -1.0)
        scal = {"den": fields["elem_den"], "pres": fields["elem_pres"],
                "sie": fields["elem_sie"][1], "vol": fields["elem_vol"],
                "mass": fields["elem_mass"], "sspd": fields["elem_sspd"], "speed": speed,
                "mat_id": np.asarray(fields["elem_mat_id"], np.float64), "elem_switch": switch}
        base = os.path.join(data, f"{NAME}.{gid:05d}")
        with open(base + ".geo", "w") as f:
            f.write("A graphics dump by Fierro \n")
            f.write("EnSight Gold geometry\n")
            f.write("node id assign\n")
            f.write("element id assign\n")
            f.write("part\n")
            f.write("%10d\n" % 1)
            f.write("Mesh\n")
            f.write("coordinates\n")
            f.write("%10d\n" % N)
            for d in range(3):
                f.write(_lines(x1[:, d]))
            f.write(etype)
            f.write("%10d\n" % E)
            f.write("".join("".join("%10d\t" % (n + 1) for n in row) + "\n"
                            for row in conn.tolist()))
        for name in SCALARS:
            with open(f"{base}.{name}", "w") as f:
                f.write("Per_elem scalar values\n")
                f.write("part\n")
                f.write("%10d\n" % 1)
                f.write(etype)
                f.write(_lines(scal[name]))
        vec = {"pos": x1, "vel": v1}
        for name in VECTORS:
            with open(f"{base}.{name}", "w") as f:
                f.write("Per_node vector values\n")
                f.write("part\n")
                f.write("%10d\n" % 1)
                f.write("block\n")
                for d in range(3):
                    f.write(_lines(vec[name][:, d]))
        self.graphics_times.append(float(time_value))
        with open(os.path.join(root, "ensight", f"{NAME}.case"), "w") as f:
            f.write("FORMAT\n")
            f.write("type: ensight gold\n")
            f.write("GEOMETRY\n")
            f.write(f"model: data/{NAME}.*****.geo\n")
            f.write("VARIABLE\n")
            for name in SCALARS:
                f.write(f"scalar per element: {name} data/{NAME}.*****.{name}\n")
            for name in VECTORS:
                f.write(f"vector per node: {name} data/{NAME}.*****.{name}\n")
            f.write("TIME\n")
            f.write("time set: 1\n")
            f.write("number of steps: %4d\n" % (gid + 1))
            f.write("filename start number: 0\n")
            f.write("filename increment: 1\n")
            f.write("time values: \n")
            f.write(_lines(self.graphics_times))
        self.graphics_id += 1
        return base


def write_state(fields, time_value, root="."):
    """``state_file`` (K/outputs.cpp:413-513): one line per element."""
    os.makedirs(os.path.join(root, "state"), exist_ok=True)
    path = os.path.join(root, "state", "elem_state_t%6.5e.txt" % time_value)
    conn = np.asarray(fields["nodes_in_elem"], np.int64)
    c = elem_mean(fields["node_coords"][1], conn)
    rad2 = np.sqrt(c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1])
    rad3 = np.sqrt(c[:, 0] * c[:, 0] + c[:, 1] * c[:, 1] + c[:, 2] * c[:, 2])
    cols = [c[:, 0], c[:, 1], c[:, 2], rad2, rad3, fields["elem_den"], fields["elem_pres"],
            fields["elem_sie"][1], fields["elem_sspd"], fields["elem_vol"], fields["elem_mass"]]
    with open(path, "w") as f:
        f.write("# state dump file\n")
        f.write("# x  y  z  radius_2D  radius_3D  den  pres  sie  sspd  vol  mass \n")
        for row in zip(*cols):
            f.write("".join("%f\t " % v for v in row) + "\n")
    return path

"""Writes geo_digests.json: SHA-256 digests of the parsed coordinates and
connectivity of reference meshes too large to commit.  Run in a container
that has the reference mounted at /root/reference."""
import hashlib
import json
import os
import sys

import numpy as np

sys.path.insert(0, os.path.join(os.path.dirname(__file__), "..", ".."))
from fierro_amd import problems as P  # noqa: E402

M = "/root/reference/Explicit-Lagrange/meshes/"
out = {}
for name in ["mesh_Sedov_16.geo", "mesh_Sedov_32.geo", "mesh_4x4x4.geo", "mesh_12x12x12.geo"]:
    x, c = P.read_ensight(M + name)
    out[name] = dict(num_nodes=int(x.shape[0]), num_elems=int(c.shape[0]),
                     coords_sha256=hashlib.sha256(np.ascontiguousarray(x).tobytes()).hexdigest(),
                     conn_sha256=hashlib.sha256(np.ascontiguousarray(c).tobytes()).hexdigest(),
                     bbox=[float(x[:, k].min()) for k in range(3)] +
                          [float(x[:, k].max()) for k in range(3)])
json.dump(out, open(os.path.join(os.path.dirname(__file__), "geo_digests.json"), "w"), indent=1)

"""Callers of include/fsgh.h compiled as C and C++ and linked against
libfsgh.so (tests/capi/): the header is valid C11 and C++17, every declared
entry point links, a host-only context driven from C gives the oracle's corner
CSR and boundary sets, and on the GPU the C++ driver -- the reference's
sgh_solve loop with the library's entry points (K/sgh_solve.cpp:125-634),
operator by operator and through fsgh_run -- runs Sedov 16^3 for 100 cycles to
the oracle's density."""
import os
import re
import subprocess

import numpy as np
import pytest

from fierro_amd import _abi, problems as P

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CAPI = os.path.join(ROOT, "tests", "capi")
LIBDIR = os.path.dirname(_abi.LIB_PATH)


def header_functions():
    txt = open(os.path.join(ROOT, "include", "fsgh.h")).read()
    txt = re.sub(r"/\*.*?\*/", "", txt, flags=re.S)
    return sorted(set(re.findall(r"\b(fsgh_\w+)\s*\(", txt)))


def compile_c(tmp_path, src, exe, std, cc):
    out = os.path.join(tmp_path, exe)
    cmd = [cc, f"-std={std}", "-O1", "-Wall", "-Wextra", "-Werror", "-I",
           os.path.join(ROOT, "include"), "-o", out, os.path.join(CAPI, src), f"-L{LIBDIR}",
           "-lfsgh", f"-Wl,-rpath,{LIBDIR}"]
    if cc == "gcc":
        cmd.insert(2, "-pedantic")
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return out


def write_mesh(path, coords, conn, bcs):
    with open(path, "w") as f:
        f.write(f"{coords.shape[0]} {conn.shape[0]}\n")
        f.write(" ".join("%.17g" % v for v in coords.ravel()) + "\n")
        f.write(" ".join(str(int(v)) for v in conn.ravel()) + "\n")
        f.write(f"{len(bcs)}\n")
        for surf, val, hydro in bcs:
            f.write(f"{int(surf)} {int(hydro)} {val:.17g}\n")


def test_c11_entry_point_table_is_the_header():
    src = open(os.path.join(CAPI, "abi_check.c")).read()
    assert sorted(set(re.findall(r"ENTRY\((fsgh_\w+)\)", src))) == header_functions()


def test_c11_host_context_matches_oracle(oracle, tmp_path):
    exe = compile_c(str(tmp_path), "abi_check.c", "abi_check", "c11", "gcc")
    spec = P.sod_tube(5, 0.125, 0.1, 1.4, ny=2, nz=3)
    coords, conn = spec.mesh_arrays()
    mesh = os.path.join(tmp_path, "mesh.txt")
    write_mesh(mesh, coords, conn, spec.bcs)
    r = subprocess.run([exe, mesh], capture_output=True, text=True, timeout=60)
    assert r.returncode == 0, r.stderr
    lines = {ln.split()[0] + ("" if ln.startswith(("cin", "entry")) else ln.split()[1]):
             ln.split()[1 if ln.startswith(("cin", "entry")) else 2:]
             for ln in r.stdout.strip().split("\n")}
    assert int(lines["entry_points"][0]) == len(header_functions())
    cx = oracle.connectivity(coords, conn)
    assert np.array_equal(np.array(lines["cin_start"], np.int64), cx["corners_in_node_start"])
    assert np.array_equal(np.array(lines["cin"], np.int64), cx["corners_in_node"])
    bs, bn, _ = oracle.bdy_sets(coords, cx, spec.bcs)
    for s in range(len(spec.bcs)):
        # same nodes in the reference's first-seen order
        assert np.array_equal(np.array(lines[f"set{s}"], np.int64), bn[bs[s]:bs[s + 1]]), s


def test_cpp_driver_compiles(tmp_path):
    compile_c(str(tmp_path), "sgh_caller.cpp", "sgh_caller", "c++17", "g++")


def write_problem(path, spec):
    coords, conn = spec.mesh_arrays()
    r = spec.run
    nmat = len(spec.materials)
    nsv = spec.num_state_vars
    sv = np.asarray(spec.state_vars, np.float64)[:, :nsv]
    fields = ["volume", "mat_id", "x1", "x2", "y1", "y2", "z1", "z2", "radius1", "radius2",
              "velocity", "u", "v", "w", "speed", "sie", "den"]
    with open(path, "w") as f:
        f.write(f"{coords.shape[0]} {conn.shape[0]}\n")
        f.write(" ".join("%.17g" % v for v in coords.ravel()) + "\n")
        f.write(" ".join(str(int(v)) for v in conn.ravel()) + "\n")
        f.write(f"{nmat}\n")
        for m in spec.materials:
            f.write(" ".join("%.17g" % v for v in m) + "\n")
        f.write(f"{nsv}\n" + " ".join("%.17g" % v for v in sv.ravel()) + "\n")
        f.write(" ".join(str(int(v)) for v in spec.mat_num_state_vars) + "\n")
        f.write(f"{len(spec.bcs)}\n")
        for surf, val, hydro in spec.bcs:
            f.write(f"{int(surf)} {int(hydro)} {val:.17g}\n")
        f.write(f"{len(spec.fills)}\n")
        for fl in spec.fills:
            f.write(" ".join("%.17g" % float(fl.get(k, 0)) for k in fields) + "\n")
        f.write(" ".join("%.17g" % float(r[k]) for k in
                         ("time_final", "dt_start", "dt_max", "dt_min", "dt_cfl",
                          "graphics_dt_ival")) + "\n")
        f.write(f"{int(r['graphics_cyc_ival'])} {int(r['cycle_stop'])}\n")
        f.write(f"{r['fuzz']:.17g} {r['small']:.17g} 2\n")


@pytest.mark.gpu
@pytest.mark.parametrize("mode", ["ops", "run"])
def test_cpp_driver_sedov16(oracle, tmp_path, mode):
    exe = os.path.join(CAPI, "sgh_caller")
    if not os.path.exists(exe):
        raise RuntimeError(f"{exe} is not built (make -C tests/capi, part of the build step)")
    spec = P.sedov(16, cycles=100)
    prob = os.path.join(tmp_path, "problem.txt")
    out = os.path.join(tmp_path, "out.txt")
    write_problem(prob, spec)
    print(f"[sgh_caller {mode}] sedov 16^3, 100 cycles", flush=True)
    r = subprocess.run([exe, prob, out, mode], capture_output=True, text=True, timeout=120)
    print(r.stdout, r.stderr, flush=True)
    assert r.returncode == 0
    head, *den = open(out).read().split("\n")
    cycles, time_value, ie, ke, outputs = head.split()
    den = np.array([float(v) for v in den if v])
    pr = oracle.OracleProblem(spec)
    ro = pr.solve(record_writes=True)
    assert int(cycles) == ro["cycles"] == 100
    assert abs(float(time_value) - ro["time_value"]) <= 1e-12 * ro["time_value"]
    te = float(ie) + float(ke)
    assert abs(te - ro["te1"]) <= 1e-12 * abs(ro["te1"])
    assert int(outputs) == len(ro["write_cycles"])
    assert den.shape == pr.elem_den.shape
    l1 = float(np.sum(np.abs(den - pr.elem_den)) / np.sum(np.abs(pr.elem_den)))
    assert l1 <= 1e-9, l1
    assert float(np.max(np.abs(den - pr.elem_den))) <= 1e-9 * float(np.max(pr.elem_den))

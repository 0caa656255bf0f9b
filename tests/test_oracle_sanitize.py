"""The CPU oracle under AddressSanitizer and UndefinedBehaviorSanitizer:
tests/oracle_sanitize_driver.c runs its set-up (generator, EnSight reader,
corner / element / patch connectivity, boundary sets, fills, masses) and
cycle loop in 3D and 2D RZ, built with gcc -fsanitize=address,undefined and
no recovery, so any out-of-bounds access, leak or undefined behaviour fails
the test."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


@pytest.mark.skipif(shutil.which("gcc") is None, reason="gcc not available")
def test_oracle_clean_under_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_sanitize")
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-fno-omit-frame-pointer", "-ffp-contract=off",
           "-fopenmp", "-fsanitize=address,undefined", "-fno-sanitize-recover=all",
           "-I", ORACLE, os.path.join(HERE, "oracle_sanitize_driver.c"),
           os.path.join(ORACLE, "sgh_oracle.c"), os.path.join(ORACLE, "rz_oracle.c"),
           "-o", exe, "-lm"]
    subprocess.run(cmd, check=True)
    env = dict(os.environ, OMP_NUM_THREADS="2",
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    geo = os.path.join(HERE, "golden", "mesh_Sedov_8.geo")
    r = subprocess.run([exe, geo], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "runtime error" not in r.stderr, r.stderr[-4000:]
    assert r.stdout.count(" ok") == 3, r.stdout

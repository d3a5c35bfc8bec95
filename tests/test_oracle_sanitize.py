"""The C oracle (oracle/c/lbm64.c) under gcc's address and undefined-
behaviour sanitizers: every boundary configuration and exported call on
small odd-sized domains, in the FP64 build and the FP32 build, with no
sanitizer report, finite results, and the density sum of the fully
periodic cases equal to the fluid node count (rho0 = 1) after 30 / 5
iterations: streaming and bounce-back move populations, never create or
drop them."""
import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
SRC = os.path.join(HERE, "c", "oracle_sanitize.c")
PERIODIC = (0, 5)


def _build(tmp_path, real):
    exe = str(tmp_path / ("oracle_sanitize_" + real))
    cmd = ["gcc", "-std=c11", "-g", "-O1", "-DO64_REAL=" + real, "-fsanitize=address,undefined",
           "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", SRC, "-o", exe, "-lm"]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and ("asan" in r.stderr or "ubsan" in r.stderr):
        pytest.skip("gcc sanitizer runtime not available: " + r.stderr[-200:])
    assert r.returncode == 0, r.stderr
    return exe


@pytest.mark.skipif(shutil.which("gcc") is None, reason="needs gcc")
@pytest.mark.parametrize("real", ["double", "float"])
def test_oracle_c_is_clean_under_sanitizers(tmp_path, real):
    exe = _build(tmp_path, real)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=120)
    assert "runtime error" not in r.stderr and "Sanitizer" not in r.stderr, r.stderr[-2000:]
    assert r.returncode == 0, r.stdout + r.stderr[-2000:]
    rows = [ln.split() for ln in r.stdout.splitlines() if ln.startswith("case")]
    assert len(rows) == 6
    tol = 1e-9 if real == "double" else 1e-4
    for row in rows:
        k, fluid, mass0, bad = int(row[1]), int(row[3]), float(row[5]), int(row[7])
        assert bad == 0
        if k in PERIODIC:
            assert abs(mass0 - fluid) <= tol * fluid, row

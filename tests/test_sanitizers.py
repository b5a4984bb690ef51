"""AddressSanitizer and UndefinedBehaviorSanitizer builds of the CPU code
(host only; no GPU code is built with a sanitizer):

* the C oracle, single-partition and partitioned engines together
  (tests/sanitize/oracle_driver.c, which includes oracle/pq_oracle.c);
* the engine's host-side frame code, pq_tableau.h over pq_algebra.h
  (tests/sanitize/tableau_driver.cpp).

Each driver runs random truncated, noisy circuits or random Clifford frames
and checks its own invariants; a sanitizer report or a mismatch fails."""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ENV = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
           OMP_NUM_THREADS="4")


def _have(cc):
    return shutil.which(cc) is not None


def _run(cmd, tmp_path, exe):
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0 and "cannot find" in (r.stderr + r.stdout) and "asan" in (r.stderr + r.stdout):
        pytest.skip("no sanitizer runtime in this toolchain")
    assert r.returncode == 0, r.stderr[-3000:]
    p = subprocess.run([exe], capture_output=True, text=True, env=ENV, timeout=300)
    assert p.returncode == 0, (p.stdout[-2000:], p.stderr[-4000:])
    assert "MISMATCH" not in p.stdout
    assert "runtime error" not in p.stderr and "AddressSanitizer" not in p.stderr, p.stderr[-4000:]
    return p.stdout


@pytest.mark.skipif(not _have("gcc"), reason="gcc not found")
def test_oracle_under_address_and_undefined_sanitizers(tmp_path):
    exe = str(tmp_path / "oracle_driver")
    out = _run(["gcc", "-std=c11", "-ffp-contract=off", "-fopenmp", *SAN,
                os.path.join(HERE, "sanitize", "oracle_driver.c"), "-o", exe, "-lm"], tmp_path, exe)
    assert out.count("ok") == 5


@pytest.mark.skipif(not _have("g++"), reason="g++ not found")
def test_host_tableau_under_address_and_undefined_sanitizers(tmp_path):
    exe = str(tmp_path / "tableau_driver")
    out = _run(["g++", "-std=c++17", *SAN, "-I", os.path.join(ROOT, "propaq_amd", "csrc"),
                os.path.join(HERE, "sanitize", "tableau_driver.cpp"), "-o", exe], tmp_path, exe)
    assert out.count("ok") == 5

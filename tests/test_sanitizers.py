"""The host-only code under AddressSanitizer and UndefinedBehaviorSanitizer.

* tests/host_asan.cpp drives the library's host sources that need no device
  (mesh, run and corner tables, observed crop and prefix rows, random table:
  q-pso_amd/csrc/qpso_tables.h; centroid: qpso_roi.h; file formats and the CPU
  visualisation render: qpso_io.cpp; synthetic scenes: qpso_scene.cpp) and the
  C oracle through ordinary, edge-of-frame and degenerate inputs, all built
  with the sanitizers and the library's own flags.
* The oracle's own CPU tests run again against a sanitizer build of the
  oracle (QPSO_ORACLE_LIB, with the sanitizer runtimes preloaded into Python).
The serving scheduler has its own sanitizer runs (tests/test_scheduler.py).
"""
import os
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = ["-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer", "-g", "-O1"]
ORACLE_FLAGS = ["-std=gnu11", "-ffp-contract=off", "-fno-fast-math", "-fwrapv", "-fopenmp"]
HOST_FLAGS = ["-std=c++17", "-ffp-contract=off", "-fwrapv"]
INC = ["-I", os.path.join(REPO, "include"), "-I", os.path.join(REPO, "q-pso_amd", "csrc"),
       "-I", os.path.join(REPO, "oracle")]


def run(cmd, **kw):
    subprocess.check_call(cmd, **kw)


def test_host_sources_under_asan_ubsan(tmp_path):
    d = str(tmp_path)
    oracle_o = os.path.join(d, "oracle.o")
    run(["gcc", *ORACLE_FLAGS, *SAN, *INC, "-c", os.path.join(REPO, "oracle", "qpso_oracle.c"), "-o", oracle_o])
    objs = [oracle_o]
    for src in ("qpso_io.cpp", "qpso_scene.cpp"):
        o = os.path.join(d, src + ".o")
        run(["g++", *HOST_FLAGS, *SAN, *INC, "-c", os.path.join(REPO, "q-pso_amd", "csrc", src), "-o", o])
        objs.append(o)
    host_o = os.path.join(d, "host.o")
    run(["g++", *HOST_FLAGS, *SAN, *INC, "-c", os.path.join(REPO, "tests", "host_asan.cpp"), "-o", host_o])
    exe = os.path.join(d, "host_asan")
    run(["g++", *SAN, "-fopenmp", "-o", exe, host_o, *objs, "-lm"])
    p = subprocess.run([exe, d], capture_output=True, text=True, timeout=300)
    assert p.returncode == 0 and "ALL PASS" in p.stdout, p.stdout + p.stderr


def _runtime(name):
    path = subprocess.run(["gcc", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return path if os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_suite_under_asan_ubsan(tmp_path):
    asan, ubsan = _runtime("libasan.so"), _runtime("libubsan.so")
    if not asan or not ubsan:
        pytest.skip("sanitizer runtimes not installed")
    lib = os.path.join(str(tmp_path), "liboracle_asan.so")
    run(["gcc", *ORACLE_FLAGS, *SAN, "-fPIC", "-shared", *INC, "-o", lib,
         os.path.join(REPO, "oracle", "qpso_oracle.c"), "-lm"])
    env = dict(os.environ, LD_PRELOAD=f"{asan}:{ubsan}", ASAN_OPTIONS="detect_leaks=0", QPSO_ORACLE_LIB=lib)
    tests = [os.path.join(REPO, "tests", t) for t in ("test_oracle.py", "test_oracle_pso.py", "test_golden.py")]
    p = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider", *tests],
                       capture_output=True, text=True, timeout=900, cwd=REPO, env=env)
    assert p.returncode == 0, p.stdout[-3000:] + p.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in p.stderr and "runtime error" not in p.stderr

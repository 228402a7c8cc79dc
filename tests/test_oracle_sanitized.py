"""The oracle's own tests rerun on an AddressSanitizer + UndefinedBehaviorSanitizer build.

The oracle (oracle/sph_oracle_capi.cpp, g++) is the checker every parity claim rests on, so an
out-of-bounds read or an undefined shift in it could pass unnoticed by value checks.  This test
compiles it with -fsanitize=address,undefined (UB aborts: -fno-sanitize-recover), then runs the
golden-vector, invariant, sink and CD10 oracle tests in a child process that loads that build
(SHAMX_ORACLE_LIB) with the ASan runtime preloaded, and fails on any sanitizer report.  CPU only:
nothing here touches the HIP library or a GPU."""
import os
import shutil
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
SRC = os.path.join(ROOT, "oracle", "sph_oracle_capi.cpp")
SUITES = ["test_oracle_golden.py", "test_oracle_invariants.py", "test_oracle_sinks.py",
          "test_oracle_cd10.py"]


def _asan_runtime(gxx):
    p = subprocess.run([gxx, "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = p.stdout.strip()
    return path if p.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


def test_oracle_under_asan_and_ubsan(tmp_path):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    rt = _asan_runtime(gxx)
    if rt is None:
        pytest.skip("libasan not available")
    lib = str(tmp_path / "libsph_oracle_san.so")
    b = subprocess.run([gxx, "-std=c++20", "-O1", "-g", "-fPIC", "-fopenmp", "-ffp-contract=off",
                        "-fno-fast-math", "-fno-omit-frame-pointer",
                        "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-shared", SRC, "-o", lib], capture_output=True, text=True)
    assert b.returncode == 0, b.stderr[-4000:]
    env = dict(os.environ)
    env.update(SHAMX_ORACLE_LIB=lib, LD_PRELOAD=rt, OMP_NUM_THREADS="4",
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu",
                        "-p", "no:cacheprovider", *[os.path.join(HERE, s) for s in SUITES]],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=900)
    out = r.stdout + r.stderr
    assert "AddressSanitizer" not in out and "runtime error" not in out, out[-6000:]
    assert r.returncode == 0, out[-6000:]
    assert " passed" in r.stdout, out[-2000:]

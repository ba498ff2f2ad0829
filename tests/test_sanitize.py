"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host-side C/C++
(CPU only):

1. tests/sanitize/check_host.cpp linked with the engine's host model
   (j1j2-public_amd/csrc/host_model.cpp) and the oracle's C sources, built
   with -fsanitize=address,undefined -fno-sanitize-recover=all: the host
   model against the oracle, and the oracle's PT/LOCAL/enumeration paths.
2. The oracle's own CPU tests (test_oracle.py, test_pt_histogram.py,
   test_local_protocols.py) rerun in a child process against a sanitizer
   build of the oracle library (libasan preloaded into the interpreter).
"""
import os
import shutil
import subprocess
import sys

import pytest

REPO = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ORACLE = os.path.join(REPO, "oracle")
CSRC = os.path.join(REPO, "j1j2-public_amd", "csrc")
SAN = ["-O1", "-g", "-fsanitize=address,undefined", "-fno-sanitize-recover=all", "-fno-omit-frame-pointer",
       "-ffp-contract=off"]

pytestmark = pytest.mark.skipif(not (shutil.which("gcc") and shutil.which("g++")), reason="needs gcc and g++")


def run(cmd, **kw):
    r = subprocess.run(cmd, capture_output=True, text=True, **kw)
    assert r.returncode == 0, " ".join(cmd) + "\n" + r.stdout[-3000:] + r.stderr[-3000:]
    return r


def test_host_model_and_oracle_under_sanitizers(tmp_path):
    objs = []
    for src in ("j1j2_oracle.c", "corr_field.c"):
        o = str(tmp_path / (src + ".o"))
        run(["gcc", "-std=c11", "-D_GNU_SOURCE", *SAN, "-I", ORACLE, "-c", os.path.join(ORACLE, src), "-o", o])
        objs.append(o)
    exe = str(tmp_path / "check_host")
    run(["g++", "-std=c++17", *SAN, "-I", CSRC, "-I", ORACLE, os.path.join(REPO, "tests", "sanitize", "check_host.cpp"),
         os.path.join(CSRC, "host_model.cpp"), *objs, "-o", exe, "-lm"])
    r = run([exe], env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1"), timeout=300)
    assert "sanitize check: ok" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr


def test_oracle_tests_under_sanitizers(tmp_path):
    lib = str(tmp_path / "libj1j2_oracle_san.so")
    run(["gcc", "-std=c11", "-D_GNU_SOURCE", "-fPIC", "-shared", *SAN, "-I", ORACLE,
         os.path.join(ORACLE, "j1j2_oracle.c"), os.path.join(ORACLE, "corr_field.c"), "-o", lib, "-lm"])
    asan = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True).stdout.strip()
    assert os.path.isabs(asan), "gcc has no libasan"
    env = dict(os.environ, J1J2_ORACLE_LIB=lib, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1")
    tests = [os.path.join(REPO, "tests", t) for t in ("test_oracle.py", "test_pt_histogram.py",
                                                    "test_local_protocols.py")]
    r = run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu", *tests],
            env=env, cwd=REPO, timeout=600)
    assert " passed" in r.stdout and "runtime error" not in r.stdout + r.stderr

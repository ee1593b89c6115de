"""The CPU oracle (oracle/cpu_ref.cpp) and the host code of the C ABI
(argument checks, Julia's promotion, workspace layout) under
AddressSanitizer and UndefinedBehaviorSanitizer.

Both builds are shared libraries loaded into a child Python process with
the sanitizer runtime preloaded (Python itself is not instrumented); any
report aborts the child, so the test sees a non-zero status and the
report."""
import os
import shutil
import subprocess
import sys
from pathlib import Path

import pytest

ROOT = Path(__file__).resolve().parent.parent
CLANG_ASAN = sorted(Path("/opt/rocm/llvm/lib/clang").glob("*/lib/linux/libclang_rt.asan-x86_64.so"))


def _gcc_lib(name):
    return subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True,
                          check=True).stdout.strip()


def _run(cmd, preload, extra_env=None, timeout=600):
    env = dict(os.environ)
    env.update(LD_PRELOAD=" ".join(preload), ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    env.update(extra_env or {})
    r = subprocess.run(cmd, cwd=ROOT, env=env, capture_output=True, text=True, timeout=timeout)
    assert r.returncode == 0, (r.stdout[-3000:], r.stderr[-6000:])
    return r


@pytest.mark.skipif(shutil.which("g++") is None, reason="needs g++")
def test_oracle_under_sanitizers(tmp_path):
    subprocess.run(["make", "-C", str(ROOT / "oracle"), "sanitize", f"SANDIR={tmp_path}"],
                   check=True, capture_output=True)
    r = _run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:xdist", "-m", "not gpu",
              "tests/test_oracle.py", "tests/test_golden.py", "tests/test_tables.py",
              "tests/test_weld_oracle.py"],
             [_gcc_lib("libasan.so"), _gcc_lib("libstdc++.so")],
             {"MSH_ORACLE_LIB": str(tmp_path / "liboracle_san.so")})
    assert " passed" in r.stdout


@pytest.mark.skipif(not CLANG_ASAN or shutil.which("/opt/rocm/bin/hipcc") is None,
                    reason="needs hipcc and its sanitizer runtime")
def test_host_planning_under_sanitizers(tmp_path):
    subprocess.run(["make", "-C", str(ROOT / "meshing.jl_amd"), "sanitize", f"SANDIR={tmp_path}"],
                   check=True, capture_output=True)
    r = _run([sys.executable, str(ROOT / "tests" / "sanitized_host_driver.py"),
              str(tmp_path / "libmeshing_hip_san.so")],
             [str(CLANG_ASAN[-1]), _gcc_lib("libstdc++.so")])
    assert "no reports" in r.stdout

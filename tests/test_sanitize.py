"""AddressSanitizer + UndefinedBehaviorSanitizer builds of the host code (CPU).

tests/sanitize/plan_driver.cpp drives the host setup (within_amd/csrc/plan.cpp):
plan building, the device layout, and the preconditioner deserialiser fed
corrupted bytes (it parses untrusted input).  tests/sanitize/oracle_driver.cpp
drives the oracle through every solver configuration.  Both are built with g++
and -fsanitize=address,undefined (no GPU code) and must finish without a
sanitizer report.
"""

import os
import shutil
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
FLAGS = ["-O1", "-g", "-std=c++17", "-fopenmp", "-fno-omit-frame-pointer", "-fsanitize=address,undefined",
         "-fno-sanitize-recover=undefined", "-ffp-contract=off"]


def build_and_run(tmp_path, name, sources):
    gxx = shutil.which("g++")
    if gxx is None:
        pytest.skip("g++ not available")
    exe = str(tmp_path / name)
    subprocess.run([gxx, *FLAGS, *sources, "-o", exe], check=True, capture_output=True, text=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1",
               OMP_NUM_THREADS="4")
    env.pop("LD_PRELOAD", None)
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_host_setup_under_sanitizers(tmp_path):
    out = build_and_run(tmp_path, "plan_driver", [os.path.join(HERE, "sanitize", "plan_driver.cpp"),
                                                  os.path.join(ROOT, "within_amd", "csrc", "plan.cpp")])
    assert "rejected" in out


def test_oracle_under_sanitizers(tmp_path):
    out = build_and_run(tmp_path, "oracle_driver", [os.path.join(HERE, "sanitize", "oracle_driver.cpp"),
                                                    os.path.join(ROOT, "oracle", "oracle.cpp")])
    assert "converged" in out

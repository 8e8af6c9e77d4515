"""Host code under the address and undefined-behaviour sanitizers: the
oracle (oracle/gtk_oracle.c) and the symbolic phase of the product
(galerkintoolkit.jl_amd/csrc/gtk_plan.cpp, through tests/host/plan_harness.cpp)
are built with -fsanitize=address,undefined (gcc/g++, no GPU code) and
driven by tests/host/sanitize_run.py in a Python process with the sanitizer
runtimes preloaded; any report fails the test."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "galerkintoolkit.jl_amd", "csrc")
SAN = ["-O1", "-g", "-fno-omit-frame-pointer", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined"]


def runtime(name):
    p = subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True).stdout.strip()
    return p if os.path.isabs(p) and os.path.exists(p) else None


@pytest.mark.skipif(runtime("libasan.so") is None or runtime("libubsan.so") is None,
                    reason="no sanitizer runtimes in this toolchain")
def test_oracle_and_planner_under_sanitizers(tmp_path):
    orc = tmp_path / "liboracle_san.so"
    plan = tmp_path / "libplan_san.so"
    subprocess.run(["gcc", *SAN, "-fPIC", "-std=gnu11", "-fopenmp", "-shared", "-o", str(orc),
                    os.path.join(ROOT, "oracle", "gtk_oracle.c"), "-lm"], check=True)
    subprocess.run(["g++", *SAN, "-fPIC", "-std=c++17", "-shared", "-I", CSRC, "-o", str(plan),
                    os.path.join(ROOT, "tests", "host", "plan_harness.cpp"), os.path.join(CSRC, "gtk_plan.cpp"),
                    "-lpthread"], check=True)
    env = dict(os.environ)
    env["LD_PRELOAD"] = runtime("libasan.so") + ":" + runtime("libubsan.so")
    env["ASAN_OPTIONS"] = "detect_leaks=0:abort_on_error=0:halt_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    env["GTK_ORACLE_LIB"] = str(orc)
    env["OMP_NUM_THREADS"] = "2"
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "host", "sanitize_run.py"), str(plan)],
                       capture_output=True, text=True, env=env, timeout=600)
    out = r.stdout + r.stderr
    assert r.returncode == 0, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]
    assert "sanitized run ok" in r.stdout

"""CPU sanitizer runs (AddressSanitizer + UBSan) of the host code.

* The C ABI's host-side argument paths: every library source is rebuilt with its host code
  instrumented (janus-sim_amd/Makefile `asan`) and linked into tests/abi_host_driver.c, which calls
  the entry points with invalid and null arguments, a create without a device, and bad snapshot files.
* The oracle: oracle/jtp_oracle.cpp rebuilt instrumented (oracle/Makefile `asan`) and the oracle's
  own tests (tests/test_oracle.py) re-run against that build in a child interpreter with the GCC
  sanitizer runtimes preloaded."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _gcc_runtime(name):
    return subprocess.run(["g++", f"-print-file-name={name}"], capture_output=True, text=True,
                          check=True).stdout.strip()


def test_abi_host_paths_under_asan(tmp_path):
    jobs = str(min(8, os.cpu_count() or 1))
    subprocess.run(["make", "-C", os.path.join(ROOT, "janus-sim_amd"), "-j", jobs, "asan"], check=True,
                   capture_output=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ROOT, "janus-sim_amd", "_build", "asan", "abi_host_driver"), str(tmp_path)],
                       capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout[-2000:], r.stderr[-4000:])
    assert r.returncode == 0
    assert "abi_host_driver: ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr


def test_oracle_under_asan():
    subprocess.run(["make", "-C", os.path.join(ROOT, "oracle"), "asan"], check=True, capture_output=True)
    pre = ":".join(_gcc_runtime(n) for n in ("libasan.so", "libubsan.so"))
    if not all(os.path.exists(p) for p in pre.split(":")):
        pytest.skip("GCC sanitizer runtimes not installed")
    env = dict(os.environ, LD_PRELOAD=pre, ASAN_OPTIONS="detect_leaks=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1",
               JTP_ORACLE_LIB=os.path.join(ROOT, "oracle", "liboracle_asan.so"))
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_oracle.py")],
                       capture_output=True, text=True, env=env, cwd=ROOT, timeout=600)
    print(r.stdout[-3000:], r.stderr[-3000:])
    assert r.returncode == 0
    assert "ERROR: AddressSanitizer" not in r.stdout + r.stderr and "runtime error" not in r.stdout + r.stderr

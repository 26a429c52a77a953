"""The C ABI under AddressSanitizer + UndefinedBehaviorSanitizer (host code).

tests/sanitize/Makefile builds a host-sanitized copy of the library and a
C++ driver (tests/sanitize/abi_driver.cpp) that calls it; __graft_entry__
.build() runs that build.  The driver overrides xerbla_ (the reference's
user-override contract, driver/others/xerbla.c) and checks the reported
names and numbers; with a device it also runs real host- and device-operand
calls and four concurrent host threads.  Any sanitizer report aborts the
driver, so a zero exit and "OK" mean a clean run.
"""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
DRIVER = os.path.join(ROOT, "tests", "sanitize", "_out", "abi_driver")


def _run(mode, timeout):
    if not os.path.exists(DRIVER):
        pytest.fail(f"{DRIVER} is missing: build it with `python -c 'import __graft_entry__ as g; g.build()'`")
    env = dict(os.environ)
    # the HIP runtime keeps allocations until exit (no leak check); the
    # sanitizer runtime need not be first in the preload list
    env["ASAN_OPTIONS"] = "detect_leaks=0:verify_asan_link_order=0:abort_on_error=1"
    env["UBSAN_OPTIONS"] = "print_stacktrace=1:halt_on_error=1"
    p = subprocess.run([DRIVER, mode], capture_output=True, text=True, timeout=timeout, env=env)
    out = p.stdout + p.stderr
    assert p.returncode == 0 and "OK" in p.stdout, out[-4000:]
    assert "ERROR: AddressSanitizer" not in out and "runtime error:" not in out, out[-4000:]


def test_sanitized_abi_cpu():
    _run("cpu", 120)


@pytest.mark.gpu
def test_sanitized_abi_gpu():
    _run("gpu", 600)

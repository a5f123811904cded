"""The kernel templates (Rys rules, quartet and derivative integrals of
eri_rys.h / grad_rys.h, instantiated for the host by tests/csrc/eri_check.cpp)
run under AddressSanitizer and UndefinedBehaviorSanitizer: every class
combination in FP64 and FP32, and the rules over the edges of their tables
(tests/csrc/san_check.cpp). Any out-of-bounds read or undefined behaviour
aborts the program with a report."""
import os
import subprocess

from tests import kernel_check as KC


def test_kernel_templates_under_asan_and_ubsan():
    srcs = [KC.SRC, KC.SAN_SRC]
    if not os.path.exists(KC.SAN_EXE) or any(os.path.getmtime(KC.SAN_EXE) < os.path.getmtime(s) for s in srcs):
        KC.build_sanitized()
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([KC.SAN_EXE], capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "san_check: ok" in r.stdout

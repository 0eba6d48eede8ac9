"""The C oracle under AddressSanitizer and UndefinedBehaviorSanitizer: every
density string (1-5 operators), the excitation lists and the contractions
on small fragments with and without a frozen core run without a sanitizer
report, and give the checksum of the plain build (oracle/sanitize_main.c)."""
import os
import subprocess

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
ORACLE = os.path.join(os.path.dirname(HERE), "oracle")


def test_oracle_is_clean_under_sanitizers():
    r = subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], capture_output=True, text=True)
    if r.returncode != 0 and "sanitize" in (r.stderr + r.stdout):
        pytest.skip("compiler without address / undefined sanitizers: " + r.stderr[-300:])
    assert r.returncode == 0, r.stderr
    plain = subprocess.run([os.path.join(ORACLE, "_san", "plain")], capture_output=True, text=True, timeout=300)
    asan = subprocess.run([os.path.join(ORACLE, "_san", "asan")], capture_output=True, text=True, timeout=300,
                          env=dict(os.environ, ASAN_OPTIONS="detect_leaks=1", UBSAN_OPTIONS="print_stacktrace=1"))
    assert plain.returncode == 0, plain.stderr
    assert asan.returncode == 0, asan.stderr[-3000:]
    assert "runtime error" not in asan.stderr and "AddressSanitizer" not in asan.stderr, asan.stderr[-3000:]
    a, b = float(plain.stdout.split()[-1]), float(asan.stdout.split()[-1])
    assert abs(a - b) <= 1e-9 * max(1.0, abs(a)), (a, b)

"""The CPU oracle under AddressSanitizer + UndefinedBehaviorSanitizer and
under ThreadSanitizer: oracle/sanitize_main.cpp drives every operator of the
oracle (including the per-thread aggregate states and their locked merges,
join builds with duplicate and NULL keys, sorts, MLB1) on seeded inputs; each
build must finish with exit status 0 and no sanitizer report."""

import os
import subprocess

import pytest

ORACLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "oracle")


@pytest.fixture(scope="module")
def sanitized():
    subprocess.run(["make", "-s", "-C", ORACLE, "sanitize"], check=True)
    return os.path.join(ORACLE, "_san")


@pytest.mark.parametrize("binary", ["oracle_asan_ubsan", "oracle_tsan"])
def test_oracle_sanitizer_clean(sanitized, binary):
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", TSAN_OPTIONS="halt_on_error=1")
    r = subprocess.run([os.path.join(sanitized, binary)], capture_output=True, text=True, env=env, timeout=600)
    report = r.stderr
    assert r.returncode == 0, report[-4000:]
    assert "all checks passed" in r.stdout
    for marker in ("AddressSanitizer", "LeakSanitizer", "ThreadSanitizer", "runtime error"):
        assert marker not in report, report[-4000:]

"""The host-side C and C++ code under AddressSanitizer and UndefinedBehaviorSanitizer.

tests/sanitize/Makefile builds the product's serial path (csrc/ppa_cpu.cpp)
and the oracle (oracle/ppa_oracle.c) with -fsanitize=address,undefined (host
compilers only) and links them into check_cpu, which runs both over small
graphs of both models and the edge cases; any sanitizer report fails it."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
SAN = os.path.join(ROOT, "tests", "sanitize")


def test_host_code_clean_under_asan_and_ubsan():
    subprocess.run(["make", "-B", "-C", SAN], check=True, capture_output=True, text=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0", UBSAN_OPTIONS="print_stacktrace=1")
    r = subprocess.run([os.path.join(SAN, "_build", "check_cpu")], capture_output=True, text=True, timeout=600,
                       env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr
    assert "sanitized host checks passed" in r.stdout

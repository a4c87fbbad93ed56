"""The host code and the CPU oracle under AddressSanitizer, LeakSanitizer and
UndefinedBehaviorSanitizer (CPU only).

tests/san/Makefile compiles the library's host sources (file readers and
writers, METIS text, conversion, generators, the initial-partition seams) and
the oracle with -fsanitize=address,undefined into tests/san/host_checks, which
drives them on small graphs and compares their outputs with the oracle's; any
memory error, leak or undefined behaviour aborts it.
"""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def test_host_code_and_oracle_under_sanitizers(tmp_path):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "tests", "san")], check=True,
                   timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", OMP_NUM_THREADS="4")
    r = subprocess.run([os.path.join(ROOT, "tests", "san", "_build", "host_checks"),
                        str(tmp_path), os.path.join(ROOT, "tests", "golden", "stub_bin")],
                       capture_output=True, text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stderr[-4000:]
    assert "host checks passed" in r.stdout

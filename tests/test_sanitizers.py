"""AddressSanitizer / UndefinedBehaviorSanitizer builds of the host code:
the oracle (oracle/scmp_oracle.c, every step order, walls, threads and the
restart path) and the driver's host side (input.dat parser, VTK writers,
model factory).  Each is built by its Makefile's `sanitize` target (gcc /
g++, host code only) and must run clean; its own checks (thread invariance,
mass conservation, parser results, VTK headers) must hold."""
import os
import subprocess

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))


def _build_and_run(folder, exe):
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, folder), "sanitize"], check=True,
                   timeout=300)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([os.path.join(ROOT, folder, "_san", exe)], capture_output=True, text=True,
                       timeout=300, env=env)
    assert r.returncode == 0, r.stdout + r.stderr
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr, r.stderr
    return r.stdout


def test_oracle_under_sanitizers():
    assert "ok" in _build_and_run("oracle", "sanitize_check")


def test_host_driver_code_under_sanitizers():
    assert "ok" in _build_and_run(os.path.join("mesobasic_amd", "host"), "sanitize_check")

"""The CPU oracle and the library's host C++ (CSV import, bin pre-sort)
under AddressSanitizer + UndefinedBehaviorSanitizer (host code only; gcc):
tests/sanitize/sanitize_main.cpp drives every entry point on random,
degenerate and malformed inputs."""
import os
import subprocess

HERE = os.path.join(os.path.dirname(os.path.abspath(__file__)), "sanitize")


def test_host_code_is_clean_under_asan_ubsan(tmp_path):
    out = tmp_path / "build"
    subprocess.run(["make", "-s", "-C", HERE, f"OUT={out}"], check=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:verify_asan_link_order=0",
               UBSAN_OPTIONS="print_stacktrace=1", OMP_NUM_THREADS="2")
    r = subprocess.run([str(out / "sanitize_host")], capture_output=True, text=True, cwd=str(tmp_path),
                       env=env, timeout=300)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]
    assert "sanitize_host: clean" in r.stdout

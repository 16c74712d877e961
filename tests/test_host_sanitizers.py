"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer, on the CPU.

The reference builds a sanitizer configuration (CMakeLists.txt:13,21-25). GPU sanitizers are not
available here, so this covers the host code that runs without a GPU: every scene's build and BVH,
the CPU arm, the `wf` CLI's host helpers and the oracle, through tests/sanitize_driver.cpp, built
with -fsanitize=address,undefined (`make -C graphics-book-tutorials_amd sanitize`). Any sanitizer
report or failed check ends the driver with a non-zero status.
"""
import os
import subprocess

from conftest import PKG


def test_host_code_is_clean_under_asan_and_ubsan(tmp_path):
    subprocess.run(["make", "-s", "sanitize"], cwd=PKG, check=True, capture_output=True, timeout=600)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([os.path.join(PKG, "_build", "sanitize", "driver"), str(tmp_path)], capture_output=True,
                       text=True, env=env, timeout=600)
    assert r.returncode == 0, r.stdout[-2000:] + r.stderr[-4000:]
    assert "ok (0 failures)" in r.stdout and "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr

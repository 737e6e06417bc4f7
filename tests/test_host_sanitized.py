"""The host-side tests (test_host.py) again, against the library's host code built with
AddressSanitizer and UndefinedBehaviorSanitizer (make -C cm_decrypt__amd asan): an invalid
memory access, a leak of a use-after-free or undefined behaviour in key handling, the
generator, Benes routing or the errorstream formatters fails this test.  CPU only; skipped
when the sanitizer build is absent."""
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "cm_decrypt__amd", "libcmdecrypt_asan.so")


def libasan():
    r = subprocess.run(["gcc", "-print-file-name=libasan.so"], capture_output=True, text=True)
    path = r.stdout.strip()
    return path if r.returncode == 0 and os.path.isabs(path) and os.path.exists(path) else None


@pytest.mark.skipif(not os.path.exists(ASAN_LIB), reason="sanitizer build absent (make -C cm_decrypt__amd asan)")
def test_host_suite_under_sanitizers():
    asan = libasan()
    if asan is None:
        pytest.skip("libasan not found")
    env = dict(os.environ, LD_PRELOAD=asan, ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1", CMD_DECRYPT_LIB=ASAN_LIB,
               CMD_DECRYPT_NO_TORCH="1")
    # the instrumented build is the one the tests load, and it calls the sanitizer runtimes
    probe = subprocess.run([sys.executable, "-c", "import cm_decrypt__amd as cm; print(cm.LIB_PATH)"], cwd=ROOT,
                           env=env, capture_output=True, text=True, timeout=300)
    assert probe.returncode == 0 and probe.stdout.strip() == ASAN_LIB, probe.stdout + probe.stderr
    syms = subprocess.run(["nm", "-D", ASAN_LIB], capture_output=True, text=True).stdout
    assert "__asan_report" in syms and "__ubsan_handle" in syms
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-p", "no:cacheprovider", "-m", "not gpu",
                        os.path.join("tests", "test_host.py")], cwd=ROOT, env=env, capture_output=True, text=True,
                       timeout=1200)
    out = r.stdout[-4000:] + r.stderr[-4000:]
    assert r.returncode == 0, out
    assert "AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, out
    assert " passed" in r.stdout, out

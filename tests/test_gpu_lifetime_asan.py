"""Object lifetimes of the C ABI with the library's host code under
AddressSanitizer and UBSan (tests/cpp/lifetime_check.cpp, built with every
-fsanitize= after -Xarch_host so the device code is unchanged): 1D contexts
whose workspaces grow, a captured multi-field graph that the growth makes
stale (refused) and that is created again, the allocation-free entry point,
a forced split and the autotuner, a two-band context whose workspace grows,
and a sampler with a surrogate network trained, used for delayed acceptance
and the surrogate transition, then destroyed before the network.  Leak
reports are off: the HIP runtime keeps allocations until the process
ends."""
import os
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
EXE = os.path.join(ROOT, "tests", "cpp", "_build", "lifetime_check_asan")

pytestmark = pytest.mark.gpu


def test_lifetimes_under_asan():
    assert os.path.exists(EXE), "build tests/cpp asan first (__graft_entry__.build)"
    print("[gpu-asan-case] lifetime_check", flush=True)
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:halt_on_error=1",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    out = subprocess.run([EXE], capture_output=True, text=True, timeout=240,
                         env=env)
    text = out.stdout + out.stderr
    assert out.returncode == 0, text[-4000:]
    assert "AddressSanitizer" not in text and "runtime error:" not in text, \
        text[-4000:]
    assert "stale graph refused" in out.stdout
    assert "lifetime_check ok" in out.stdout

"""The host runtime of the C ABI under AddressSanitizer and UBSan: builds
pico_amd/_build/asan/libpico_gpu.so (make -C pico_amd asan; host code
instrumented, kernels as usual) and runs tests/asan_host_paths.py against
it with the sanitizer runtime preloaded.  CPU only."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
RUNTIME = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))


@pytest.mark.skipif(not RUNTIME, reason="no ASan runtime in this ROCm installation")
def test_host_paths_under_asan_and_ubsan():
    subprocess.run(["make", "-s", "-C", os.path.join(ROOT, "pico_amd"), "asan"], check=True, timeout=900)
    lib = os.path.join(ROOT, "pico_amd", "_build", "asan", "libpico_gpu.so")
    env = dict(os.environ, LD_PRELOAD=RUNTIME[-1], ASAN_OPTIONS="detect_leaks=0:abort_on_error=0",
               UBSAN_OPTIONS="print_stacktrace=1:halt_on_error=1")
    r = subprocess.run([sys.executable, os.path.join(ROOT, "tests", "asan_host_paths.py"), lib], env=env,
                       capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "host paths ok" in r.stdout
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]

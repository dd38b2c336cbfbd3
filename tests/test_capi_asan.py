"""The C-ABI layer's host code under AddressSanitizer and UBSan: the CPU
tests of the C ABI (argument checks, configuration refusals, creation
failing without a device and releasing what it built) run against the
host-sanitizer build of the library (`make -C guacho_amd asan`: gm_capi and
gm_batch instrumented, the kernels unchanged), with the clang ASan runtime
preloaded; any report aborts the run."""
import glob
import os
import subprocess
import sys

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
ASAN_LIB = os.path.join(ROOT, "guacho_amd", "_asan", "libguacho_amd.so")
ASAN_RT = sorted(glob.glob("/opt/rocm/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))

CHECK = """
import ctypes, os
from guacho_amd import _capi
_capi.lib()
maps = open('/proc/self/maps').read()
assert os.environ['GUACHO_AMD_LIB'] in maps, 'sanitizer build not loaded'
assert 'libclang_rt.asan' in maps, 'ASan runtime not loaded'
print('sanitizer build loaded')
"""


@pytest.mark.skipif(not os.path.exists(ASAN_LIB) or not ASAN_RT, reason="host-sanitizer build not present")
def test_capi_cpu_tests_under_asan_and_ubsan():
    env = dict(os.environ, GUACHO_AMD_LIB=ASAN_LIB, LD_PRELOAD=ASAN_RT[-1],
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1")
    r = subprocess.run([sys.executable, "-c", CHECK], cwd=ROOT, env=env, capture_output=True, text=True, timeout=300)
    assert r.returncode == 0 and "sanitizer build loaded" in r.stdout, r.stdout + r.stderr
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider",
                        os.path.join(ROOT, "tests", "test_capi_cpu.py")],
                       cwd=ROOT, env=env, capture_output=True, text=True, timeout=600)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error" not in r.stderr, r.stderr[-3000:]

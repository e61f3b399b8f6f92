"""Host code under AddressSanitizer and UndefinedBehaviorSanitizer (CPU only).

build() also makes a variant of the library, the transfer server and the
transfer client whose host-side code (API layers, parameter validation, RF
state, DAS constants, active-tap count, batch set-up, synthetic generator,
wire protocol) is instrumented with both sanitizers (GPU code is not: each
-fsanitize= applies to the host pass only).  The CPU suite's host-side tests
run again against it in a child process with the sanitizer runtime preloaded;
any report aborts that process and fails this test.
"""
import os
import subprocess
import sys
from pathlib import Path

import pytest

from cuda_ultrasound_toolkit_amd import build as B

ROOT = Path(__file__).resolve().parent.parent
SAN = B.SAN_DIR
HOST_TESTS = ["tests/test_abi.py", "tests/test_active_taps.py", "tests/test_host.py", "tests/test_golden.py",
              "tests/test_transfer.py"]


def _env():
    rt = B.sanitizer_runtime()
    if rt is None or not (SAN / "libcuda_toolkit.so").exists():
        pytest.skip("host-sanitizer build missing (build() makes it when the clang sanitizer runtime exists)")
    env = dict(os.environ)
    env.update(LD_PRELOAD=str(rt), ASAN_OPTIONS="detect_leaks=0:halt_on_error=1:abort_on_error=0",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", CUT_LIBRARY=str(SAN / "libcuda_toolkit.so"),
               CUT_TRANSFER_SERVER=str(SAN / "cut_transfer_server"),
               CUT_TRANSFER_CLIENT=str(SAN / "libcut_transfer.so"), PYTHONDONTWRITEBYTECODE="1")
    return env


def test_sanitized_library_is_the_one_loaded():
    code = ("from cuda_ultrasound_toolkit_amd import _lib as L; L.load(); "
            "m = open('/proc/self/maps').read(); "
            "assert '_build_san/libcuda_toolkit.so' in m and 'libclang_rt.asan' in m; print('ok')")
    r = subprocess.run([sys.executable, "-c", code], cwd=ROOT, env=_env(), capture_output=True, text=True,
                       timeout=300)
    assert r.returncode == 0 and "ok" in r.stdout, r.stderr[-3000:]


def test_host_tests_clean_under_sanitizers():
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-x", "-m", "not gpu", "-p", "no:cacheprovider",
                        *HOST_TESTS], cwd=ROOT, env=_env(), capture_output=True, text=True, timeout=600)
    tail = (r.stdout[-2000:] + r.stderr[-4000:])
    assert r.returncode == 0, tail
    assert "ERROR: AddressSanitizer" not in r.stderr and "runtime error:" not in r.stderr, tail

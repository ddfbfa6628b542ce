"""The C ABI's argument checks under AddressSanitizer (CPU only): the
library's host code built with -fsanitize=address (tests/asan_build.py)
runs every test of tests/test_abi.py in a child process with the sanitizer
runtime preloaded; an invalid read or write in a check, or in the status
strings, aborts the child with a report."""
import os
import subprocess
import sys

import pytest

HERE = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, HERE)
import asan_build  # noqa: E402


def _gpu_visible():
    try:
        import torch
        return torch.cuda.is_available()
    except Exception:
        return False


@pytest.mark.skipif(_gpu_visible(), reason="the ABI checks pass made-up device pointers; CPU-only")
@pytest.mark.skipif(not os.path.exists(asan_build.HIPCC), reason="hipcc not installed")
def test_abi_checks_under_address_sanitizer():
    lib = asan_build.build()
    rt = asan_build.asan_runtime()
    if rt is None:
        pytest.skip("no clang AddressSanitizer runtime")
    env = dict(os.environ, FFIELDS_LIB=lib, LD_PRELOAD=rt,
               ASAN_OPTIONS="detect_leaks=0:abort_on_error=1:detect_odr_violation=0:alloc_dealloc_mismatch=0")
    r = subprocess.run([sys.executable, "-m", "pytest", "-q", "-p", "no:cacheprovider", os.path.join(HERE, "test_abi.py")],
                       env=env, capture_output=True, text=True, cwd=os.path.dirname(HERE), timeout=900)
    assert r.returncode == 0, r.stdout[-3000:] + r.stderr[-3000:]
    assert "AddressSanitizer" not in r.stderr
    assert " passed" in r.stdout

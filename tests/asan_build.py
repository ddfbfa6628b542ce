"""AddressSanitizer build of libffields.so's host code, for the CPU checks of
the C ABI's argument validation (tests/test_abi_asan.py).

Host code is compiled with AddressSanitizer (`-Xarch_host` before each
sanitizer flag, so the GPU code is not instrumented); the GPU code is
compiled without optimisation to keep the build short -- it is never run
from this library.  Output: factor-fields_amd/_build/asan/libffields_asan.so.
Load it with the clang AddressSanitizer runtime preloaded (see asan_runtime).

    python tests/asan_build.py
"""
import concurrent.futures
import glob
import os
import subprocess
import sys

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "factor-fields_amd", "csrc")
OUT = os.path.join(ROOT, "factor-fields_amd", "_build", "asan")
LIB = os.path.join(OUT, "libffields_asan.so")
HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]
FLAGS = ["--offload-arch=gfx950", "-O1", "-Xarch_device", "-O0", "-g", "-std=c++17", "-fPIC", "-ffp-contract=off",
         "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops", "-Wno-unused-result",
         "-I", os.path.join(ROOT, "include"), "-I", CSRC] + SAN


def asan_runtime():
    """The clang AddressSanitizer runtime that hipcc's instrumentation needs."""
    out = subprocess.run([HIPCC, "--print-file-name=libclang_rt.asan-x86_64.so"], capture_output=True, text=True)
    p = out.stdout.strip()
    if p and os.path.isabs(p) and os.path.exists(p):
        return p
    hits = sorted(glob.glob("/opt/rocm/lib/llvm/lib/clang/*/lib/linux/libclang_rt.asan-x86_64.so"))
    return hits[-1] if hits else None


def _compile(src):
    obj = os.path.join(OUT, os.path.basename(src)[:-4] + ".o")
    if os.path.exists(obj) and os.path.getmtime(obj) > max(
            os.path.getmtime(f) for f in glob.glob(os.path.join(CSRC, "*")) + [os.path.abspath(__file__)]):
        return obj
    r = subprocess.run([HIPCC] + FLAGS + ["-c", src, "-o", obj], capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"hipcc failed for {src}:\n{r.stderr[-4000:]}")
    return obj


def build():
    os.makedirs(OUT, exist_ok=True)
    srcs = sorted(glob.glob(os.path.join(CSRC, "*.hip")))
    with concurrent.futures.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(_compile, srcs))
    r = subprocess.run([HIPCC, "--offload-arch=gfx950", "-shared", "-fPIC"] + SAN[:2] + ["-o", LIB] + objs,
                       capture_output=True, text=True)
    if r.returncode:
        raise RuntimeError(f"link failed:\n{r.stderr[-4000:]}")
    return LIB


if __name__ == "__main__":
    print(build())
    print(asan_runtime())
    sys.exit(0)

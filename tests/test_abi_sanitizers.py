"""The C-ABI's host code under AddressSanitizer and UndefinedBehaviorSanitizer
(CPU only): every library source is compiled with both sanitizers on its host
pass (the device pass as usual, at -O1; no kernel runs), linked into a shared
library, and the C
driver tests/abi_host_check.c -- sizing functions, the host tile packer over
synthetic records, the argument checks of the device entry points -- runs
against it with the sanitizer runtimes and must print ALL OK.  Leak detection is
off: the HIP runtime keeps allocations until exit."""
import concurrent.futures as cf
import os
import shutil
import subprocess

import pytest

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "jax-dips_amd", "csrc")
HIPCC = "/opt/rocm/bin/hipcc"
CLANG = "/opt/rocm/lib/llvm/bin/clang"
# the sanitizer flags apply to the host compilation only (each directly after -Xarch_host)
SAN = ["-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fsanitize=undefined",
       "-Xarch_host", "-fno-sanitize-recover=undefined"]


@pytest.mark.skipif(not (os.path.exists(HIPCC) and os.path.exists(CLANG)), reason="hipcc / clang not installed")
def test_host_abi_under_asan_and_ubsan(tmp_path):
    srcs = sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))

    def compile_one(src):
        obj = str(tmp_path / (os.path.basename(src) + ".o"))
        cmd = [HIPCC, "-O1", "-g", "-fno-omit-frame-pointer", "-std=c++17", "-fPIC",
               "--offload-arch=gfx950", *SAN, "-Wno-unused-command-line-argument",
               "-I", os.path.join(ROOT, "include"), "-c", src, "-o", obj]
        r = subprocess.run(cmd, capture_output=True, text=True)
        assert r.returncode == 0, r.stderr[-3000:]
        return obj

    with cf.ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as ex:
        objs = list(ex.map(compile_one, srcs))
    lib = str(tmp_path / "libdips_host.so")
    r = subprocess.run([HIPCC, "-shared", "-fPIC", "--offload-arch=gfx950", *SAN, *objs, "-L/opt/rocm/lib",
                        "-Wl,-rpath,/opt/rocm/lib", "-lrocprofiler-sdk-roctx", "-o", lib],
                       capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    exe = str(tmp_path / "abi_host_check")
    # the driver is plain C on a line of its own, sanitized with the same clang (which links the runtimes)
    r = subprocess.run([CLANG, "-g", "-O1", "-fsanitize=address,undefined", "-fno-sanitize-recover=undefined",
                        "-I", os.path.join(ROOT, "include"), os.path.join(ROOT, "tests", "abi_host_check.c"),
                        lib, f"-Wl,-rpath,{tmp_path}", "-o", exe], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-3000:]
    env = dict(os.environ, ASAN_OPTIONS="detect_leaks=0:abort_on_error=0:halt_on_error=1",
               UBSAN_OPTIONS="halt_on_error=1:print_stacktrace=1", HIP_VISIBLE_DEVICES="")
    r = subprocess.run([exe], capture_output=True, text=True, env=env, timeout=300)
    print(r.stdout, r.stderr[-4000:])
    assert r.returncode == 0 and "ALL OK" in r.stdout
    assert "runtime error" not in r.stderr and "AddressSanitizer" not in r.stderr

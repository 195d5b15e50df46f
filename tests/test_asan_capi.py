"""AddressSanitizer / LeakSanitizer run of the C-ABI's host code.

tests/cpp/asan_capi.cpp is linked against the library variant whose host code
is instrumented (physgrad_amd/_asan/libphysgrad_mpm.so, built by
physgrad_amd.build.build_asan).  Without a GPU it exercises argument validation
and the clean-up of contexts whose device allocations fail; on a GPU it drives
every entry-point family once on small scenes, including the autotune cache
parser over a file with malformed lines.  Any heap error or leak in the
library's host code fails the run.
"""
import os
import subprocess

import pytest

from conftest import ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "asan_capi.cpp")
CLANG = "/opt/rocm/llvm/bin/clang++"


def _asan_lib(allow_build):
    from physgrad_amd import build as pgbuild
    if pgbuild.asan_stale():
        if not allow_build:
            pytest.skip("instrumented library not built (python -m physgrad_amd.build asan)")
        pgbuild.build_asan(verbose=False)
    return pgbuild.ASAN_LIB


def _driver(tmp_path, allow_build):
    lib = _asan_lib(allow_build)
    exe = os.path.join(str(tmp_path), "asan_capi")
    libdir = os.path.dirname(lib)
    cmd = [CLANG, "-std=c++17", "-g", "-fsanitize=address", "-fno-omit-frame-pointer", "-D__HIP_PLATFORM_AMD__",
           "-I/opt/rocm/include", f"-I{os.path.join(ROOT, 'include')}", SRC, lib, "-L/opt/rocm/lib", "-lamdhip64",
           f"-Wl,-rpath,{libdir}:/opt/rocm/lib", "-o", exe]
    r = subprocess.run(cmd, capture_output=True, text=True)
    assert r.returncode == 0, r.stderr
    return exe


def _run(exe, tmp_path):
    cache = os.path.join(str(tmp_path), "autotune.txt")
    with open(cache, "w") as fh:
        fh.write("x" * 1000 + " 64 1280 256\n")                  # longer than the parser's line buffer
        fh.write("\n   \nslots=1;blocks=1;n=1;scenes=1\n")        # blank lines, key without values
        fh.write("slots=1;blocks=1;n=1;scenes=1 7 99999 300\n")   # values out of range
        fh.write("%s %d\n")                                      # format characters
    supp = os.path.join(str(tmp_path), "lsan.supp")
    with open(supp, "w") as fh:
        # allocations the HIP/HSA runtimes keep until exit are theirs, not the library's
        fh.write("leak:libamdhip64\nleak:libhsa-runtime64\nleak:libhsakmt\nleak:librocprofiler\n")
    env = dict(os.environ, PG_AUTOTUNE_CACHE=cache,
               ASAN_OPTIONS="detect_leaks=1:abort_on_error=0:halt_on_error=1:exitcode=23",
               LSAN_OPTIONS=f"suppressions={supp}")
    r = subprocess.run([exe], capture_output=True, text=True, timeout=110, env=env)
    report = r.stdout + r.stderr
    assert "AddressSanitizer" not in report and "LeakSanitizer" not in report, report[-6000:]
    assert r.returncode == 0, report[-6000:]
    assert "asan driver OK" in r.stdout, report[-6000:]
    return r.stdout


def test_asan_capi_host(tmp_path):
    """Argument checks and failed-create clean-up; no device calls succeed here."""
    import torch
    if torch.cuda.is_available():
        pytest.skip("the GPU variant runs the device workload")
    out = _run(_driver(tmp_path, allow_build=True), tmp_path)
    assert "no device" in out


@pytest.mark.gpu
def test_asan_capi_gpu(gpu, tmp_path):
    print("asan driver: device workload", flush=True)
    out = _run(_driver(tmp_path, allow_build=False), tmp_path)
    assert "device workload done" in out

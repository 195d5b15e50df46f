"""Build the HIP library libphysgrad_mpm.so for gfx950 (in-tree).

Every .hip source under csrc/ is compiled by hipcc to an object, in parallel,
and linked into physgrad_amd/libphysgrad_mpm.so.  The library has a plain
C-ABI (include/physgrad_mpm.h) and no PyTorch dependency.

build_asan() builds the same sources into physgrad_amd/_asan/libphysgrad_mpm.so
with the host code instrumented by AddressSanitizer (-Xarch_host keeps the
device code as it is).  An executable linked
against it must itself be linked with -fsanitize=address, which brings the
sanitizer runtime; tests/cpp/asan_capi.cpp is such a driver.

    python -m physgrad_amd.build          # or physgrad_amd.build.build()
    python -m physgrad_amd.build asan     # the instrumented variant as well
"""
from __future__ import annotations

import concurrent.futures as cf
import os
import shutil
import subprocess
import sys

HERE = os.path.dirname(os.path.abspath(__file__))
ROOT = os.path.dirname(HERE)
CSRC = os.path.join(HERE, "csrc")
OBJ = os.path.join(HERE, "_build")
LIB = os.path.join(HERE, "libphysgrad_mpm.so")
ASAN_DIR = os.path.join(HERE, "_asan")
ASAN_LIB = os.path.join(ASAN_DIR, "libphysgrad_mpm.so")

HIPCC = os.environ.get("HIPCC", "/opt/rocm/bin/hipcc")
ARCH = os.environ.get("PG_OFFLOAD_ARCH", "gfx950")

# -packed-fp32-ops: no v_pk_{add,mul,fma}_f32 from the compiler (hipcc flag).
HIP_FLAGS = [
    f"--offload-arch={ARCH}", "-O3", "-fPIC", "-std=c++17",
    "-Xclang", "-target-feature", "-Xclang", "-packed-fp32-ops",
    "-Wno-unused-command-line-argument",
    f"-I{os.path.join(ROOT, 'include')}", f"-I{CSRC}",
]


def sources():
    return sorted(os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".hip"))


# host code only: each sanitizer flag goes behind -Xarch_host
ASAN_FLAGS = ["-Xarch_host", "-g", "-Xarch_host", "-fsanitize=address", "-Xarch_host", "-fno-omit-frame-pointer"]


def _compile(src: str, objdir: str = OBJ, extra=()) -> str:
    obj = os.path.join(objdir, os.path.basename(src) + ".o")
    cmd = [HIPCC] + HIP_FLAGS + list(extra) + ["-c", src, "-o", obj]
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"hipcc failed on {src}:\n{r.stdout}\n{r.stderr}")
    return obj


def _build(lib: str, objdir: str, extra=(), link_extra=(), verbose: bool = True) -> str:
    if not os.path.exists(HIPCC):
        raise RuntimeError(f"hipcc not found at {HIPCC}")
    os.makedirs(objdir, exist_ok=True)
    srcs = sources()
    jobs = min(len(srcs), int(os.environ.get("MAX_JOBS", "8")), 16)
    with cf.ThreadPoolExecutor(max_workers=max(jobs, 1)) as ex:
        objs = list(ex.map(lambda s: _compile(s, objdir, extra), srcs))
    tmp = lib + ".tmp"
    cmd = [HIPCC, f"--offload-arch={ARCH}", "-shared", "-fPIC"] + list(link_extra) + ["-o", tmp] + objs
    r = subprocess.run(cmd, capture_output=True, text=True)
    if r.returncode != 0:
        raise RuntimeError(f"link failed:\n{r.stdout}\n{r.stderr}")
    shutil.move(tmp, lib)
    if verbose:
        print(f"built {lib} from {len(srcs)} sources", file=sys.stderr)
    return lib


def build(verbose: bool = True) -> str:
    """Compile all HIP sources and link the shared library; returns its path."""
    return _build(LIB, OBJ, verbose=verbose)


def build_asan(verbose: bool = True) -> str:
    """The library with AddressSanitizer on its host code; returns its path."""
    return _build(ASAN_LIB, os.path.join(ASAN_DIR, "obj"), ASAN_FLAGS, ["-Xarch_host", "-fsanitize=address"],
                  verbose=verbose)


def asan_stale() -> bool:
    """True if the instrumented library is missing or older than a source."""
    if not os.path.exists(ASAN_LIB):
        return True
    t = os.path.getmtime(ASAN_LIB)
    deps = sources() + [os.path.join(CSRC, f) for f in os.listdir(CSRC) if f.endswith(".h")]
    deps.append(os.path.join(ROOT, "include", "physgrad_mpm.h"))
    return any(os.path.getmtime(d) > t for d in deps)


if __name__ == "__main__":
    build()
    if len(sys.argv) > 1 and sys.argv[1] == "asan":
        build_asan()

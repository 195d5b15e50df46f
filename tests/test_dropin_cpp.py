"""The C++ drop-in of the reference CPU solver (include/physgrad/material_point_method.h)
compiles against the C-ABI and, on the GPU, reproduces the reference fixtures."""
import os
import subprocess

import numpy as np
import pytest

from conftest import GOLDEN, ROOT

SRC = os.path.join(ROOT, "tests", "cpp", "dropin_refcompat.cpp")


def build(tmp_path):
    exe = os.path.join(str(tmp_path), "dropin_refcompat")
    lib_dir = os.path.join(ROOT, "physgrad_amd")
    cmd = ["g++", "-std=c++17", "-O2", f"-I{os.path.join(ROOT, 'include')}", SRC, f"-L{lib_dir}",
           "-lphysgrad_mpm", f"-Wl,-rpath,{lib_dir}", "-o", exe]
    subprocess.run(cmd, check=True, capture_output=True, text=True)
    return exe


def test_dropin_compiles_and_links(tmp_path):
    exe = build(tmp_path)
    assert os.path.exists(exe)


@pytest.mark.gpu
def test_dropin_matches_reference_fixture(gpu, tmp_path):
    f = np.load(os.path.join(GOLDEN, "refcompat_box512_o2.npz"))
    exe = build(tmp_path)
    n = len(f["m"])
    steps = 10
    inp = os.path.join(str(tmp_path), "in.bin")
    out = os.path.join(str(tmp_path), "out.bin")
    with open(inp, "wb") as fh:
        np.array([*f["dims"], int(f["order"]), steps, n], np.int32).tofile(fh)
        np.array([*f["cell"], *f["origin"], f["dt"], *f["gravity"]], np.float32).tofile(fh)
        f["x0"].astype(np.float32).tofile(fh)
        f["v0"].astype(np.float32).tofile(fh)
        f["m"].astype(np.float32).tofile(fh)
    r = subprocess.run([exe, inp, out], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    G = int(np.prod(f["dims"]))
    data = np.fromfile(out, np.float32)
    x, v = data[:3 * n], data[3 * n:6 * n]
    gm = data[6 * n:6 * n + G]
    gp = data[6 * n + G:6 * n + 4 * G]

    def rel(a, b):
        return float(np.abs(a - b).max() / max(np.abs(b).max(), 1e-30))

    assert rel(x, f["x_10"].reshape(-1)) <= 1e-5
    assert rel(v, f["v_10"].reshape(-1)) <= 1e-5
    assert rel(gm, f["grid_mass_10"]) <= 1e-5
    assert rel(gp, f["grid_mom_10"].reshape(-1)) <= 1e-5
    assert "steps 10" in r.stdout

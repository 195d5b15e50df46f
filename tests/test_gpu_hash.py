"""GPU spatial hash (cell ids + stable radix sort) bit-exact against the
reference SpatialHashGrid fixtures and the counting-sort oracle."""
import ctypes
import glob
import os

import numpy as np
import pytest
import torch

from conftest import GOLDEN

pytestmark = pytest.mark.gpu
FILES = sorted(glob.glob(os.path.join(GOLDEN, "hash_*.npz")))


def gpu_hash(pos, cell, lo, res):
    from physgrad_amd import _lib
    L = _lib.load()
    n = pos.shape[0]
    total = int(np.prod(res.astype(np.uint64)))
    d_pos = torch.as_tensor(pos.reshape(-1), device="cuda")
    ids = torch.empty(n, dtype=torch.int32, device="cuda")
    idx = torch.empty(n, dtype=torch.int32, device="cuda")
    st = torch.empty(total, dtype=torch.int32, device="cuda")
    ct = torch.empty(total, dtype=torch.int32, device="cuda")
    ws_bytes = L.pg_spatial_hash_workspace_size(n)
    ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
    lo_c = (ctypes.c_float * 3)(*[float(a) for a in lo])
    res_c = (ctypes.c_uint * 3)(*[int(r) for r in res])
    stream = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    rc = L.pg_spatial_hash_build(ctypes.c_void_p(d_pos.data_ptr()), n, float(cell), lo_c, res_c,
                                 ctypes.c_void_p(ids.data_ptr()), ctypes.c_void_p(idx.data_ptr()),
                                 ctypes.c_void_p(st.data_ptr()), ctypes.c_void_p(ct.data_ptr()),
                                 ctypes.c_void_p(ws.data_ptr()), ws_bytes, stream)
    _lib.check(rc)
    torch.cuda.synchronize()
    u = lambda t: t.cpu().numpy().view(np.uint32)  # noqa: E731
    return u(ids), u(idx), u(st), u(ct)


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_hash_vs_reference_fixtures(gpu, path):
    f = np.load(path)
    print("case", os.path.basename(path), flush=True)
    res = np.ceil((f["hi"] - f["lo"]) / f["cell"]).astype(np.uint32)
    ids, idx, st, ct = gpu_hash(f["pos"], f["cell"], f["lo"], res)
    assert np.array_equal(ids, f["ids"])
    assert np.array_equal(idx, f["idx"])
    assert np.array_equal(st, f["start"])
    assert np.array_equal(ct, f["count"])


@pytest.mark.parametrize("n", [1, 4095, 4097, 1 << 20])
def test_hash_large_vs_oracle(gpu, oracle_lib, n):
    O = oracle_lib
    rng = np.random.default_rng(n)
    pos = rng.random((n, 3)).astype(np.float32)
    res = np.array([128, 128, 128], np.uint32)
    ids, idx, st, ct = gpu_hash(pos, 1 / 128, (0, 0, 0), res)
    want_ids = O.cell_ids(pos, 1 / 128, (0, 0, 0), res)
    w_idx, w_st, w_ct = O.counting_sort(want_ids, 128 ** 3)
    assert np.array_equal(ids, want_ids)
    assert np.array_equal(idx, w_idx)
    assert np.array_equal(st, w_st) and np.array_equal(ct, w_ct)

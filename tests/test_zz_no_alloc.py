"""GPU: the allocation-free spatial hash (pg_spatial_hash_build_noalloc)
against pg_spatial_hash_build with a caller workspace, bit-exact, on two calls
in a row (the second smaller, so stale scratch contents from the first would
show), plus the size limit."""
import ctypes

import numpy as np
import pytest
import torch

pytestmark = pytest.mark.gpu


def _ptr(t):
    return ctypes.c_void_p(t.data_ptr())


def build(pos, cell, res, noalloc):
    from physgrad_amd import _lib
    L = _lib.load()
    n = pos.shape[0]
    total = int(np.prod(res.astype(np.uint64)))
    d_pos = torch.as_tensor(pos.reshape(-1), device="cuda")
    out = [torch.full((k,), -7, dtype=torch.int32, device="cuda") for k in (n, n, total, total)]
    lo = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    r = (ctypes.c_uint * 3)(*[int(v) for v in res])
    st = ctypes.c_void_p(torch.cuda.current_stream().cuda_stream)
    if noalloc:
        rc = L.pg_spatial_hash_build_noalloc(_ptr(d_pos), n, float(cell), lo, r, *[_ptr(t) for t in out], st)
    else:
        ws_bytes = L.pg_spatial_hash_workspace_size(n)
        ws = torch.empty(max(ws_bytes, 1), dtype=torch.uint8, device="cuda")
        rc = L.pg_spatial_hash_build(_ptr(d_pos), n, float(cell), lo, r, *[_ptr(t) for t in out], _ptr(ws),
                                     ws_bytes, st)
    _lib.check(rc)
    torch.cuda.synchronize()
    return [t.cpu().numpy().view(np.uint32) for t in out]


def test_noalloc_matches_workspace_build_on_two_calls(gpu):
    from physgrad_amd import _lib
    rng = np.random.default_rng(21)
    cases = [(_lib.SPATIAL_HASH_NOALLOC_MAX, 1 / 128, np.array([128, 128, 128], np.uint32)),
             (5000, 1 / 40, np.array([40, 37, 45], np.uint32))]
    for n, cell, res in cases:
        print("noalloc case", n, flush=True)
        pos = (rng.random((n, 3)) * (res * cell)).astype(np.float32)
        want = build(pos, cell, res, noalloc=False)
        got = build(pos, cell, res, noalloc=True)
        for a, b, name in zip(got, want, ("cell_ids", "particle_indices", "cell_start", "cell_count")):
            assert np.array_equal(a, b), (n, name)
        # and the result is the stable sort of the ids
        assert np.array_equal(got[1], np.argsort(got[0], kind="stable"))


def test_noalloc_rejects_more_than_the_maximum(gpu):
    from physgrad_amd import _lib
    L = _lib.load()
    n = _lib.SPATIAL_HASH_NOALLOC_MAX + 1
    d = torch.zeros(3 * n, dtype=torch.float32, device="cuda")
    o = torch.zeros(n, dtype=torch.int32, device="cuda")
    c = torch.zeros(8, dtype=torch.int32, device="cuda")
    lo = (ctypes.c_float * 3)(0.0, 0.0, 0.0)
    r = (ctypes.c_uint * 3)(2, 2, 2)
    rc = L.pg_spatial_hash_build_noalloc(_ptr(d), n, 0.5, lo, r, _ptr(o), _ptr(o), _ptr(c), _ptr(c), None)
    assert rc == _lib.PG_ERR_INVALID
    assert b"exceed the maximum" in L.pg_mpm_last_error(None)

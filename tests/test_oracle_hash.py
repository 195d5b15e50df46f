"""Cell keys + counting sort oracle vs the reference SpatialHashGrid fixtures."""
import glob
import os

import numpy as np
import pytest

from conftest import GOLDEN

FILES = sorted(glob.glob(os.path.join(GOLDEN, "hash_*.npz")))


@pytest.mark.parametrize("path", FILES, ids=[os.path.basename(p) for p in FILES])
def test_hash_oracle_bitwise(oracle_lib, path):
    O = oracle_lib
    f = np.load(path)
    res = np.ceil((f["hi"] - f["lo"]) / f["cell"]).astype(np.uint32)
    ids = O.cell_ids(f["pos"], f["cell"], f["lo"], res)
    assert np.array_equal(ids, f["ids"])
    idx, st, ct = O.counting_sort(ids, int(np.prod(res.astype(np.uint64))))
    assert np.array_equal(idx, f["idx"])
    assert np.array_equal(st, f["start"])
    assert np.array_equal(ct, f["count"])


def test_morton_known_values(oracle_lib):
    O = oracle_lib
    # bit interleave x0 y0 z0 x1 y1 z1 ...
    assert O.lib().orc_morton3(1, 0, 0) == 1
    assert O.lib().orc_morton3(0, 1, 0) == 2
    assert O.lib().orc_morton3(0, 0, 1) == 4
    assert O.lib().orc_morton3(3, 0, 0) == 9
    assert O.lib().orc_morton3(1023, 1023, 1023) == (1 << 30) - 1


def test_sort_is_stable_and_grouped(oracle_lib):
    O = oracle_lib
    rng = np.random.default_rng(5)
    ids = rng.integers(0, 50, 1000).astype(np.uint32)
    idx, st, ct = O.counting_sort(ids, 50)
    np.testing.assert_array_equal(idx, np.argsort(ids, kind="stable"))
    np.testing.assert_array_equal(ct, np.bincount(ids, minlength=50))

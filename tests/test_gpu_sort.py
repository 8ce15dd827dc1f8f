"""The GPU introsort emulation reproduces libstdc++ std::sort's permutation (ties included)."""
import ctypes

import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests.gpu_common import log

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


def gpu_sort(ctx, a):
    a = np.array(a, dtype=np.uint64, copy=True)
    f = pma.lib().pma_testing_sort_desc
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    assert f(ctx._h, a.ctypes.data_as(ctypes.c_void_p), a.size) == 0
    return a


def records(keys):
    keys = np.asarray(keys, dtype=np.uint64)
    return (keys << np.uint64(32)) | np.arange(keys.size, dtype=np.uint64)


def cases():
    rng = np.random.default_rng(3)
    out = {"empty": [], "one": [5], "sixteen": rng.integers(0, 3, 16), "seventeen": rng.integers(0, 3, 17)}
    for n in [50, 1000, 100_000, 2_000_000]:
        for nk in [1, 3, 40, 1 << 20]:
            out[f"n{n}_k{nk}"] = rng.integers(0, nk, n)
    out["ascending"] = np.arange(300_000)
    out["descending"] = np.arange(300_000)[::-1]
    out["organ_pipe"] = np.concatenate([np.arange(100_000), np.arange(100_000)[::-1]])
    out["sawtooth"] = np.arange(500_000) % 97
    return out


CASES = cases()


@pytest.mark.parametrize("name", sorted(CASES))
def test_sort_matches_libstdcxx(ctx, name):
    log(f"sort case {name}")
    r = records(CASES[name])
    assert np.array_equal(gpu_sort(ctx, r), oracle.sort_desc(r))

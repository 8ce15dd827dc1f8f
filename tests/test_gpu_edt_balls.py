"""GPU parity of the EDT and containing-ball stages against the CPU oracle (bit-exact)."""
import numpy as np
import pytest

import porousmediaanalysis_amd as pma
from tests import example_fixture as ex
from tests import volumes
from tests.gpu_common import assert_front_equal, log, oracle_front, run_front
from tests.volume_utils import sha256_array

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


CASES = dict(volumes.edge_cases())
CASES["spheres_48"] = volumes.spheres((48, 48, 48), 40, 3, 7, 1)
CASES["gauss_noncubic"] = volumes.gaussian((29, 41, 67), 2.0, 0.65, 2)
CASES["gauss_wide"] = volumes.gaussian((20, 70, 130), 3.0, 0.75, 3)
CASES["noise"] = volumes.random_noise((17, 19, 23), 0.15, 4)


@pytest.mark.parametrize("name", sorted(CASES))
def test_front_matches_oracle(ctx, name):
    log(f"edt+balls case {name} shape {CASES[name].shape}")
    assert_front_equal(run_front(ctx, CASES[name]), oracle_front(CASES[name]))


def test_front_example_fixture(ctx):
    log("edt+balls Docs/Example")
    solid = ex.solid()
    g = run_front(ctx, solid)
    d = ex.digests()
    assert sha256_array(g["e"]) == d["edt2"]
    assert sha256_array(g["cr"]) == d["cr2"]
    assert g["max_e"] == 2561 and g["n_max"] == 380507
    assert ctx.text(pma.TEXT_EDT_COUNTS) == ex.text("euclideanDistanceSquareCounts.txt")
    assert_front_equal(g, oracle_front(solid))


def test_edt_longest_rows(ctx):
    """Rows of the largest Y pass Y accepts (16 320 values and 32 padding entries either side in
    64 KB of LDS); the solid lies only at y = 7, so the y searches run far past the padding.  One
    more value along y is refused before any work."""
    from oracle import edt as oracle_edt
    solid = np.zeros((2, 3, 16320), np.uint8)
    solid[:, :, 7] = 1
    log("edt rows of 16320")
    ctx.load(pma.VOL_SOLID, solid)
    ctx.edt()
    e, m, h = oracle_edt(solid)
    assert np.array_equal(ctx.read(pma.VOL_EDT2), e)
    assert ctx.statistics()[0] == m
    assert np.array_equal(ctx.histogram(), h)
    log("edt rows of 16321 (refused)")
    ctx.load(pma.VOL_SOLID, np.zeros((2, 3, 16321), np.uint8))
    with pytest.raises(pma.PmaError, match="Y must be <= 16320"):
        ctx.edt()


def test_generators_deterministic(ctx):
    log("generators")
    ctx.generate_gaussian((40, 36, 44), 3.0, 0.7, 5)
    a = ctx.read(pma.VOL_SOLID)
    ctx.generate_gaussian((40, 36, 44), 3.0, 0.7, 5)
    b = ctx.read(pma.VOL_SOLID)
    assert np.array_equal(a, b)
    assert abs((a == 0).mean() - 0.7) < 1e-3
    ctx.generate_spheres((64, 64, 64), 120, 6, 12, 9)
    s = ctx.read(pma.VOL_SOLID)
    assert 0.2 < (s != 0).mean() < 0.8

"""GPU parity of the pore seeds / pore propagation stage and of empty-pore removal.

The stage is fed the oracle's EDT, containing radii and maximum-ball mask, exactly as the
reference's -apt mode reads the folders the earlier stages wrote; PoreThroatIds, pores.txt
and sharedPixelTypes.txt must match bit for bit, before and after empty-pore removal.
"""
import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests import example_fixture as ex
from tests import volumes
from tests.gpu_common import log
from tests.volume_utils import canonical_relabel, sha256_array

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


def oracle_front(solid):
    e, max_e, hist = oracle.edt(solid, threads=4)
    cr, cx, cy, cz, mask, nmax = oracle.containing_balls(e, threads=4)
    return e, cr, mask


def gpu_pores(ctx, e, cr, mask, d3q7):
    ctx.set_stencil(d3q7)
    ctx.load(pma.VOL_EDT2, e)
    ctx.load(pma.VOL_CR2, cr)
    ctx.load(pma.VOL_MASK, mask)
    ctx.pores()
    first = ctx.read(pma.VOL_PT), ctx.text(pma.TEXT_PORES), ctx.text(pma.TEXT_THROATS)
    ctx.remove_empty_pores()
    return first, (ctx.read(pma.VOL_PT), ctx.text(pma.TEXT_PORES), ctx.text(pma.TEXT_THROATS))


def assert_same(g, o, what):
    pt, ptxt, ktxt = g
    opt, optxt, oktxt = o
    assert ptxt == optxt, f"{what}: pores.txt differs"
    assert ktxt == oktxt, f"{what}: sharedPixelTypes.txt differs"
    if not np.array_equal(pt, opt):
        bad = np.argwhere(pt != opt)
        raise AssertionError(f"{what}: {len(bad)} labels differ, first {tuple(bad[0])}: "
                             f"{pt[tuple(bad[0])]} vs {opt[tuple(bad[0])]}")


CASES = dict(volumes.edge_cases())
CASES["spheres_48"] = volumes.spheres((48, 48, 48), 40, 3, 7, 1)
CASES["spheres_dense"] = volumes.spheres((40, 52, 44), 90, 2, 6, 11)
CASES["gauss_s4"] = volumes.gaussian((48, 56, 64), 4.0, 0.7, 3)
CASES["gauss_s3_low"] = volumes.gaussian((40, 40, 40), 3.0, 0.5, 12)
CASES["noise"] = volumes.random_noise((17, 19, 23), 0.15, 4)
CASES["noise_many_pores"] = volumes.random_noise((40, 48, 56), 0.45, 6)  # > 129 pores: the +127 ID jump


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_pores_match_oracle(ctx, name, d3q7):
    log(f"pores case {name} d3q7={d3q7} shape {CASES[name].shape}")
    e, cr, mask = oracle_front(CASES[name])
    first, removed = gpu_pores(ctx, e, cr, mask, d3q7)
    o = oracle.pores(mask, e, cr, d3q7)
    assert_same(first, o, "pores")
    assert_same(removed, oracle.remove_empty_pores(*o), "after empty-pore removal")


def test_pores_example_partition(ctx):
    """Fed the fixture's own maximum-ball mask, the pore partition equals the fixture's."""
    log("pores Docs/Example")
    solid = ex.solid()
    e, cr, _ = oracle_front(solid)
    first, removed = gpu_pores(ctx, e, cr, ex.mask(), False)
    assert sha256_array(canonical_relabel(removed[0])) == ex.digests()["pt_partition"]
    o = oracle.pores(ex.mask(), e, cr)
    assert_same(first, o, "pores")
    assert_same(removed, oracle.remove_empty_pores(*o), "after empty-pore removal")

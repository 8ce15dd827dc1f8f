"""GPU parity of the pore seeds / pore propagation stage and of empty-pore removal.

The stage is fed the oracle's EDT, containing radii and maximum-ball mask, exactly as the
reference's -apt mode reads the folders the earlier stages wrote; PoreThroatIds, pores.txt
and sharedPixelTypes.txt must match bit for bit, before and after empty-pore removal.
"""
import ctypes

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


def set_ball_order(ctx, order):
    """Test hook: the next pma_pores takes the maximum balls in this order instead of sorting."""
    order = np.ascontiguousarray(order, dtype=np.uint32)
    f = pma.lib().pma_testing_set_ball_order
    f.argtypes = [ctypes.c_void_p, ctypes.c_void_p, ctypes.c_size_t]
    f.restype = ctypes.c_int
    assert f(ctx._h, order.ctypes.data_as(ctypes.c_void_p), order.size) == 0, pma.lib().pma_last_error()


def test_pores_example_bit_exact_in_the_fixture_seed_order(ctx):
    """The reference's own PoreThroatIds, pores.txt and sharedPixelTypes.txt, throat-key IDs
    included, once the seeds are taken in the fixture's order (MSVC's std::sort breaks r2 ties
    differently from libstdc++)."""
    log("pores Docs/Example in the fixture's seed order")
    solid = ex.solid()
    e, cr, _ = oracle_front(solid)
    mask = ex.mask()
    set_ball_order(ctx, ex.ball_order(mask, e))
    first, removed = gpu_pores(ctx, e, cr, mask, False)
    pt, ptxt, ktxt = removed
    assert ptxt == ex.text("pores.txt")
    assert ktxt == ex.text("sharedPixelTypes.txt")
    assert sha256_array(pt) == ex.digests()["pt"]


def test_watershed_example_bit_exact_from_the_fixture_pores(ctx):
    """Fed the fixture's pores.txt and the (bit-exact) containing radii, the watershed writes the
    reference's own PoreThroatIdsWatershed, poresWatershed.txt and sharedPixelTypesWatershed.txt."""
    log("watershed Docs/Example from the fixture's pores.txt")
    solid = ex.solid()
    e, max_e, _ = oracle.edt(solid, threads=4)
    cr, *_ = oracle.containing_balls(e, threads=4)
    ctx.set_stencil(False)
    ctx.load(pma.VOL_EDT2, e)
    ctx.load(pma.VOL_CR2, cr)
    ctx.set_text(pma.TEXT_STATS, ex.text("intermediateStatistics.txt"))
    ctx.set_text(pma.TEXT_PORES, ex.text("pores.txt"))
    ctx.watershed()
    assert ctx.text(pma.TEXT_PORES_WS) == ex.text("poresWatershed.txt")
    assert ctx.text(pma.TEXT_THROATS_WS) == ex.text("sharedPixelTypesWatershed.txt")
    assert sha256_array(ctx.read(pma.VOL_PTW)) == ex.digests()["ptw"]


@pytest.mark.parametrize("d3q7", [False, True])
def test_pores_in_a_given_ball_order_match_oracle(ctx, d3q7):
    """The order hook against the oracle's: r2 descending with ties in random order opens other
    seeds (and so other IDs and keys) than libstdc++'s tie order, and both sides must still agree
    bit for bit."""
    vol = volumes.gaussian((48, 56, 64), 4.0, 0.7, 3)
    log(f"pores in a random tie order d3q7={d3q7}")
    e, cr, mask = oracle_front(vol)
    balls = np.flatnonzero(mask.reshape(-1) == 255)
    r2 = e.reshape(-1)[balls].astype(np.int64)
    order = balls[np.lexsort((np.random.default_rng(5).random(balls.size), -r2))].astype(np.uint32)
    set_ball_order(ctx, order)
    first, removed = gpu_pores(ctx, e, cr, mask, d3q7)
    o = oracle.pores_in_order(e, cr, order, d3q7)
    assert_same(first, o, "pores")
    assert o[1] != oracle.pores(mask, e, cr, d3q7)[1]  # the order did change the seeds
    assert_same(removed, oracle.remove_empty_pores(*o), "after empty-pore removal")


def test_ball_order_is_checked(ctx):
    """An order that misses a maximum ball, or raises r2, is rejected; the hook's order is used
    once, and the next call sorts again."""
    vol = volumes.gaussian((24, 24, 24), 2.0, 0.7, 8)
    e, cr, mask = oracle_front(vol)
    balls = np.flatnonzero(mask.reshape(-1) == 255)
    order = balls[np.argsort(-e.reshape(-1)[balls].astype(np.int64), kind="stable")].astype(np.uint32)
    ctx.set_stencil(False)
    ctx.load(pma.VOL_EDT2, e)
    ctx.load(pma.VOL_CR2, cr)
    ctx.load(pma.VOL_MASK, mask)
    set_ball_order(ctx, order[:-1])
    with pytest.raises(pma.PmaError, match="ball order"):
        ctx.pores()
    set_ball_order(ctx, order[::-1].copy())
    with pytest.raises(pma.PmaError, match="records_in_order"):
        ctx.pores()
    ctx.pores()
    assert ctx.text(pma.TEXT_PORES) == oracle.pores(mask, e, cr)[1]

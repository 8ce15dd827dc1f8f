"""GPU parity of the watershed stage against the CPU oracle (labels and text bit-exact).

The stage is fed the oracle's containing radii, statistics and pores.txt, exactly as the
reference's watershed reads the files the earlier stages wrote.
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


def oracle_inputs(solid, d3q7, mask=None):
    e, max_e, hist = oracle.edt(solid, threads=4)
    cr, cx, cy, cz, m, nmax = oracle.containing_balls(e, threads=4)
    pt, ptxt, ktxt = oracle.pores(m if mask is None else mask, e, cr, d3q7)
    pt, ptxt, ktxt = oracle.remove_empty_pores(pt, ptxt, ktxt)
    return cr, max_e, nmax, ptxt


def gpu_watershed(ctx, cr, max_e, nmax, ptxt, d3q7):
    ctx.set_stencil(d3q7)
    ctx.load(pma.VOL_CR2, cr)
    ctx.set_text(pma.TEXT_STATS, f"Max Euclidean distance square: {max_e}\nMaximum balls count: {nmax}\n")
    ctx.set_text(pma.TEXT_PORES, ptxt)
    ctx.watershed()
    return ctx.read(pma.VOL_PTW), ctx.text(pma.TEXT_PORES_WS), ctx.text(pma.TEXT_THROATS_WS)


def check(ctx, solid, d3q7=False):
    cr, max_e, nmax, ptxt = oracle_inputs(solid, d3q7)
    ptw, wp, wk = gpu_watershed(ctx, cr, max_e, nmax, ptxt, d3q7)
    optw, owp, owk = oracle.watershed(cr, max_e, ptxt, d3q7)
    assert wp == owp, "poresWatershed.txt differs"
    assert wk == owk, "sharedPixelTypesWatershed.txt differs"
    if not np.array_equal(ptw, optw):
        bad = np.argwhere(ptw != optw)
        raise AssertionError(f"{len(bad)} labels differ, first {tuple(bad[0])}: {ptw[tuple(bad[0])]} vs {optw[tuple(bad[0])]}")


CASES = dict(volumes.edge_cases())
CASES["spheres_48"] = volumes.spheres((48, 48, 48), 40, 3, 7, 1)
CASES["gauss_s2"] = volumes.gaussian((29, 41, 67), 2.0, 0.65, 2)
CASES["gauss_s4"] = volumes.gaussian((48, 56, 64), 4.0, 0.7, 3)
CASES["noise"] = volumes.random_noise((17, 19, 23), 0.15, 4)


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_watershed_matches_oracle(ctx, name, d3q7):
    log(f"watershed case {name} d3q7={d3q7}")
    check(ctx, CASES[name], d3q7)


def test_watershed_example_partition(ctx):
    """Fed the fixture's own maximum-ball mask, the watershed partition equals the fixture's."""
    log("watershed Docs/Example")
    solid = ex.solid()
    cr, max_e, nmax, ptxt = oracle_inputs(solid, False, mask=ex.mask())
    ptw, wp, wk = gpu_watershed(ctx, cr, max_e, nmax, ptxt, False)
    assert sha256_array(canonical_relabel(ptw)) == ex.digests()["ptw_partition"]
    optw, owp, owk = oracle.watershed(cr, max_e, ptxt)
    assert np.array_equal(ptw, optw) and wp == owp and wk == owk

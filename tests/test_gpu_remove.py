"""GPU parity of empty-pore removal (EmptyPoresRemover.cpp:51-244) against the oracle.

Fed a PoreThroatIds volume and the two tables (see tests/removal_inputs.py: the oracle's pore
stage output with some pore volumes set to 0), the relabelled volume, pores.txt and
sharedPixelTypes.txt must match the oracle's bit for bit.
"""
import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests.gpu_common import log
from tests.removal_inputs import CASES, removal_input

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
@pytest.mark.parametrize("every", [2, 3])
def test_remove_matches_oracle(ctx, name, d3q7, every):
    log(f"remove case {name} d3q7={d3q7} every={every}")
    pt, ptxt, ktxt = removal_input(name, d3q7, every)
    opt, optxt, oktxt = oracle.remove_empty_pores(pt, ptxt, ktxt)
    assert optxt.count("\n") < ptxt.count("\n"), "the input drops no pore"
    ctx.load(pma.VOL_PT, pt)
    ctx.set_text(pma.TEXT_PORES, ptxt)
    ctx.set_text(pma.TEXT_THROATS, ktxt)
    ctx.remove_empty_pores()
    assert ctx.text(pma.TEXT_PORES) == optxt, "pores.txt differs"
    assert ctx.text(pma.TEXT_THROATS) == oktxt, "sharedPixelTypes.txt differs"
    g = ctx.read(pma.VOL_PT)
    if not np.array_equal(g, opt):
        bad = np.argwhere(g != opt)
        raise AssertionError(f"{len(bad)} labels differ, first {tuple(bad[0])}: {g[tuple(bad[0])]} vs {opt[tuple(bad[0])]}")

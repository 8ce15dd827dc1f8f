"""The benchmark's workload at 192^3 and 256^3 (and 160^3 for the 6-neighbour watershed): the
GPU Gaussian-field generator (sigma 8, porosity 0.65), then EDT, containing balls, pores,
empty-pore removal, watershed and geodesic distances with cavities, each bit-exact against the
oracle.  At this size the containing-ball stage takes the paths the 1024^3 run takes (balls
packed in the pair sort, chord widths from the LDS table, 16-bit winner slots) and the
geodesic runs a few hundred chained rounds."""
import os

import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests.gpu_common import assert_front_equal, log

pytestmark = pytest.mark.gpu

THREADS = min(os.cpu_count() or 1, 16)


@pytest.mark.parametrize("seed", [20261203, 7])
def test_gaussian_192_matches_oracle(seed):
    n = 192
    log(f"bench workload {n}^3 seed {seed}")
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((n, n, n), 8.0, 0.65, seed)
        solid = ctx.read(pma.VOL_SOLID)
        ctx.edt()
        ctx.containing_balls()
        ctx.geodesic()
        m, nmax = ctx.statistics()
        g = dict(e=ctx.read(pma.VOL_EDT2), max_e=m, hist=ctx.histogram(), cr=ctx.read(pma.VOL_CR2),
                 cx=ctx.read(pma.VOL_CX), cy=ctx.read(pma.VOL_CY), cz=ctx.read(pma.VOL_CZ),
                 mask=ctx.read(pma.VOL_MASK), n_max=nmax)
        dist, cav, text = ctx.read(pma.VOL_SPD), ctx.read(pma.VOL_CAVITY), ctx.text(pma.TEXT_CAVITIES)
    assert abs((solid == 0).mean() - 0.65) < 1e-3
    e, max_e, hist = oracle.edt(solid, threads=THREADS)
    cr, cx, cy, cz, mask, n_max = oracle.containing_balls(e, threads=THREADS)
    assert_front_equal(g, dict(e=e, max_e=max_e, hist=hist, cr=cr, cx=cx, cy=cy, cz=cz, mask=mask, n_max=n_max))
    odist, ocav, otext = oracle.geodesic(solid)
    assert np.array_equal(dist.view(np.uint32), odist.view(np.uint32)), "distances differ"
    assert np.array_equal(cav, ocav), "cavity ids differ"
    assert text == otext


def test_gaussian_192_removal_watershed():
    """The same workload through empty-pore removal and the watershed on the GPU, fed the oracle's
    pore output: labels and text bit-exact against the oracle.  Each stage runs twice on one context; the second run's device time is logged."""
    n = 192
    log(f"removal and watershed on the bench workload {n}^3")
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((n, n, n), 8.0, 0.65, 20261203)
        solid = ctx.read(pma.VOL_SOLID)
    e, max_e, _ = oracle.edt(solid, threads=THREADS)
    cr, _, _, _, mask, n_max = oracle.containing_balls(e, threads=THREADS)
    pt, ptxt, ktxt = oracle.pores(mask, e, cr)
    opt, optxt, oktxt = oracle.remove_empty_pores(pt, ptxt, ktxt)
    optw, owp, owk = oracle.watershed(cr, max_e, optxt)
    times = {}
    with pma.Context(0) as ctx:
        for _ in range(2):
            ctx.load(pma.VOL_PT, pt)
            ctx.set_text(pma.TEXT_PORES, ptxt)
            ctx.set_text(pma.TEXT_THROATS, ktxt)
            ctx.remove_empty_pores()
            times["remove_empty_pores"] = ctx.stage_info(pma.STAGE_REMOVE)["device_ms"]
            assert np.array_equal(ctx.read(pma.VOL_PT), opt), "removal labels differ"
            assert ctx.text(pma.TEXT_PORES) == optxt and ctx.text(pma.TEXT_THROATS) == oktxt
        for _ in range(2):
            ctx.load(pma.VOL_CR2, cr)
            ctx.set_text(pma.TEXT_STATS, f"Max Euclidean distance square: {max_e}\nMaximum balls count: {n_max}\n")
            ctx.set_text(pma.TEXT_PORES, optxt)
            ctx.watershed()
            times["watershed"] = ctx.stage_info(pma.STAGE_WATERSHED)["device_ms"]
            assert np.array_equal(ctx.read(pma.VOL_PTW), optw), "watershed labels differ"
            assert ctx.text(pma.TEXT_PORES_WS) == owp and ctx.text(pma.TEXT_THROATS_WS) == owk
    log(f"{n}^3: {optxt.count(chr(10)) - 1} pores; device ms: " +
        ", ".join(f"{k} {v:.2f}" for k, v in times.items()))


def test_gaussian_160_watershed_d3q7():
    """The watershed on the same generator at 160^3 with the 6-neighbour stencil (8 lanes per
    item in the round kernels), fed the oracle's D3Q7 pores after empty-pore removal: labels and
    text bit-exact against the oracle, on two runs in one context."""
    n = 160
    log(f"watershed d3q7 on the bench workload {n}^3")
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((n, n, n), 8.0, 0.65, 20261203)
        solid = ctx.read(pma.VOL_SOLID)
    e, max_e, _ = oracle.edt(solid, threads=THREADS)
    cr, _, _, _, mask, n_max = oracle.containing_balls(e, threads=THREADS)
    pt, ptxt, ktxt = oracle.pores(mask, e, cr, True)
    _, optxt, _ = oracle.remove_empty_pores(pt, ptxt, ktxt)
    optw, owp, owk = oracle.watershed(cr, max_e, optxt, True)
    with pma.Context(0) as ctx:
        ctx.set_stencil(True)
        for _ in range(2):
            ctx.load(pma.VOL_CR2, cr)
            ctx.set_text(pma.TEXT_STATS, f"Max Euclidean distance square: {max_e}\nMaximum balls count: {n_max}\n")
            ctx.set_text(pma.TEXT_PORES, optxt)
            ctx.watershed()
            assert np.array_equal(ctx.read(pma.VOL_PTW), optw), "watershed labels differ"
            assert ctx.text(pma.TEXT_PORES_WS) == owp and ctx.text(pma.TEXT_THROATS_WS) == owk
        ms = ctx.stage_info(pma.STAGE_WATERSHED)["device_ms"]
    log(f"{n}^3 d3q7: {optxt.count(chr(10)) - 1} pores; watershed device ms {ms:.2f}")


def test_gaussian_256_pores_watershed():
    """The pore stage, empty-pore removal and the watershed of the GPU pipeline on the benchmark's
    generator at 256^3, against the oracle run on the GPU's own EDT, containing radii and
    maximum-ball mask (the stage inputs of the reference's -apt mode).  At this size plateau
    chains are long enough that the flattening of the plateau forest meets concurrent walks
    through the voxels it writes (a flatten that path-halves there left non-root parents)."""
    n = 256
    log(f"pores, removal and watershed on the bench workload {n}^3")
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((n, n, n), 8.0, 0.65, 20261203)
        ctx.edt()
        ctx.containing_balls()
        ctx.pores()
        pt, ptxt, ktxt = ctx.read(pma.VOL_PT), ctx.text(pma.TEXT_PORES), ctx.text(pma.TEXT_THROATS)
        ctx.remove_empty_pores()
        rpt, rptxt, rktxt = ctx.read(pma.VOL_PT), ctx.text(pma.TEXT_PORES), ctx.text(pma.TEXT_THROATS)
        ctx.watershed()
        ptw, wp, wk = ctx.read(pma.VOL_PTW), ctx.text(pma.TEXT_PORES_WS), ctx.text(pma.TEXT_THROATS_WS)
        e, cr, mask = ctx.read(pma.VOL_EDT2), ctx.read(pma.VOL_CR2), ctx.read(pma.VOL_MASK)
        max_e, _ = ctx.statistics()
        ms = {k: ctx.stage_info(s)["device_ms"] for k, s in
              [("pores", pma.STAGE_PORES), ("watershed", pma.STAGE_WATERSHED)]}
    opt, optxt, oktxt = oracle.pores(mask, e, cr)
    assert ptxt == optxt, "pores.txt differs"
    assert ktxt == oktxt, "sharedPixelTypes.txt differs"
    assert np.array_equal(pt, opt), "PoreThroatIds differ"
    opt, optxt, oktxt = oracle.remove_empty_pores(opt, optxt, oktxt)
    assert np.array_equal(rpt, opt) and rptxt == optxt and rktxt == oktxt, "after removal"
    optw, owp, owk = oracle.watershed(cr, max_e, optxt)
    assert wp == owp and wk == owk, "watershed text differs"
    assert np.array_equal(ptw, optw), "PoreThroatIdsWatershed differ"
    log(f"{n}^3: {optxt.count(chr(10)) - 1} pores, {oktxt.count(chr(10)) - 1} throat keys; device ms " +
        ", ".join(f"{k} {v:.1f}" for k, v in ms.items()))

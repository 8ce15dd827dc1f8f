"""The batch runner (pma_run_batch) and its autotuner (pma_tune_batch), through the C ABI.

A batch runs the whole pipeline on every subvolume on a pool of worker contexts.  Whatever the
worker count and knobs, every output volume and text of every subvolume must equal what one
context gives for it alone, bit for bit; the first two subvolumes are also checked against the
oracle.
"""
import threading

import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests import volumes
from tests.gpu_common import log

pytestmark = pytest.mark.gpu

VOLS = [("edt2", pma.VOL_EDT2), ("cr2", pma.VOL_CR2), ("cx", pma.VOL_CX), ("cy", pma.VOL_CY), ("cz", pma.VOL_CZ),
        ("mask", pma.VOL_MASK), ("pt", pma.VOL_PT), ("ptw", pma.VOL_PTW), ("dist", pma.VOL_SPD),
        ("cavity", pma.VOL_CAVITY)]
TEXTS = [("pores", pma.TEXT_PORES), ("throats", pma.TEXT_THROATS), ("pores_ws", pma.TEXT_PORES_WS),
         ("throats_ws", pma.TEXT_THROATS_WS), ("cavities", pma.TEXT_CAVITIES)]


def make_batch(n, seed):
    """Uneven subvolumes, alternately sphere unions and Gaussian fields."""
    rng = np.random.default_rng(seed)
    out = []
    for i in range(n):
        shape = tuple(int(v) for v in rng.integers(20, 56, 3))
        if i % 2 == 0:
            count = max(1, round(300 * np.prod(shape) / 128 ** 3 * 4))
            out.append(volumes.spheres(shape, count, 3, 7, seed + i))
        else:
            out.append(volumes.gaussian(shape, float(rng.uniform(2, 4)), float(rng.uniform(0.5, 0.8)), seed + i))
    return out


def outputs(ctx):
    out = {k: ctx.read(w) for k, w in VOLS}
    out.update({k: ctx.text(w) for k, w in TEXTS})
    return out


def compare(g, o, what):
    for k, _ in TEXTS:
        assert g[k] == o[k], f"{what}: {k} differs"
    for k, _ in VOLS:
        a, b = g[k], o[k]
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        assert np.array_equal(a, b), f"{what}: {k} differs in {int((a != b).sum())} voxels"


@pytest.fixture(scope="module")
def batch():
    return make_batch(10, 300)


@pytest.fixture(scope="module", params=[False, True], ids=["d3q27", "d3q7"])
def single(request, batch):
    """Each subvolume on one context, alone."""
    d3q7 = request.param
    res = []
    with pma.Context(0) as ctx:
        ctx.set_stencil(d3q7)
        for v in batch:
            ctx.load(pma.VOL_SOLID, v)
            ctx.run_all()
            res.append(outputs(ctx))
    return d3q7, res


def run(batch, d3q7, config=None):
    got, lock = {}, threading.Lock()

    def done(i, ctx):
        r = outputs(ctx)
        with lock:
            got[i] = r

    pma.run_batch(batch, d3q7, config, done)
    assert sorted(got) == list(range(len(batch)))
    return got


@pytest.mark.parametrize("config", [None, pma.BatchConfig(1, 0.0, 0, 0), pma.BatchConfig(4, 0.0, 0, 0),
                                    pma.BatchConfig(3, 8.0, 1024, 512), pma.BatchConfig(16, 32.0, 16384, 256)],
                         ids=["default", "one_worker", "four_workers", "knobs", "more_workers_than_volumes"])
def test_batch_matches_one_context(batch, single, config):
    d3q7, ref = single
    log(f"batch of {len(batch)} d3q7={d3q7} config={None if config is None else config.as_dict()}")
    got = run(batch, d3q7, config)
    for i in range(len(batch)):
        compare(got[i], ref[i], f"subvolume {i} {batch[i].shape}")


def test_batch_matches_oracle(batch, single):
    d3q7, ref = single
    for i in range(2):
        log(f"oracle pipeline on subvolume {i} {batch[i].shape} d3q7={d3q7}")
        compare(ref[i], oracle.pipeline(batch[i], d3q7, threads=4), f"subvolume {i} vs the oracle")


def test_batch_reports_a_failing_subvolume(batch):
    """A subvolume that cannot run (here: no voxels) fails alone: the others are done."""
    vols = [batch[0], np.zeros((0, 8, 8), np.uint8), batch[1]]
    seen = []
    with pytest.raises(pma.PmaError, match="subvolume 1"):
        pma.run_batch(vols, False, pma.BatchConfig(2, 0.0, 0, 0), lambda i, ctx: seen.append(i))
    assert sorted(seen) == [0, 2]


def test_tune_batch(batch, single):
    d3q7, ref = single
    log(f"autotuner on 6 subvolumes d3q7={d3q7}")
    best, vps = pma.tune_batch(batch[:6], d3q7)
    assert best.workers in (1, 2, 4) and best.geodesic_bucket in (8.0, 16.0, 32.0)
    assert best.flood_grid in (1024, 4096, 16384) and best.replay_memory_mb >= 256
    assert vps > 0
    got = run(batch[:6], d3q7, best)
    for i in range(6):
        compare(got[i], ref[i], f"subvolume {i} with the tuned configuration")


def test_tuning_knobs():
    with pma.Context(0) as ctx:
        assert ctx.tuning(pma.TUNE_GEODESIC_BUCKET) == 16.0
        assert ctx.tuning(pma.TUNE_FLOOD_GRID) == 4096
        assert ctx.tuning(pma.TUNE_REPLAY_MEMORY_MB) == 128 * 1024
        ctx.set_tuning(pma.TUNE_GEODESIC_BUCKET, 4)
        ctx.set_tuning(pma.TUNE_FLOOD_GRID, 512)
        ctx.set_tuning(pma.TUNE_REPLAY_MEMORY_MB, 1024)
        assert (ctx.tuning(pma.TUNE_GEODESIC_BUCKET), ctx.tuning(pma.TUNE_FLOOD_GRID),
                ctx.tuning(pma.TUNE_REPLAY_MEMORY_MB)) == (4.0, 512.0, 1024.0)
        for which, bad in [(pma.TUNE_GEODESIC_BUCKET, 0), (pma.TUNE_FLOOD_GRID, 32), (pma.TUNE_REPLAY_MEMORY_MB, 1)]:
            with pytest.raises(pma.PmaError, match="out of range"):
                ctx.set_tuning(which, bad)

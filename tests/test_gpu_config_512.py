"""The 512^3 configuration: seeded N(0,1) noise blurred with sigma 4, void = the lowest 70 % of the
field (porosity 0.70), D3Q27.  EDT, containing balls (radii, centres, mask) and geodesic distances
with cavities are compared bit for bit with the oracle on the whole volume; the pore and watershed
stages are checked through size-independent properties (the oracle's serial pore DFS at this size
is too slow for a test; their bit-exact comparison runs up to 256^3)."""
import os

import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests.gpu_common import log

pytestmark = pytest.mark.gpu

N = 512


@pytest.fixture(scope="module")
def run():
    log(f"{N}^3 sigma 4 porosity 0.70: all six stages on the GPU")
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((N, N, N), 4.0, 0.70, 512)
        ctx.run_all()
        out = {k: ctx.read(getattr(pma, k)) for k in
               ["VOL_SOLID", "VOL_EDT2", "VOL_CR2", "VOL_CX", "VOL_CY", "VOL_CZ", "VOL_MASK", "VOL_SPD",
                "VOL_CAVITY", "VOL_PT", "VOL_PTW"]}
        out["max_e"], out["n_max"] = ctx.statistics()
        out["cavities"] = ctx.text(pma.TEXT_CAVITIES)
        out["pores"] = ctx.text(pma.TEXT_PORES)
        out["pores_ws"] = ctx.text(pma.TEXT_PORES_WS)
    return out


def test_porosity(run):
    assert abs(float((run["VOL_SOLID"] == 0).mean()) - 0.70) < 1e-4


def test_edt_and_balls_match_oracle(run):
    threads = os.cpu_count() or 1
    log(f"oracle EDT and containing balls at {N}^3 on {threads} threads")
    e, max_e, _ = oracle.edt(run["VOL_SOLID"], threads)
    assert run["max_e"] == max_e
    assert np.array_equal(run["VOL_EDT2"], e)
    cr, cx, cy, cz, mask, n_max = oracle.containing_balls(e, threads)
    assert run["n_max"] == n_max
    for k, v in [("VOL_CR2", cr), ("VOL_CX", cx), ("VOL_CY", cy), ("VOL_CZ", cz), ("VOL_MASK", mask)]:
        assert np.array_equal(run[k], v), k


def test_geodesic_matches_oracle(run):
    log(f"oracle geodesic at {N}^3")
    dist, cav, text = oracle.geodesic(run["VOL_SOLID"])
    assert np.array_equal(run["VOL_SPD"].view(np.uint32), dist.view(np.uint32))
    assert np.array_equal(run["VOL_CAVITY"], cav)
    assert run["cavities"] == text


@pytest.mark.parametrize("which", ["VOL_PT", "VOL_PTW"])
def test_pore_labels_cover_the_void(run, which):
    lab, solid = run[which], run["VOL_SOLID"]
    assert np.array_equal(lab == 0, solid != 0)
    rows = (run["pores"] if which == "VOL_PT" else run["pores_ws"]).strip().splitlines()[1:]
    assert len(rows) > 100
    ids = np.array([int(r.split()[0]) for r in rows])
    assert np.isin(ids, np.unique(lab)).all() or which == "VOL_PTW"

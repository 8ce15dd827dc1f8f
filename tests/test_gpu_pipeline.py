"""GPU parity of the whole pipeline (the reference's -all then -spd) against the oracle.

Every output volume and text file must match bit for bit: EDT, containing radii and
centres, maximum-ball mask, PoreThroatIds, pores.txt / sharedPixelTypes.txt (after
empty-pore removal), PoreThroatIdsWatershed and its two text files, geodesic distances,
cavity IDs and cavityDescriptions.txt.
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

VOLS = [("edt2", pma.VOL_EDT2), ("cr2", pma.VOL_CR2), ("cx", pma.VOL_CX), ("cy", pma.VOL_CY), ("cz", pma.VOL_CZ),
        ("mask", pma.VOL_MASK), ("pt", pma.VOL_PT), ("ptw", pma.VOL_PTW), ("dist", pma.VOL_SPD),
        ("cavity", pma.VOL_CAVITY)]
TEXTS = [("pores", pma.TEXT_PORES), ("throats", pma.TEXT_THROATS), ("pores_ws", pma.TEXT_PORES_WS),
         ("throats_ws", pma.TEXT_THROATS_WS), ("cavities", pma.TEXT_CAVITIES)]


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


def gpu_pipeline(ctx, solid, d3q7):
    ctx.set_stencil(d3q7)
    ctx.load(pma.VOL_SOLID, solid)
    ctx.run_all()
    out = {k: ctx.read(w) for k, w in VOLS}
    out.update({k: ctx.text(w) for k, w in TEXTS})
    return out


def compare(g, o):
    for k, _ in TEXTS:
        assert g[k] == o[k], f"{k} differs:\n--- gpu\n{g[k][:2000]}\n--- oracle\n{o[k][:2000]}"
    for k, _ in VOLS:
        a, b = g[k], o[k]
        if a.dtype == np.float32:
            a, b = a.view(np.uint32), b.view(np.uint32)
        if not np.array_equal(a, b):
            bad = np.argwhere(a != b)
            raise AssertionError(f"{k}: {len(bad)} voxels differ, first (z,x,y)={tuple(bad[0])}: "
                                 f"gpu {a[tuple(bad[0])]} oracle {b[tuple(bad[0])]}")


CASES = dict(volumes.edge_cases())
CASES["spheres_48"] = volumes.spheres((48, 48, 48), 40, 3, 7, 1)
CASES["spheres_dense"] = volumes.spheres((40, 52, 44), 90, 2, 6, 11)
CASES["gauss_s2"] = volumes.gaussian((29, 41, 67), 2.0, 0.65, 2)
CASES["gauss_s4"] = volumes.gaussian((48, 56, 64), 4.0, 0.7, 3)
CASES["gauss_s3_low"] = volumes.gaussian((40, 40, 40), 3.0, 0.5, 12)
CASES["noise"] = volumes.random_noise((17, 19, 23), 0.15, 4)
CASES["noise_many_pores"] = volumes.random_noise((40, 48, 56), 0.45, 6)  # > 129 pores: the +127 ID jump


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_pipeline_matches_oracle(ctx, name, d3q7):
    log(f"pipeline case {name} d3q7={d3q7} shape {CASES[name].shape}")
    compare(gpu_pipeline(ctx, CASES[name], d3q7), oracle.pipeline(CASES[name], d3q7, threads=4))


def test_pipeline_sphere_union_128(ctx):
    """The CPU-runnable configuration: 128^3, 300 spheres, radii in [6, 12]."""
    log("pipeline spheres 128^3")
    ctx.generate_spheres((128, 128, 128), 300, 6.0, 12.0, 2024)
    solid = ctx.read(pma.VOL_SOLID)
    compare(gpu_pipeline(ctx, solid, False), oracle.pipeline(solid, False, threads=8))


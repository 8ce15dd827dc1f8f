"""GPU parity of the geodesic stage against the CPU oracle (float32 bit-exact)."""
import numpy as np
import pytest

import oracle
import porousmediaanalysis_amd as pma
from tests import example_fixture as ex
from tests import volumes
from tests.gpu_common import log
from tests.volume_utils import sha256_array

pytestmark = pytest.mark.gpu


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


def run(ctx, solid, d3q7=False):
    ctx.load(pma.VOL_SOLID, solid)
    ctx.set_stencil(d3q7)
    ctx.geodesic()
    return ctx.read(pma.VOL_SPD), ctx.read(pma.VOL_CAVITY), ctx.text(pma.TEXT_CAVITIES)


def check(ctx, solid, d3q7=False):
    dist, cav, text = run(ctx, solid, d3q7)
    odist, ocav, otext = oracle.geodesic(solid, d3q7)
    assert np.array_equal(dist.view(np.uint32), odist.view(np.uint32)), "distances differ"
    assert np.array_equal(cav, ocav), "cavity ids differ"
    assert text == otext


CASES = dict(volumes.edge_cases())
CASES["spheres_40"] = volumes.spheres((40, 44, 36), 60, 3, 6, 5)
CASES["gauss"] = volumes.gaussian((33, 47, 52), 2.0, 0.55, 6)
# many small cavities: the cavity loop restarts from box shells
CASES["noise"] = volumes.random_noise((20, 21, 22), 0.62, 7)


def _enclosed():
    v = np.zeros((21, 23, 19), np.uint8)
    c = np.array([10, 11, 9])
    z, x, y = np.meshgrid(np.arange(21), np.arange(23), np.arange(19), indexing="ij")
    cheb = np.maximum(np.maximum(abs(z - c[0]), abs(x - c[1])), abs(y - c[2]))
    v[cheb == 4] = 1  # closed shell: the centre's cavity reaches no wall
    v[:, 5, :] = 1    # and a wall splitting the outside
    return v


def _split():
    v = volumes.random_noise((16, 18, 20), 0.2, 8)
    v[:, 9, :] = 1  # centre row solid: the start is found on box shells
    v[:, :, 3] = 1
    return v


CASES["enclosed"] = _enclosed()
CASES["split"] = _split()


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_geodesic_matches_oracle(ctx, name, d3q7):
    log(f"geodesic case {name} d3q7={d3q7}")
    check(ctx, CASES[name], d3q7)


def test_geodesic_example_fixture(ctx):
    log("geodesic Docs/Example")
    dist, cav, text = run(ctx, ex.solid())
    assert sha256_array(dist) == ex.digests()["spd"]
    assert sha256_array(cav) == ex.digests()["cavity"]
    assert text == ex.text("cavityDescriptions.txt")

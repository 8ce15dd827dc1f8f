"""The reference's folder contract on the GPU: -all, -edt, -rib, -apt and -spd over a folder with
config.txt and Images/, through pma_run_folder and the pma console program.

On the Docs/Example folder (its own config.txt and five input slices) the written
EuclideanDistanceSquares, ContainingBallRadiiSquares, ShortestPathDistances and CavityIds slices
read back to the fixture's volumes bit for bit, and the text files equal the fixture's.  Fed the
fixture's own IsMaximumBall slices between -rib and -apt (as when resuming a folder the
reference wrote), the PoreThroatIds and PoreThroatIdsWatershed partitions equal the fixture's.
Anisotropic resolutions are resampled as the reference does."""
import os
import shutil
import subprocess

import numpy as np
import pytest
from PIL import Image

import oracle
import porousmediaanalysis_amd as pma
from tests import example_fixture as ex
from tests.gpu_common import log
from tests.volume_utils import canonical_relabel, sha256_array

pytestmark = pytest.mark.gpu

PMA_CLI = os.path.join(os.path.dirname(pma.LIB_PATH), "pma")


@pytest.fixture(scope="module")
def ctx():
    with pma.Context(0) as c:
        yield c


def example_folder(tmp_path):
    d = tmp_path / "example"
    d.mkdir()
    shutil.copytree(os.path.join(ex.DIR, "Images"), d / "Images")
    shutil.copyfile(os.path.join(ex.DIR, "config.txt"), d / "config.txt")
    return d


def text(path):
    with open(path, "rb") as f:
        return f.read().decode()


def test_example_all_then_spd(ctx, tmp_path):
    d = example_folder(tmp_path)
    log("folder -all on Docs/Example")
    ctx.run_folder(str(d), pma.MODE_ALL)
    dig = ex.digests()
    assert sha256_array(pma.read_tiff_stack(str(d / "EuclideanDistanceSquares"))) == dig["edt2"]
    assert sha256_array(pma.read_tiff_stack(str(d / "ContainingBallRadiiSquares"))) == dig["cr2"]
    assert text(d / "euclideanDistanceSquareCounts.txt") == ex.text("euclideanDistanceSquareCounts.txt")
    names = sorted(os.listdir(d / "Images"))
    for sub in ["EuclideanDistanceSquares", "ContainingBallRadiiSquares", "ContainingBallCoordinatesX",
                "ContainingBallCoordinatesY", "ContainingBallCoordinatesZ", "IsMaximumBall", "PoreThroatIds",
                "PoreThroatIdsWatershed"]:
        assert sorted(os.listdir(d / sub)) == names, sub
    # the written slices hold the context's volumes (the GPU tests pin those against the oracle)
    for sub, vol in [("IsMaximumBall", pma.VOL_MASK), ("PoreThroatIds", pma.VOL_PT),
                     ("PoreThroatIdsWatershed", pma.VOL_PTW), ("ContainingBallCoordinatesX", pma.VOL_CX)]:
        assert np.array_equal(pma.read_tiff_stack(str(d / sub)), ctx.read(vol)), sub
    for name, t in [("pores.txt", pma.TEXT_PORES), ("sharedPixelTypes.txt", pma.TEXT_THROATS),
                    ("poresWatershed.txt", pma.TEXT_PORES_WS), ("sharedPixelTypesWatershed.txt", pma.TEXT_THROATS_WS),
                    ("intermediateStatistics.txt", pma.TEXT_STATS)]:
        assert text(d / name) == ctx.text(t), name
    assert text(d / "intermediateStatistics.txt") == "Max Euclidean distance square: 2561\nMaximum balls count: 380507\n"
    # an independent decoder reads the same 8-bit mask (MinIsWhite tags kept: PIL inverts them)
    im = Image.open(d / "IsMaximumBall" / names[0])
    assert im.tag_v2[262] == 0 and np.array_equal(255 - np.array(im), ctx.read(pma.VOL_MASK)[0])
    log("folder -spd on Docs/Example")
    ctx.run_folder(str(d), pma.MODE_SPD)
    assert sha256_array(pma.read_tiff_stack(str(d / "ShortestPathDistances")).view(np.float32)) == dig["spd"]
    assert sha256_array(pma.read_tiff_stack(str(d / "CavityIds"))) == dig["cavity"]
    assert text(d / "cavityDescriptions.txt") == ex.text("cavityDescriptions.txt")


def test_example_resume_with_the_fixtures_mask(ctx, tmp_path):
    """-edt, -rib, then the fixture's IsMaximumBall slices replace ours (a folder the reference
    wrote with 7 threads), then -apt: the pore and watershed partitions are the fixture's."""
    d = example_folder(tmp_path)
    log("folder -edt, -rib on Docs/Example")
    ctx.run_folder(str(d), pma.MODE_EDT)
    assert text(d / "intermediateStatistics.txt") == "Max Euclidean distance square: 2561\nMaximum balls count: 0\n"
    ctx.run_folder(str(d), pma.MODE_RIB)
    pma.write_tiff_stack(str(d / "IsMaximumBall"), str(d / "Images"), ex.mask())
    log("folder -apt with the fixture's mask")
    ctx.run_folder(str(d), pma.MODE_APT)
    dig = ex.digests()
    assert sha256_array(canonical_relabel(pma.read_tiff_stack(str(d / "PoreThroatIds")))) == dig["pt_partition"]
    assert sha256_array(canonical_relabel(pma.read_tiff_stack(str(d / "PoreThroatIdsWatershed")))) == \
        dig["ptw_partition"]


def _write_images(d, solid):
    (d / "Images").mkdir(parents=True)
    for z in range(solid.shape[0]):
        Image.fromarray(solid[z]).save(d / "Images" / f"slice{z:02d}.tif", compression="tiff_lzw")


def _config(d, res, d3q7=0):
    (d / "config.txt").write_text(f"Resolution of initial images: {res[0]} {res[1]} {res[2]}\n"
                                  "Available memory in megabytes: 1500\n"
                                  "Max ratio of pore radius to seed radius: 10.0\n"
                                  f"Use D3Q7 stencil: {d3q7}\n"
                                  "Number of threads: 1\n")


@pytest.mark.parametrize("res", [(1, 1, 2), (2, 1, 1), (1.5, 1, 1)])
def test_anisotropic_resolution_resamples(ctx, tmp_path, res):
    """Images/ becomes Original/; the new slices repeat nearest voxels (ImageResampler.cpp)."""
    from tests import volumes
    solid = volumes.spheres((6, 20, 24), 10, 2.0, 4.0, 5)
    d = tmp_path / "aniso"
    _write_images(d, solid)
    _config(d, res)
    log(f"folder -edt with resolution {res}")
    ctx.run_folder(str(d), pma.MODE_EDT)
    Z, X, Y = solid.shape
    rz, rx, ry = res[2], res[0], res[1]
    nz, nx, ny = int(Z * rz + 0.5), int(X * rx + 0.5), int(Y * ry + 0.5)
    iz = (np.arange(nz) / rz).astype(int)
    ix = (np.arange(nx) / rx).astype(int)
    iy = (np.arange(ny) / ry).astype(int)
    expect = solid[iz][:, ix][:, :, iy]
    assert (d / "Original").is_dir() and len(os.listdir(d / "Original")) == Z
    assert np.array_equal(pma.read_tiff_stack(str(d / "Images")), expect)
    if rz == 2:
        assert sorted(os.listdir(d / "Images"))[:2] == ["slice00.tif", "slice00_1.tif"]
    e, m, _ = oracle.edt(expect)
    assert np.array_equal(pma.read_tiff_stack(str(d / "EuclideanDistanceSquares")), e)
    # a second pass does not resample again
    ctx.run_folder(str(d), pma.MODE_SPD)
    assert len(os.listdir(d / "Images")) == nz


def test_console_program(tmp_path):
    """pma -spd over a base folder runs every folder below it that holds config.txt."""
    from tests import volumes
    base = tmp_path / "base"
    for k, sub in enumerate(["a", "b/c"]):
        d = base / sub
        _write_images(d, volumes.spheres((4, 16, 18), 6, 2.0, 3.5, 11 + k))
        _config(d, (1, 1, 1), d3q7=k)
    log("pma -spd")
    r = subprocess.run([PMA_CLI, "-spd", str(base)], capture_output=True, text=True, timeout=120)
    assert r.returncode == 0, r.stderr
    for k, sub in enumerate(["a", "b/c"]):
        solid = pma.read_tiff_stack(str(base / sub / "Images"))
        dist, cav, ctext = oracle.geodesic(solid, bool(k))
        got = pma.read_tiff_stack(str(base / sub / "ShortestPathDistances"))
        assert np.array_equal(got, dist.view(np.uint32))
        assert text(base / sub / "cavityDescriptions.txt") == ctext
    r = subprocess.run([PMA_CLI, "-xyz"], capture_output=True, text=True, timeout=60)
    assert r.returncode == 2 and "-spd" in r.stdout
    r = subprocess.run([PMA_CLI, "-rib", str(base)], capture_output=True, text=True, timeout=60)
    assert r.returncode == 1 and "EuclideanDistanceSquares" in r.stderr

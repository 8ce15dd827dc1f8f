"""Host emulation of the EDT, containing-ball, pore, removal, watershed and geodesic stages.

The library's own .hip sources (scan, libstdc++-sort emulation, prefix trie, EDT, containing balls,
pores, removal, watershed, geodesic) are translated to host C++ (tests/hostemu/translate.py), built with
AddressSanitizer and UndefinedBehaviorSanitizer and run on the oracle's inputs: every kernel
executes its real body, thread by thread (wave kernels as fibers), with exact-size buffers, so
an out-of-bounds access fails the test.  Outputs must equal the oracle bit for bit.  rocPRIM
sorts are replaced by std::stable_sort, the one device helper not emulated here.
"""
import fcntl
import hashlib
import os
import shutil
import subprocess
from concurrent.futures import ThreadPoolExecutor

import numpy as np
import pytest

import oracle
from tests import example_fixture as ex
from tests import removal_inputs, volumes
from tests.volume_utils import canonical_relabel, sha256_array

ROOT = os.path.dirname(os.path.dirname(os.path.abspath(__file__)))
CSRC = os.path.join(ROOT, "porousmediaanalysis_amd", "csrc")
EMU = os.path.join(ROOT, "tests", "hostemu")
SOURCES = ["scan", "sort", "lists", "edt", "balls", "pores", "remove", "watershed", "geodesic"]
SANITIZED = ["-std=c++20", "-O1", "-g", "-fsanitize=address,undefined", "-fno-omit-frame-pointer",
             "-fno-sanitize-recover=undefined", f"-I{EMU}"]
FAST = ["-std=c++20", "-O2", f"-I{EMU}"]
ENV = {"PATH": "/usr/bin:/bin",
       "ASAN_OPTIONS": "detect_leaks=0:detect_stack_use_after_return=0",
       "UBSAN_OPTIONS": "print_stacktrace=1"}


def _digest():
    h = hashlib.sha256()
    for p in [os.path.join(CSRC, f + ".hip") for f in SOURCES] + \
             [os.path.join(EMU, f) for f in ["shim.h", "shim.cpp", "main.cpp", "translate.py"]]:
        with open(p, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def _build(tmp_path_factory, flags, tag):
    if shutil.which("g++") is None:
        pytest.skip("g++ is not available")
    out = tmp_path_factory.getbasetemp().parent / f"pma_hostemu_{tag}_{_digest()}"
    out.mkdir(parents=True, exist_ok=True)
    # pytest-xdist workers share this folder: one builds, the others wait on the lock
    with open(out / "build.lock", "w") as lock:
        fcntl.flock(lock, fcntl.LOCK_EX)
        return _build_locked(out, flags)


def _build_locked(out, flags):
    exe = out / "hostemu"
    if exe.exists():
        return str(exe)
    import sys
    sys.path.insert(0, EMU)
    from translate import translate
    jobs = []
    for f in SOURCES:
        with open(os.path.join(CSRC, f + ".hip")) as src:
            (out / f"{f}.cpp").write_text(translate(src.read()))
        jobs.append(["g++", *flags, "-c", str(out / f"{f}.cpp"), "-o", str(out / f"{f}.o")])
    jobs.append(["g++", *flags, "-c", os.path.join(EMU, "main.cpp"), "-o", str(out / "main.o")])
    jobs.append(["g++", "-c", os.path.join(EMU, "shim.cpp"), "-o", str(out / "shim.o")])
    with ThreadPoolExecutor(max_workers=min(8, os.cpu_count() or 1)) as pool:
        for r in pool.map(lambda cmd: subprocess.run(cmd, capture_output=True, text=True), jobs):
            assert r.returncode == 0, r.stderr[-4000:]
    objs = [str(out / f"{f}.o") for f in SOURCES + ["main", "shim"]]
    link = [f for f in flags if f.startswith("-fsanitize=")]
    r = subprocess.run(["g++", *link, "-o", str(exe) + ".tmp", *objs, "-lpthread"], capture_output=True, text=True)
    assert r.returncode == 0, r.stderr[-4000:]
    os.replace(str(exe) + ".tmp", exe)
    return str(exe)


@pytest.fixture(scope="session")
def hostemu(tmp_path_factory):
    """Sanitized build: every access is bounds-checked."""
    # 64 relaxing waves per launch: each round launches the whole grid, and the emulation runs
    # every workgroup, even those with nothing to do
    return _build(tmp_path_factory, SANITIZED + ["-DPMA_GEO_RELAX_GRID=64u"], "asan")


@pytest.fixture(scope="session")
def hostemu_slabs(tmp_path_factory):
    """Sanitized build with small work limits: containing-ball pair lists cut into slabs of at most
    4000 pairs, 32-bit winner slots, ranks (not packed balls) in the pair sort, far tables read
    from global memory by the centre prune, geodesic relaxation on 2 waves with 32-entry push
    buffers, DFS replay stacks of 8 frames in LDS and 64 in global memory (so frames move between
    them and replays overflow and are repeated), visited voxels on 4 pages of 64 voxels (so pages
    run out and replays are repeated), the per-thread
    merge of pore sets starting from a pool of one entry per plateau (so it overflows and is
    repeated), and skip traversals that give up after listing 3 plateaus."""
    return _build(tmp_path_factory, SANITIZED + ["-DPMA_CB_SLAB_PAIRS=4000ull", "-DPMA_CB_WIN16_MAX=0u",
                                                 "-DPMA_CB_GZ=1", "-DPMA_CB_GX=3", "-DPMA_CB_SMALL_R2=0u",
                                                 "-DPMA_GEO_RELAX_GRID=2u", "-DPMA_GEO_PUSH_BUF=32u",
                                                 "-DPMA_CB_PACKED=0", "-DPMA_CB_PRUNE_FAR_LDS=0u",
                                                 "-DPMA_PORES_STACK_CAP=64ull", "-DPMA_PORES_RING=8u", "-DPMA_UNION_LDS_MAX=0u",
                                                 "-DPMA_PORES_VIS_MIN_BITS=6u", "-DPMA_PORES_VIS_PAGES=4u",
                                                 "-DPMA_PORES_TRAV_CAP=3ull", "-DPMA_PORES_POOL_MIN=1ull",
                                                 "-DPMA_PORES_POOL_MAX=1ull", "-DPMA_PORES_POOL_FLOOR=16ull",
                                                 "-DPMA_PORES_PCACHE=4u", "-DPMA_PORES_VCACHE=4u"], "slabs")


@pytest.fixture(scope="session")
def hostemu_fast(tmp_path_factory):
    """Optimised build without sanitizers, for the larger fixture volume; its containing-ball pair
    sort keeps cells and packed balls apart (the form used where they do not fit one word)."""
    return _build(tmp_path_factory, FAST + ["-DPMA_CB_FUSED=0"], "fast")


def _run(exe, stage, payload, tmp_path, shape):
    inp, vol, txt = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "out.txt"
    inp.write_bytes(payload)
    r = subprocess.run([exe, stage, str(inp), str(vol), str(txt)], capture_output=True, text=True, env=ENV,
                       timeout=600)
    assert r.returncode == 0, f"{stage}: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    pores, throats = txt.read_text().split("@@\n")
    return np.fromfile(vol, np.uint32).reshape(shape), pores, throats


def emu_pores(exe, tmp_path, e, cr, mask, d3q7, order=None):
    Z, X, Y = e.shape
    payload = np.array([X, Y, Z, int(d3q7)], np.int32).tobytes() + e.astype(np.uint32).tobytes() + \
        cr.astype(np.uint32).tobytes() + mask.astype(np.uint8).tobytes()
    if order is not None:
        payload += np.array([len(order)], np.uint64).tobytes() + np.asarray(order, np.uint32).tobytes()
    return _run(exe, "pores", payload, tmp_path, e.shape)


def emu_remove(exe, tmp_path, pt, ptxt, ktxt):
    Z, X, Y = pt.shape
    payload = np.array([X, Y, Z], np.int32).tobytes() + pt.astype(np.uint32).tobytes() + \
        (ptxt.replace("\r", "") + "@@\n" + ktxt.replace("\r", "")).encode()
    return _run(exe, "remove", payload, tmp_path, pt.shape)


def emu_watershed(exe, tmp_path, cr, max_e, ptxt, d3q7):
    Z, X, Y = cr.shape
    rows = [list(map(int, line.split())) for line in ptxt.replace("\r", "").split("\n")[1:] if line.strip()]
    payload = np.array([X, Y, Z, int(d3q7), max_e], np.int32).tobytes() + cr.astype(np.uint32).tobytes() + \
        np.array([len(rows)], np.int32).tobytes() + np.array(rows, np.int32).reshape(-1, 7).tobytes()
    return _run(exe, "watershed", payload, tmp_path, cr.shape)


def emu_balls(exe, tmp_path, e, max_e):
    Z, X, Y = e.shape
    inp, vol, txt = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "out.txt"
    inp.write_bytes(np.array([X, Y, Z, max_e], np.int32).tobytes() + e.astype(np.uint32).tobytes())
    r = subprocess.run([exe, "balls", str(inp), str(vol), str(txt)], capture_output=True, text=True, env=ENV,
                       timeout=600)
    assert r.returncode == 0, f"balls: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    n = e.size
    raw = np.fromfile(vol, np.uint8)
    u32 = raw[:16 * n].view(np.uint32)
    out = [u32[k * n:(k + 1) * n].reshape(e.shape) for k in range(4)]
    out.append(raw[16 * n:].reshape(e.shape))
    out.append(int(txt.read_text()))
    return out


def front(solid):
    e, max_e, _ = oracle.edt(solid, threads=4)
    cr, _, _, _, mask, _ = oracle.containing_balls(e, threads=4)
    return e, max_e, cr, mask


def assert_same(g, o, what):
    assert g[1] == o[1], f"{what}: pore table differs"
    assert g[2] == o[2], f"{what}: throat table differs"
    if not np.array_equal(g[0], o[0]):
        bad = np.argwhere(g[0] != o[0])
        raise AssertionError(f"{what}: {len(bad)} labels differ, first {tuple(bad[0])}")


CASES = dict(volumes.edge_cases())
CASES["spheres_48"] = volumes.spheres((48, 48, 48), 40, 3, 7, 1)
CASES["gauss_s2"] = volumes.gaussian((29, 41, 67), 2.0, 0.65, 2)
CASES["gauss_s3_low"] = volumes.gaussian((40, 40, 40), 3.0, 0.5, 12)
CASES["noise"] = volumes.random_noise((17, 19, 23), 0.15, 4)
CASES["noise_many_pores"] = volumes.random_noise((24, 26, 28), 0.45, 6)


BALL_CASES = dict(CASES)
BALL_CASES["gauss_s4"] = volumes.gaussian((48, 56, 64), 4.0, 0.7, 3)
BALL_CASES["gauss_ragged"] = volumes.gaussian((19, 70, 130), 3.0, 0.75, 3)


def _far_void(Y):
    # void columns 48 voxels wide in x and solid only at z = 0: squared 2-D distances up to 71^2
    # (the virtual planes x = -X and 2X-1), so pass Z searches beyond its staged halo
    s = np.zeros((150, 48, Y), np.uint8)
    s[0] = 1
    s[120, 3, 5] = 1
    return s


EDT_CASES = dict(BALL_CASES)
EDT_CASES["far_void_y12"] = _far_void(12)
EDT_CASES["far_void_y13"] = _far_void(13)
# solid only at y = 0: rows of 100 along y whose nearest solid lies up to 99 away, so pass Y
# searches beyond its LDS padding
EDT_CASES["far_void_rows"] = np.zeros((6, 48, 100), np.uint8)
EDT_CASES["far_void_rows"][:, :, 0] = 1


@pytest.mark.parametrize("name", sorted(EDT_CASES))
def test_edt(hostemu, tmp_path, name):
    solid = EDT_CASES[name]
    Z, X, Y = solid.shape
    inp, vol, txt = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "out.txt"
    inp.write_bytes(np.array([X, Y, Z], np.int32).tobytes() + solid.astype(np.uint8).tobytes())
    r = subprocess.run([hostemu, "edt", str(inp), str(vol), str(txt)], capture_output=True, text=True, env=ENV,
                       timeout=600)
    assert r.returncode == 0, f"edt: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    e = np.fromfile(vol, np.uint32).reshape(solid.shape)
    words = txt.read_text().split()
    oe, om, oh = oracle.edt(solid, threads=4)
    if not np.array_equal(e, oe):
        bad = np.argwhere(e != oe)
        raise AssertionError(f"{len(bad)} squared distances differ, first {tuple(bad[0])}: "
                             f"{e[tuple(bad[0])]} vs {oe[tuple(bad[0])]}")
    assert int(words[0]) == om, "maximum differs"
    assert np.array_equal(np.array(words[1:], np.uint64), oh), "histogram differs"


def check_balls(binary, tmp_path, solid):
    e, max_e, _ = oracle.edt(solid, threads=4)
    g = emu_balls(binary, tmp_path, e, max_e)
    o = oracle.containing_balls(e, threads=4)
    for k, what in enumerate(["CR2", "CX", "CY", "CZ", "mask"]):
        if not np.array_equal(g[k], o[k]):
            bad = np.argwhere(g[k] != o[k])
            raise AssertionError(f"{what}: {len(bad)} voxels differ, first {tuple(bad[0])}: "
                                 f"{g[k][tuple(bad[0])]} vs {o[k][tuple(bad[0])]}")
    assert g[5] == o[5], "maximum-ball count differs"


@pytest.mark.parametrize("name", sorted(BALL_CASES))
def test_containing_balls(hostemu, tmp_path, name):
    check_balls(hostemu, tmp_path, BALL_CASES[name])


@pytest.mark.parametrize("name", ["gauss_s4", "gauss_ragged", "spheres_48", "noise_many_pores"])
def test_containing_balls_separate_keys(hostemu_fast, tmp_path, name):
    check_balls(hostemu_fast, tmp_path, BALL_CASES[name])


GEO_CASES = dict(volumes.edge_cases())
GEO_CASES["spheres_40"] = volumes.spheres((40, 44, 36), 60, 3, 6, 5)
GEO_CASES["gauss"] = volumes.gaussian((33, 47, 52), 2.0, 0.55, 6)
GEO_CASES["noise"] = volumes.random_noise((20, 21, 22), 0.62, 7)


def check_geodesic(binary, tmp_path, solid, d3q7):
    Z, X, Y = solid.shape
    inp, vol, txt = tmp_path / "in.bin", tmp_path / "out.bin", tmp_path / "out.txt"
    inp.write_bytes(np.array([X, Y, Z, int(d3q7)], np.int32).tobytes() + solid.astype(np.uint8).tobytes())
    r = subprocess.run([binary, "geodesic", str(inp), str(vol), str(txt)], capture_output=True, text=True, env=ENV,
                       timeout=600)
    assert r.returncode == 0, f"geodesic: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    raw = np.fromfile(vol, np.uint32)
    dist, cav = raw[:solid.size].reshape(solid.shape), raw[solid.size:].reshape(solid.shape)
    od, oc, ot = oracle.geodesic(solid, d3q7)
    assert np.array_equal(dist, od.view(np.uint32)), "distances differ"
    assert np.array_equal(cav, oc), "cavity ids differ"
    assert txt.read_text() == ot.replace("\r", ""), "cavityDescriptions.txt differs"


@pytest.mark.parametrize("name", sorted(GEO_CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_geodesic(hostemu, tmp_path, name, d3q7):
    check_geodesic(hostemu, tmp_path, GEO_CASES[name], d3q7)


@pytest.mark.parametrize("name", ["gauss", "spheres_40"])
def test_geodesic_small_grid(hostemu_slabs, tmp_path, name):
    """The same distances when each relaxing wave takes many blocks in turn and flushes its
    buffer of activated blocks many times (2 waves per launch, 32-entry buffer)."""
    check_geodesic(hostemu_slabs, tmp_path, GEO_CASES[name], False)


def test_containing_balls_in_slabs(hostemu_slabs, tmp_path):
    """The same result when the (cell, ball) pairs are processed in many z-slabs, with 32-bit winner
    slots, pair cells of 1 x 3 tiles, ranks gathered after the pair sort and the general integer
    square root."""
    solid = BALL_CASES["gauss_s4"]
    e, max_e, _ = oracle.edt(solid, threads=4)
    g = emu_balls(hostemu_slabs, tmp_path, e, max_e)
    o = oracle.containing_balls(e, threads=4)
    for k in range(5):
        assert np.array_equal(g[k], o[k])
    assert g[5] == o[5]


@pytest.mark.parametrize("name", ["gauss_s2", "gauss_s3_low", "noise_many_pores", "spheres_48"])
def test_pores_small_replay_stacks(hostemu_slabs, tmp_path, name):
    """DFS replay stacks of 64 frames: replays overflow, are repeated with room for their whole
    reach, and the labels and tables are still the oracle's."""
    e, max_e, cr, mask = front(CASES[name])
    assert_same(emu_pores(hostemu_slabs, tmp_path, e, cr, mask, False), oracle.pores(mask, e, cr, False), "pores")


@pytest.mark.parametrize("name", sorted(CASES))
@pytest.mark.parametrize("d3q7", [False, True])
def test_pores_removal_watershed(hostemu, tmp_path, name, d3q7):
    e, max_e, cr, mask = front(CASES[name])
    o = oracle.pores(mask, e, cr, d3q7)
    assert_same(emu_pores(hostemu, tmp_path, e, cr, mask, d3q7), o, "pores")
    o2 = oracle.remove_empty_pores(*o)
    assert_same(emu_remove(hostemu, tmp_path, *o), o2, "empty-pore removal")
    ow = oracle.watershed(cr, max_e, o2[1], d3q7)
    assert_same(emu_watershed(hostemu, tmp_path, cr, max_e, o2[1], d3q7), ow, "watershed")


def test_example_partitions(hostemu_fast, tmp_path):
    """Fed the Docs/Example mask, the pore and watershed partitions equal the fixture's."""
    hostemu = hostemu_fast
    solid = ex.solid()
    e, max_e, cr, _ = front(solid)
    o = oracle.pores(ex.mask(), e, cr)
    g = emu_pores(hostemu, tmp_path, e, cr, ex.mask(), False)
    assert_same(g, o, "pores")
    o2 = oracle.remove_empty_pores(*o)
    g2 = emu_remove(hostemu, tmp_path, *o)
    assert_same(g2, o2, "empty-pore removal")
    assert sha256_array(canonical_relabel(g2[0])) == ex.digests()["pt_partition"]
    gw = emu_watershed(hostemu, tmp_path, cr, max_e, o2[1], False)
    assert_same(gw, oracle.watershed(cr, max_e, o2[1], False), "watershed")
    assert sha256_array(canonical_relabel(gw[0])) == ex.digests()["ptw_partition"]


def test_example_bit_exact_in_the_fixture_seed_order(hostemu_fast, tmp_path):
    """With the fixture's seed order replayed, the pore labels and tables, throat-key IDs included,
    are the reference's own files."""
    e, max_e, cr, _ = front(ex.solid())
    g = emu_pores(hostemu_fast, tmp_path, e, cr, ex.mask(), False, ex.ball_order(ex.mask(), e))
    pt, ptxt, ktxt = emu_remove(hostemu_fast, tmp_path, *g)
    assert ptxt == ex.text("pores.txt")
    assert ktxt == ex.text("sharedPixelTypes.txt")
    assert sha256_array(pt) == ex.digests()["pt"]


@pytest.mark.parametrize("d3q7", [False, True])
def test_pores_in_a_given_ball_order(hostemu, tmp_path, d3q7):
    e, max_e, cr, mask = front(CASES["gauss_s3_low"])
    balls = np.flatnonzero(mask.reshape(-1) == 255)
    r2 = e.reshape(-1)[balls].astype(np.int64)
    order = balls[np.lexsort((np.random.default_rng(5).random(balls.size), -r2))].astype(np.uint32)
    assert_same(emu_pores(hostemu, tmp_path, e, cr, mask, d3q7, order), oracle.pores_in_order(e, cr, order, d3q7),
                "pores")


@pytest.mark.parametrize("name", sorted(removal_inputs.CASES))
@pytest.mark.parametrize("d3q7", [False, True])
@pytest.mark.parametrize("every", [2, 3])
def test_remove_crafted(hostemu, tmp_path, name, d3q7, every):
    """Empty-pore removal with some pore volumes set to 0 (pores dropped, throats folded and merged)."""
    pt, ptxt, ktxt = removal_inputs.removal_input(name, d3q7, every)
    o = oracle.remove_empty_pores(pt, ptxt, ktxt)
    assert o[1].count("\n") < ptxt.count("\n")
    assert_same(emu_remove(hostemu, tmp_path, pt, ptxt, ktxt), o, "empty-pore removal")


def test_pores_one_context(hostemu, tmp_path):
    """The pore stage over the first GPU pore test cases in their order (the faulting run's prefix
    and the next case), on one context: scratch buffers keep the previous case's contents, as in
    one process that runs them all.  (The whole sequence, both stencils, also passes: ~10 min.)"""
    from tests import test_gpu_pores as gp

    order = [(False, name) for name in sorted(gp.CASES)][:4]
    lines, expected = [], []
    for k, (d3q7, name) in enumerate(order):
        solid = gp.CASES[name]
        e, max_e, cr, mask = front(solid)
        Z, X, Y = e.shape
        inp = tmp_path / f"in{k}.bin"
        inp.write_bytes(np.array([X, Y, Z, int(d3q7)], np.int32).tobytes() + e.astype(np.uint32).tobytes() +
                        cr.astype(np.uint32).tobytes() + mask.astype(np.uint8).tobytes())
        lines.append(f"{inp} {tmp_path / f'out{k}.bin'} {tmp_path / f'out{k}.txt'}")
        expected.append(oracle.pores(mask, e, cr, d3q7))
    (tmp_path / "list.txt").write_text("\n".join(lines) + "\n")
    r = subprocess.run([hostemu, "pores_seq", str(tmp_path / "list.txt"), "-", "-"], capture_output=True, text=True,
                       env=ENV, timeout=1200)
    assert r.returncode == 0, f"pores_seq: rc {r.returncode}\n{r.stdout[-2000:]}\n{r.stderr[-6000:]}"
    for k, (d3q7, name) in enumerate(order):
        pores, throats = (tmp_path / f"out{k}.txt").read_text().split("@@\n")
        pt = np.fromfile(tmp_path / f"out{k}.bin", np.uint32).reshape(gp.CASES[name].shape)
        assert_same((pt, pores, throats), expected[k], f"pores {name} d3q7={d3q7}")

"""Pins the oracle's EDT and containing balls against brute force and scipy on tiny volumes.

A plausible mistake (a dropped boundary term, a wrong start value, a transposed axis, the
wrong tie-break) changes at least one of these volumes.
"""
import numpy as np
import pytest
from scipy.ndimage import distance_transform_edt

import oracle
from tests import volumes


def brute_edt(solid):
    """min over solids and the virtual planes x = -X and x = 2X-1 of the squared distance."""
    Z, X, Y = solid.shape
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    best = np.minimum((xx + X) ** 2, (2 * X - 1 - xx) ** 2).astype(np.int64)
    for z, x, y in zip(*np.nonzero(solid)):
        best = np.minimum(best, (xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2)
    best[solid != 0] = 0
    return best.astype(np.uint32)


def brute_balls(e):
    """Max r2 over covering balls; the first centre in x, y, z scan order among equals."""
    Z, X, Y = e.shape
    cr = np.zeros(e.shape, np.int64)
    best_scan = np.full(e.shape, np.iinfo(np.int64).max)
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    centres = sorted(zip(*np.nonzero(e)), key=lambda t: (t[1], t[2], t[0]))  # x, y, z scan order
    for z, x, y in centres:
        r2 = int(e[z, x, y])
        cov = ((xx - x) ** 2 + (yy - y) ** 2 + (zz - z) ** 2 <= r2) & (e > 0)
        scan = (x * Y + y) * Z + z
        upd = cov & ((r2 > cr) | ((r2 == cr) & (scan < best_scan)))
        cr[upd] = r2
        best_scan[upd] = scan
    return cr.astype(np.uint32), best_scan


CASES = dict(volumes.edge_cases())
CASES["spheres"] = volumes.spheres((10, 14, 12), 6, 1.5, 3.5, 5)
CASES["noise"] = volumes.random_noise((9, 11, 10), 0.2, 9)


@pytest.mark.parametrize("name", sorted(CASES))
def test_edt_brute_force(name):
    solid = CASES[name]
    e, m, h = oracle.edt(solid)
    assert np.array_equal(e, brute_edt(solid))
    assert m == (int(e.max()) if e.size else 0)


def test_edt_matches_scipy_with_virtual_planes():
    solid = volumes.spheres((12, 16, 14), 8, 1.5, 4.0, 11)
    Z, X, Y = solid.shape
    # pad x with X voids on each side, then a solid plane: planes at x = -X - 1 + 1 ... place
    # the planes exactly at x = -X and x = 2X - 1.
    pad = np.zeros((Z, 3 * X + 1, Y), bool)
    pad[:, X:2 * X, :] = solid != 0
    pad[:, 0, :] = True          # x = -X
    pad[:, 3 * X, :] = True      # x = 2X
    pad[:, 3 * X - 1, :] = True  # x = 2X - 1
    d = distance_transform_edt(~pad)
    ref = np.rint(d[:, X:2 * X, :] ** 2).astype(np.uint32)
    e, _, _ = oracle.edt(solid)
    assert np.array_equal(e, ref)


def test_edt_histogram_counts_only_inner_balls():
    solid = volumes.spheres((14, 15, 16), 10, 1.5, 3.5, 3)
    e, m, h = oracle.edt(solid)
    Z, X, Y = e.shape
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    r = np.sqrt(e.astype(np.float64))
    inner = (e > 0) & ~((r > xx) | (r > yy) | (r > zz) | (xx + r > X - 1) | (yy + r > Y - 1) | (zz + r > Z - 1))
    ref = np.bincount(e[inner].astype(np.int64), minlength=m + 1)
    assert np.array_equal(h, ref)


@pytest.mark.parametrize("name", sorted(CASES))
def test_containing_balls_brute_force(name):
    solid = CASES[name]
    e, _, _ = oracle.edt(solid)
    cr, cx, cy, cz, mask, nmax = oracle.containing_balls(e)
    ref_cr, ref_scan = brute_balls(e)
    assert np.array_equal(cr, ref_cr)
    Z, X, Y = e.shape
    void = e > 0
    scan = (cx.astype(np.int64) * Y + cy) * Z + cz
    assert np.array_equal(scan[void], ref_scan[void])
    centres = set(zip(cz[void], cx[void], cy[void]))
    ref_mask = np.where(void, 127, 0).astype(np.uint8)
    for z, x, y in centres:
        ref_mask[z, x, y] = 255
    assert np.array_equal(mask, ref_mask)
    assert nmax == len(centres)


def test_containing_balls_threads_agree():
    solid = volumes.gaussian((20, 22, 24), 2.0, 0.6, 4)
    e, _, _ = oracle.edt(solid)
    one = oracle.containing_balls(e, threads=1)
    many = oracle.containing_balls(e, threads=4)
    for a, b in zip(one[:5], many[:5]):
        assert np.array_equal(a, b)


# ---------------------------------------------------------------------------------------------
# Pores, removal, watershed and geodesic: brute force on tiny volumes and cases worked out by hand.

from collections import deque  # noqa: E402

from scipy import ndimage  # noqa: E402


def _structure(d3q7):
    return ndimage.generate_binary_structure(3, 1 if d3q7 else 3)


def _offsets(d3q7):
    return [(dz, dx, dy) for dz in (-1, 0, 1) for dx in (-1, 0, 1) for dy in (-1, 0, 1)
            if (dz, dx, dy) != (0, 0, 0) and (not d3q7 or abs(dz) + abs(dx) + abs(dy) == 1)]


def parse_pores(text):
    rows = {}
    for line in text.strip().splitlines()[1:]:
        pid, x, y, z, r2, bnd, vol = map(int, line.split())
        rows[pid] = dict(x=x, y=y, z=z, r2=r2, boundary=bnd, volume=vol)
    return rows


def parse_keys(text):
    rows = {}
    for line in text.strip().splitlines()[1:]:
        kid, vol, members = line.split(";")
        rows[int(kid)] = dict(volume=int(vol), pores=[int(v) for v in members.split()])
    return rows


def pore_id(k):
    return 127 + 255 * (k // 129) + k % 129


def reach_bfs(cr, seed, d3q7):
    """Voxels reachable from seed through void voxels of non-increasing CR (a set of (z, x, y))."""
    Z, X, Y = cr.shape
    seen = {seed}
    todo = deque([seed])
    while todo:
        z, x, y = todo.popleft()
        for dz, dx, dy in _offsets(d3q7):
            n = (z + dz, x + dx, y + dy)
            if not (0 <= n[0] < Z and 0 <= n[1] < X and 0 <= n[2] < Y) or n in seen:
                continue
            if 0 < cr[n] <= cr[z, x, y]:
                seen.add(n)
                todo.append(n)
    return seen


PORE_CASES = {
    # void = union of balls in solid, so that there are several pores and throats
    "void_balls": (volumes.spheres((10, 14, 12), 14, 1.5, 3.0, 5) == 0).astype(np.uint8) * 255,
    "void_balls_b": (volumes.spheres((12, 12, 12), 18, 1.5, 2.5, 17) == 0).astype(np.uint8) * 255,
    "noise": volumes.random_noise((10, 11, 12), 0.4, 9),
    "gauss": volumes.gaussian((12, 12, 12), 1.5, 0.6, 3),
}


def _cb(solid):
    e, max_e, _ = oracle.edt(solid)
    cr, _, _, _, mask, _ = oracle.containing_balls(e)
    return e, max_e, cr, mask


@pytest.mark.parametrize("d3q7", [False, True])
@pytest.mark.parametrize("name", sorted(PORE_CASES))
def test_pore_seeds_are_first_sorted_balls_of_regional_max_plateaus(name, d3q7):
    """Seeds = one per regional-maximum CR plateau (plateaus labelled by scipy), each the first
    maximum ball of its plateau in libstdc++ std::sort order; pores numbered in that order."""
    e, _, cr, mask = _cb(PORE_CASES[name])
    _, ptxt, _ = oracle.pores(mask, e, cr, d3q7)
    pores = parse_pores(ptxt)
    st = _structure(d3q7)
    plateau = np.zeros(cr.shape, np.int64)
    nlab = 0
    for v in np.unique(cr[cr > 0]):
        lab, k = ndimage.label(cr == v, st)
        plateau[lab > 0] = lab[lab > 0] + nlab
        nlab += k
    nbmax = ndimage.maximum_filter(cr, footprint=st, mode="constant", cval=0)
    regional = {q for q in range(1, nlab + 1) if np.all(nbmax[plateau == q] == cr[plateau == q])}
    idx = np.flatnonzero(mask.ravel() == 255)
    rec = (e.ravel()[idx].astype(np.uint64) << np.uint64(32)) | idx.astype(np.uint64)
    order = oracle.sort_desc(rec)  # libstdc++ std::sort, the reference's own dependency
    seeds, taken = [], set()
    for r in order:
        i = int(r & np.uint64(0xFFFFFFFF))
        q = int(plateau.ravel()[i])
        if q in regional and q not in taken:
            taken.add(q)
            seeds.append(np.unravel_index(i, cr.shape))
    assert taken == regional
    assert len(pores) == len(seeds)
    for k, (z, x, y) in enumerate(seeds):
        p = pores[pore_id(k)]
        assert (p["z"], p["x"], p["y"]) == (z, x, y)
        assert p["r2"] == cr[z, x, y] == e[z, x, y]


@pytest.mark.parametrize("d3q7", [False, True])
@pytest.mark.parametrize("name", sorted(PORE_CASES))
def test_pore_sets_volumes_and_boundaries_by_reachability(name, d3q7):
    """Each voxel's label holds exactly the pores whose seed reaches it (BFS over non-increasing
    CR), listed [second, first, third, ...]; volumes count labels; a pore is a boundary pore when
    its reach touches the image border."""
    e, _, cr, mask = _cb(PORE_CASES[name])
    pt, ptxt, ktxt = oracle.pores(mask, e, cr, d3q7)
    pores, keys = parse_pores(ptxt), parse_keys(ktxt)
    ids = sorted(pores)
    sets = {}
    Z, X, Y = cr.shape
    for k, pid in enumerate(ids):
        p = pores[pid]
        reach = reach_bfs(cr, (p["z"], p["x"], p["y"]), d3q7)
        for v in reach:
            sets.setdefault(v, []).append(pid)
        border = any(z in (0, Z - 1) or x in (0, X - 1) or y in (0, Y - 1) for z, x, y in reach)
        assert p["boundary"] == int(border), pid
    for z, x, y in zip(*np.nonzero(cr)):
        members = sets[(z, x, y)]
        label = int(pt[z, x, y])
        if len(members) == 1:
            assert label == members[0]
        else:
            assert keys[label]["pores"] == [members[1], members[0]] + members[2:]
    assert np.all(pt[cr == 0] == 0)
    for pid, p in pores.items():
        assert p["volume"] == int(np.sum(pt == pid))
    for kid, kk in keys.items():
        assert kk["volume"] == int(np.sum(pt == kid)) > 0


@pytest.mark.parametrize("d3q7", [False, True])
def test_pore_keys_numbered_in_dfs_order_worked_by_hand(d3q7):
    """A row of 9 voxels, CR 36 9 4 9 16 9 4 9 25: pores 127 (y 0), 128 (y 8), 129 (y 4).  Pore
    129's DFS tries y + 1 before y - 1 (neighbours pushed in x, y, z loop order, last pushed first),
    so it meets pore 128's voxel at y 6 first: key UINT_MAX - 1 = [129, 128], then UINT_MAX - 2 =
    [129, 127] at y 2."""
    cr = np.array([36, 9, 4, 9, 16, 9, 4, 9, 25], np.uint32).reshape(1, 1, 9)
    mask = np.full(cr.shape, 127, np.uint8)
    mask[0, 0, [0, 4, 8]] = 255
    pt, ptxt, ktxt = oracle.pores(mask, cr.copy(), cr, d3q7)
    assert pt.ravel().tolist() == [127, 127, 4294967293, 129, 129, 129, 4294967294, 128, 128]
    pores = parse_pores(ptxt)
    assert [(pid, p["y"], p["r2"], p["volume"], p["boundary"]) for pid, p in sorted(pores.items())] == \
        [(127, 0, 36, 2, 1), (128, 8, 25, 2, 1), (129, 4, 16, 3, 1)]
    assert parse_keys(ktxt) == {4294967293: dict(volume=1, pores=[129, 127]),
                                4294967294: dict(volume=1, pores=[129, 128])}


def test_remove_empty_pores_worked_by_hand():
    """Pores 128 and 130 are empty.  Keys: ...91 [127 129] and ...90 [129 130 127] become
    duplicates of ...94 [129 127] once sorted and stripped, and merge into the largest ID ...94
    (its volume is not summed); ...93 [128 127] folds into pore 127; ...92 [130 128] loses both
    pores and is dropped without a remap, so its voxels keep their label."""
    ptxt = ("PoreId SeedX SeedY SeedZ SeedRadiusSquare IsBoundaryPore Volume\n"
            "127 0 0 0 9 0 5\n128 0 1 0 4 0 0\n129 0 2 0 4 1 3\n130 0 3 0 1 0 0\n")
    ktxt = ("SharedPixelTypeId; Volume; PoreId1 PoreId2 ... \n"
            "4294967290; 2; 129 130 127\n4294967291; 4; 127 129\n4294967292; 1; 130 128\n"
            "4294967293; 1; 128 127\n4294967294; 2; 129 127\n")
    pt = np.array([0, 127, 129, 4294967291, 4294967292, 4294967293, 4294967294, 4294967290, 127, 0],
                  np.uint32).reshape(1, 1, 10)
    out, p2, k2 = oracle.remove_empty_pores(pt.copy(), ptxt, ktxt)
    assert out.ravel().tolist() == [0, 127, 129, 4294967294, 4294967292, 127, 4294967294, 4294967294, 127, 0]
    assert sorted(parse_pores(p2)) == [127, 129]
    assert parse_pores(p2)[129] == dict(x=0, y=2, z=0, r2=4, boundary=1, volume=3)
    assert parse_keys(k2) == {4294967294: dict(volume=2, pores=[127, 129])}


@pytest.mark.parametrize("d3q7", [False, True])
def test_watershed_fifo_worked_by_hand(d3q7):
    """A row of 7 voxels, CR 9 4 1 1 1 4 9, seeds 127 (y 0) and 128 (y 6) at depth 9.  The FIFOs
    give y 1, y 5 (depth 4), then y 2, y 4, then y 3 claimed by 127 (pushed after 128's y 4 at
    depth 1).  128 at y 4 then meets y 3: key UINT_MAX - 1 = [128, 127]; 127 at y 3 keeps
    propagating and meets y 4: key UINT_MAX - 2 = [127, 128].  Volumes 3 and 2."""
    cr = np.array([9, 4, 1, 1, 1, 4, 9], np.uint32).reshape(1, 1, 7)
    ptxt = ("PoreId SeedX SeedY SeedZ SeedRadiusSquare IsBoundaryPore Volume\n"
            "127 0 0 0 9 1 3\n128 0 6 0 9 1 3\n")
    ptw, p2, k2 = oracle.watershed(cr, 9, ptxt, d3q7)
    assert ptw.ravel().tolist() == [127, 127, 127, 4294967294, 4294967293, 128, 128]
    pores = parse_pores(p2)
    assert (pores[127]["volume"], pores[128]["volume"]) == (3, 2)
    assert parse_keys(k2) == {4294967293: dict(volume=1, pores=[127, 128]),
                              4294967294: dict(volume=1, pores=[128, 127])}


def _bellman_ford_f32(solid, start, d3q7):
    """float32 label correcting from start until nothing changes: d(v) = min(d(v), fl32(d(u) + w))."""
    Z, X, Y = solid.shape
    inf = np.float32(np.finfo(np.float32).max)
    d = np.full(solid.shape, inf, np.float32)
    d[start] = np.float32(0)
    pad = np.full((Z + 2, X + 2, Y + 2), inf, np.float32)
    void = solid == 0
    while True:
        pad[1:-1, 1:-1, 1:-1] = d
        best = d.copy()
        for dz, dx, dy in _offsets(d3q7):
            w = np.float32(1.0) if d3q7 else np.sqrt(np.float32(abs(dz) + abs(dx) + abs(dy)))
            nb = pad[1 + dz:1 + dz + Z, 1 + dx:1 + dx + X, 1 + dy:1 + dy + Y]
            cand = np.where(nb < inf, (nb + w).astype(np.float32), inf)
            best = np.minimum(best, cand)
        best[~void] = inf
        if np.array_equal(best, d):
            return d
        d = best


def _shelled(shape, p_solid, seed, radii):
    """Noise with solid spherical shells around the centre: the centre cavity reaches no wall, so
    the geodesic restarts from box shells."""
    solid = volumes.random_noise(shape, p_solid, seed)
    Z, X, Y = shape
    zz, xx, yy = np.meshgrid(np.arange(Z), np.arange(X), np.arange(Y), indexing="ij")
    c = [min(int(0.5 * n + 0.5), n - 1) for n in (X, Y, Z)]
    r = np.sqrt((xx - c[0]) ** 2 + (yy - c[1]) ** 2 + (zz - c[2]) ** 2)
    for r0 in radii:
        solid[(r >= r0) & (r < r0 + 1.5)] = 255
    return solid


GEO_CASES = {
    "noise": volumes.random_noise((8, 10, 9), 0.3, 4),
    "noise_fragmented": volumes.random_noise((10, 10, 10), 0.72, 21),
    "shell": _shelled((11, 12, 13), 0.2, 5, [2.5]),
    "two_shells": _shelled((14, 14, 14), 0.15, 6, [2.0, 4.5]),
}


@pytest.mark.parametrize("d3q7", [False, True])
@pytest.mark.parametrize("name", sorted(GEO_CASES))
def test_geodesic_equals_float32_bellman_ford(name, d3q7):
    """Distances of every cavity equal float32 Bellman-Ford from that cavity's seed (a float64
    Dijkstra would not match); cavity ids are 127 k; solids -1 / 0, unreached FLT_MAX / UINT_MAX."""
    solid = GEO_CASES[name]
    dist, cav, text = oracle.geodesic(solid, d3q7)
    rows = [line.split() for line in text.strip().splitlines()[1:]]
    seen = np.zeros(solid.shape, bool)
    for k, row in enumerate(rows):
        cid, x, y, z = int(row[0]), int(row[1]), int(row[2]), int(row[3])
        assert cid == 127 * (k + 1)
        bf = _bellman_ford_f32(solid, (z, x, y), d3q7)
        mine = bf < np.finfo(np.float32).max
        assert not np.any(mine & seen)
        assert np.array_equal(dist[mine].view(np.uint32), bf[mine].view(np.uint32))
        assert np.all(cav[mine] == cid) and int(row[5]) == int(mine.sum())
        seen |= mine
    assert np.all(dist[solid != 0] == -1) and np.all(cav[solid != 0] == 0)
    rest = (solid == 0) & ~seen
    assert np.all(dist[rest] == np.finfo(np.float32).max) and np.all(cav[rest] == 0xFFFFFFFF)


def test_geodesic_start_on_box_shell_worked_by_hand():
    """5 x 5 x 5 solid but (x 1, y 3, z 3).  Centre (3, 3, 3) is solid; shell 1 has no void;
    shell 2 scans x = 1 first (y, z from 1) and finds (1, 3, 3).  It reaches no wall and shell 3
    is not searched (the search stops below the largest half size, 3): one cavity."""
    solid = np.ones((5, 5, 5), np.uint8)
    solid[3, 1, 3] = 0  # [z][x][y]
    dist, cav, text = oracle.geodesic(solid, False)
    assert text.strip().splitlines()[1:] == ["127 1 3 3 2.000000 1 0 0"]
    assert cav[3, 1, 3] == 127 and dist[3, 1, 3] == 0


@pytest.mark.parametrize("d3q7", [False, True])
def test_geodesic_restart_in_split_volume_worked_by_hand(d3q7):
    """7 x 7 x 7 with a solid plane x = 3.  The centre (4, 4, 4) seeds cavity 127, which reaches
    5 walls (not x = 0); the restart scans shells from half size 2, whose first void voxel not yet
    reached is (2, 2, 2): cavity 254, 5 walls.  Shell 3 finds nothing more."""
    solid = np.zeros((7, 7, 7), np.uint8)
    solid[:, 3, :] = 255
    dist, cav, text = oracle.geodesic(solid, d3q7)
    assert text.strip().splitlines()[1:] == ["127 4 4 4 0.000000 147 5 0", "254 2 2 2 3.464102 147 5 0"]
    assert np.all(cav[:, 4:, :] == 127) and np.all(cav[:, :3, :] == 254)

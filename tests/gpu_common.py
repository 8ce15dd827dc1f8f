"""Shared helpers of the GPU parity tests (these tests call the C ABI through ctypes)."""
import sys

import numpy as np

import oracle
import porousmediaanalysis_amd as pma


def log(msg):
    print(msg, flush=True)
    sys.stdout.flush()


def run_front(ctx, solid):
    """GPU EDT + containing balls on a host volume; returns a dict of results."""
    ctx.load(pma.VOL_SOLID, solid)
    ctx.edt()
    ctx.containing_balls()
    m, n = ctx.statistics()
    return dict(e=ctx.read(pma.VOL_EDT2), max_e=m, hist=ctx.histogram(), cr=ctx.read(pma.VOL_CR2),
                cx=ctx.read(pma.VOL_CX), cy=ctx.read(pma.VOL_CY), cz=ctx.read(pma.VOL_CZ),
                mask=ctx.read(pma.VOL_MASK), n_max=n)


def oracle_front(solid):
    e, m, h = oracle.edt(solid, threads=4)
    cr, cx, cy, cz, mask, n = oracle.containing_balls(e, threads=4)
    return dict(e=e, max_e=m, hist=h, cr=cr, cx=cx, cy=cy, cz=cz, mask=mask, n_max=n)


def assert_front_equal(g, o):
    assert g["max_e"] == o["max_e"]
    assert np.array_equal(g["hist"], o["hist"])
    for k in ["e", "cr", "cx", "cy", "cz", "mask"]:
        if not np.array_equal(g[k], o[k]):
            bad = np.argwhere(g[k] != o[k])
            raise AssertionError(f"{k}: {len(bad)} voxels differ, first at (z,x,y)={tuple(bad[0])}: "
                                 f"gpu {g[k][tuple(bad[0])]} oracle {o[k][tuple(bad[0])]}")
    assert g["n_max"] == o["n_max"]

"""Inputs for the empty-pore removal stage that make it do work.

With the maximum-ball masks of these volumes no pore ends up empty, so the removal would only
copy its input.  These cases take the oracle's pore stage output and set the Volume column of
every k-th pore row to 0, as the reference's remover reads it from pores.txt: those pores are
dropped, throats left with one pore are folded into it and throats with equal lists merge.
"""
import oracle
from tests import volumes


def zero_volumes(pores_text, every):
    lines = pores_text.replace("\r", "").split("\n")
    out = [lines[0]]
    k = 0
    for line in lines[1:]:
        if line.strip():
            f = line.split()
            if k % every == every - 1:
                f[-1] = "0"
            k += 1
            line = " ".join(f)
        out.append(line)
    return "\n".join(out)


CASES = {
    "gauss_s4": volumes.gaussian((48, 56, 64), 4.0, 0.7, 3),
    "gauss_s3_low": volumes.gaussian((40, 40, 40), 3.0, 0.5, 12),
    "spheres_dense": volumes.spheres((40, 52, 44), 90, 2, 6, 11),
    "noise_many_pores": volumes.random_noise((40, 48, 56), 0.45, 6),
}


def removal_input(name, d3q7, every):
    """(PT volume, pores.txt with some volumes zeroed, sharedPixelTypes.txt)."""
    solid = CASES[name]
    e, max_e, _ = oracle.edt(solid, threads=4)
    cr, cx, cy, cz, mask, nmax = oracle.containing_balls(e, threads=4)
    pt, ptxt, ktxt = oracle.pores(mask, e, cr, d3q7)
    return pt, zero_volumes(ptxt, every), ktxt

"""Benchmark of the pore-throat + geodesic pipeline on one MI355X.

One step = the stages of the reference's -all mode and -spd that run on the GPU, on one
synthetic volume resident in HBM: EDT -> containing balls -> geodesic distances with
cavities.  The pore-propagation stage (and with it empty-pore removal and the watershed, which
consume its pores) is switched off on the GPU (see the pores.hip gate), so it is not in the
step.  Default workload: 1024^3 voxels, i.i.d. N(0,1) noise blurred with sigma = 8 (truncated
at 4 sigma), void = lowest 65 % of the field (porosity 0.65), generated on the GPU (not
timed).

Prints one JSON line:
  value      voxels per second of whole steps (input resident in HBM before timing),
  roofline   the dominant instrumented kernel: algorithmic bytes per launch / its average
             launch time from HIP events recorded on the context's stream in the timed
             region, against the MI355X HBM peak of 8 TB/s,
  cpu_baseline  the CPU oracle (a restatement of the reference with its OpenMP threading in
             the EDT and containing-ball loops, serial Dijkstra for the distances) timed on the
             same stages over a bounded sample of the same generator (--cpu-size^3).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 1024] [--no-cpu]
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxels/s, 1024³ EDT→watershed→geodesic; batched & surrogate vs plain at MAE"
HBM_PEAK_GBS = 8000.0

# Algorithmic (compulsory) bytes per voxel of each instrumented kernel that sweeps the whole
# volume once per step (in one launch, or in z-slabs over several): each input read once,
# each output written once.
KERNEL_BYTES_PER_VOXEL = {
    "edt_pass_x": 1 + 4,        # u8 solid -> u32 1-D distance square
    "edt_pass_y": 4 + 4,        # u32 -> u32
    "edt_pass_z": 4 + 4,        # u32 -> u32 EDT square (+ histogram)
    "prune_centres": 4 + 1,     # u32 EDT -> u8 mask / candidate flag
    "cover_tiles": 4 + 16,      # u32 EDT -> u32 CR, CX, CY, CZ (over all the launches of a step)
    "count_equal_u8": 1,
    "cc_union": 4 + 4,          # u32 CR -> u32 plateau parent
    "cc_not_max": 4 + 4,
    "plateau_edges": 4 + 4,
    "write_labels": 4 + 4,      # u32 plateau index -> u32 PoreThroatIds
    "geo_init": 1 + 4 + 4,       # u8 solid -> f32 distance, u32 cavity id
}
# Per block visit of one relax_blocks launch: the 8^3 distances read and written once (the
# one-voxel halo is shared with the neighbouring blocks and not counted).
RELAX_BYTES_PER_BLOCK = 512 * (4 + 4)
STAGES = ("edt", "containing_balls", "geodesic")


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--sigma", type=float, default=8.0)
    ap.add_argument("--porosity", type=float, default=0.65)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-size", type=int, default=128)
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stencil", choices=["d3q27", "d3q7"], default="d3q27")
    return ap.parse_args()


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_baseline(args):
    """Oracle EDT, containing balls and geodesic distances on a bounded sample, timed here."""
    import oracle
    import porousmediaanalysis_amd as pma

    n = args.cpu_size
    with pma.Context(0) as ctx:
        ctx.generate_gaussian((n, n, n), args.sigma, args.porosity, args.seed + 1)
        solid = ctx.read(pma.VOL_SOLID)
    threads = args.cpu_threads or min(16, os.cpu_count() or 1)
    progress(f"cpu baseline: oracle EDT, containing balls, geodesic on {n}^3 with {threads} threads")
    t0 = time.perf_counter()
    e, _, _ = oracle.edt(solid, threads)
    oracle.containing_balls(e, threads)
    oracle.geodesic(solid, args.stencil == "d3q7")
    dt = time.perf_counter() - t0
    return {"value": n ** 3 / dt, "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"{n}^3 voxels of the same Gaussian-field generator (sigma={args.sigma:g}, "
                      f"porosity={args.porosity:g}): EDT + containing balls (OpenMP) + geodesic (serial "
                      f"Dijkstra), {dt:.1f} s"}


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import porousmediaanalysis_amd as pma

    S = args.size
    shape = (S, S, S)
    nvox = S ** 3
    with pma.Context(0) as ctx:
        progress(f"generating {S}^3 Gaussian field sigma={args.sigma} porosity={args.porosity}")
        ctx.generate_gaussian(shape, args.sigma, args.porosity, args.seed)
        ctx.set_stencil(args.stencil == "d3q7")
        def step():
            ctx.edt()
            ctx.containing_balls()
            ctx.geodesic()

        for w in range(args.warmup):
            t0 = time.perf_counter()
            step()
            progress(f"warmup step {w + 1}/{args.warmup}: {time.perf_counter() - t0:.3f} s")
        pma.lib().pma_set_kernel_timing(ctx._h, 1)
        stage_ms = {name: 0.0 for name in STAGES}
        stage_items = {name: 0 for name in STAGES}
        t0 = time.perf_counter()
        for s in range(args.steps):
            ts = time.perf_counter()
            step()
            for k, name in enumerate(pma.STAGE_NAMES):
                if name in stage_ms:
                    info = ctx.stage_info(k)
                    stage_ms[name] += info["device_ms"]
                    stage_items[name] += info["work_items"]
            progress(f"step {s + 1}/{args.steps}: {time.perf_counter() - ts:.3f} s")
        elapsed = time.perf_counter() - t0
        kernels = kernel_times(pma, ctx)
        pma.lib().pma_set_kernel_timing(ctx._h, 0)
        max_e, n_max = ctx.statistics()

    ms_per_step = 1000.0 * elapsed / args.steps
    dom = max(kernels.items(), key=lambda kv: kv[1][0]) if kernels else None
    roofline = None
    if dom:
        name, (tot_ms, launches) = dom
        avg_ms = tot_ms / max(launches, 1)
        bpv = KERNEL_BYTES_PER_VOXEL.get(name)
        per_launch = bpv * nvox * args.steps / launches if bpv is not None and launches else None
        if name == "relax_blocks" and launches:
            per_launch = RELAX_BYTES_PER_BLOCK * stage_items["geodesic"] / launches
        if per_launch is not None:
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4),
                        "traffic": measured_traffic(name, S),
                        "bytes_per_launch": round(per_launch), "avg_launch_ms": round(avg_ms, 4)}
        else:
            roofline = {"bound": "hbm", "kernel": name, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": None, "traffic": None, "avg_launch_ms": round(avg_ms, 4), "launches": launches}
    # algorithmic bytes of the step: solid in (1 B), EDT2, CR2, CX, CY, CZ, mask, distance and
    # cavity id out (4 + 16 + 1 + 8 B) per voxel
    pipeline_gbs = 30.0 * nvox / (ms_per_step * 1e-3) / 1e9
    result = {
        "metric": METRIC,
        "value": nvox * args.steps / elapsed,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(ms_per_step, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+f32",
        "data": "synthetic",
        "config": {"workload": f"{S}^3 Gaussian field sigma={args.sigma:g} porosity={args.porosity:g} "
                               f"({args.stencil}): EDT, containing balls, geodesic distances with cavities",
                   "stages": list(STAGES),
                   "not_in_step": "pore propagation, empty-pore removal, watershed (pore stage switched off on "
                                  "the GPU after an unexplained illegal memory access)",
                   "size": S, "sigma": args.sigma, "porosity": args.porosity,
                   "seed": args.seed, "max_edt2": max_e, "maximum_balls": n_max},
        "stages_ms": {k: round(v / args.steps, 3) for k, v in stage_ms.items()},
        # EDT: voxels; containing balls: (ball, tile) pairs; geodesic: block relaxations
        "stage_work_items": {k: v // args.steps for k, v in stage_items.items()},
        "kernels_ms": {k: {"total_ms_per_step": round(v[0] / args.steps, 3), "launches_per_step": v[1] / args.steps}
                       for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][0])},
        "pipeline_algorithmic_GBs": round(pipeline_gbs, 1),
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args)
    print(json.dumps(result))


def measured_traffic(kernel, size):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if size != 1024 or not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(kernel)
    return round(entry["traffic_bytes"]) if entry else None


def kernel_times(pma, ctx):
    import ctypes

    L = pma.lib()
    L.pma_list_kernel_times.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.c_size_t,
                                        ctypes.POINTER(ctypes.c_size_t)]
    L.pma_get_kernel_time.argtypes = [ctypes.c_void_p, ctypes.c_char_p, ctypes.POINTER(ctypes.c_double),
                                      ctypes.POINTER(ctypes.c_uint64)]
    n = ctypes.c_size_t()
    L.pma_list_kernel_times(ctx._h, None, 0, ctypes.byref(n))
    buf = ctypes.create_string_buffer(n.value)
    L.pma_list_kernel_times(ctx._h, buf, n.value, ctypes.byref(n))
    out = {}
    for name in buf.value.decode().split("\n"):
        if not name:
            continue
        ms, cnt = ctypes.c_double(), ctypes.c_uint64()
        L.pma_get_kernel_time(ctx._h, name.encode(), ctypes.byref(ms), ctypes.byref(cnt))
        out[name] = (ms.value, cnt.value)
    return out


if __name__ == "__main__":
    main()

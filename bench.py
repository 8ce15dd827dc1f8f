"""Benchmark of the pore-throat + geodesic pipeline on one MI355X.

One step = the reference's -all mode followed by -spd, all on the GPU, on one synthetic
volume resident in HBM: EDT -> containing balls -> pore seeds and pore propagation ->
empty-pore removal -> watershed -> geodesic distances with cavities.  Default workload:
1024^3 voxels, i.i.d. N(0,1) noise blurred with sigma = 8 (truncated at 4 sigma), void = lowest
65 % of the field (porosity 0.65), generated on the GPU (not timed).

Prints one JSON line:
  value      voxels per second of whole steps (input resident in HBM before timing),
  stages_ms  device time of each stage (HIP events around the stage on the context stream),
  stage_roofline  per stage: compulsory bytes (each input read once, each output written
             once, per voxel) / stage time, against the MI355X HBM peak of 8 TB/s,
  roofline   the dominant instrumented kernel: algorithmic bytes per launch / its average
             launch time from HIP events recorded on the context's stream, in one extra step
             after the timed ones (the timed steps run without per-kernel events),
  cpu_baseline  the CPU oracle (a restatement of the reference: OpenMP in the EDT and
             containing-ball loops as in the reference, serial elsewhere) running all six
             stages on bounded samples of the same generators.

With --workload batched: 256 subvolumes, each side drawn from U{48..192}, half sphere unions
(round(300 V / 128^3) spheres, radii in [6, 12]) and half Gaussian fields (sigma in [2, 6],
porosity in [0.5, 0.8]), host-resident; the whole pipeline on each, once as a per-volume loop on
one context and once through pma_run_batch with the configuration pma_tune_batch picks on a
sample of 24 of them.  value = voxels per second of the batch run (fresh contexts, uploads and
allocations included in both timings).

With --workload surrogate: the screening network trained on --train-crops crops of --crop^3
(half sphere unions, half Gaussian fields) labelled by the exact pipeline, then run over
--test-crops fresh crops; the exact stages rerun on every 20th crop give the radius and ratio MAE
and max error, and a sampled crop whose radius MAE exceeds --bound falls back to its exact fields.
value = voxels per second of the screening (EDT, network, the sampled exact reruns).

The plain run's line also carries "batched" (64 subvolumes) and "surrogate" (96 training and 120
fresh 64^3 crops) objects from reduced runs of those workloads (--no-extras skips them).

Usage: python bench.py [--gpus 1] [--steps K] [--warmup W] [--size 1024] [--no-cpu] [--no-extras]
                       [--workload plain|batched|surrogate]
"""
import argparse
import json
import os
import platform
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

METRIC = "voxels/s, 1024³ EDT→watershed→geodesic; batched & surrogate vs plain at MAE"
HBM_PEAK_GBS = 8000.0

STAGES = ("edt", "containing_balls", "pores", "remove_empty_pores", "watershed", "geodesic")
# Compulsory bytes per voxel of each stage (inputs read once, outputs written once):
#   EDT: u8 solid in, u32 EDT square out; containing balls: EDT in, CR, CX, CY, CZ (u32) and
#   the u8 mask out; pores: mask, EDT and CR in for the seeds (9), CR in and PT out for the
#   propagation (8); removal: PT in and out; watershed: CR in, PTW out; geodesic: solid in,
#   f32 distance and u32 cavity id out.
STAGE_BYTES_PER_VOXEL = {"edt": 5, "containing_balls": 21, "pores": 17, "remove_empty_pores": 8,
                         "watershed": 8, "geodesic": 9}
# Instrumented kernels that sweep the whole volume once per launch.
KERNEL_BYTES_PER_VOXEL = {
    "edt_pass_x": 1 + 4,        # u8 solid -> u32 1-D distance square
    "edt_pass_y": 4 + 4,        # u32 -> u32
    "edt_pass_z": 4 + 4,        # u32 -> u32 EDT square (+ histogram)
    "prune_centres": 4 + 1,     # u32 EDT -> u8 candidate flag
    "cover_tiles": 4 + 16 + 1,  # u32 EDT -> u32 CR, CX, CY, CZ and the u8 mask
    "count_equal_u8": 1,
    "cc_union": 4 + 4,          # u32 CR -> u32 plateau parent
    "cc_not_max": 4 + 4,
    "plateau_edges": 4 + 4,
    "write_labels": 4 + 4,      # u32 plateau index -> u32 PoreThroatIds
}
# Kernels that carry a whole stage: the stage's compulsory bytes per voxel, spread over the
# kernel's launches in a step.  dfs_replay is the pore propagation (CR in, PT out: 8 B/voxel);
# ws_resolve does the watershed's claims and labels (CR in, PTW out: 8 B/voxel).
STAGE_KERNEL_BYTES_PER_VOXEL = {"dfs_replay": 8, "ws_resolve": 8}
# Per block visit of one relax_blocks launch: the 8^3 distances read and written once (the
# one-voxel halo is shared with the neighbouring blocks and not counted).
RELAX_BYTES_PER_BLOCK = 512 * (4 + 4)


def parse():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=3)
    ap.add_argument("--warmup", type=int, default=1)
    ap.add_argument("--size", type=int, default=1024)
    ap.add_argument("--sigma", type=float, default=8.0)
    ap.add_argument("--porosity", type=float, default=0.65)
    ap.add_argument("--seed", type=int, default=20261203)
    ap.add_argument("--cpu-sizes", type=int, nargs="+", default=[128, 192])
    ap.add_argument("--cpu-cap", type=float, default=30.0, help="wall seconds per CPU stage")
    ap.add_argument("--cpu-threads", type=int, default=0)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--stencil", choices=["d3q27", "d3q7"], default="d3q27")
    ap.add_argument("--workload", choices=["plain", "batched", "surrogate"], default="plain")
    ap.add_argument("--batch-count", type=int, default=256)
    ap.add_argument("--crop", type=int, default=64, help="surrogate crop side")
    ap.add_argument("--train-crops", type=int, default=192)
    ap.add_argument("--test-crops", type=int, default=256)
    ap.add_argument("--epochs", type=int, default=8)
    ap.add_argument("--bound", type=float, default=3.0, help="radius MAE bound (voxels) of a sampled crop")
    ap.add_argument("--no-extras", action="store_true",
                    help="plain workload: skip the reduced batched and screening runs added to its line")
    return ap.parse_args()


def progress(msg):
    print(f"[bench] {msg}", file=sys.stderr, flush=True)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return platform.processor() or "unknown"


def oracle_stages(oracle, solid, d3q7, threads):
    """All six oracle stages on one volume; returns per-stage wall seconds."""
    t = {}
    t0 = time.perf_counter()
    e, max_e, _ = oracle.edt(solid, threads)
    t["edt"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    cr, _, _, _, mask, _ = oracle.containing_balls(e, threads)
    t["containing_balls"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    pt, ptxt, ktxt = oracle.pores(mask, e, cr, d3q7)
    t["pores"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    pt, ptxt, ktxt = oracle.remove_empty_pores(pt, ptxt, ktxt)
    t["remove_empty_pores"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.watershed(cr, max_e, ptxt, d3q7)
    t["watershed"] = time.perf_counter() - t0
    t0 = time.perf_counter()
    oracle.geodesic(solid, d3q7)
    t["geodesic"] = time.perf_counter() - t0
    return t


def cpu_baseline(args):
    """The oracle's six stages, timed on this host: samples of the benchmark's generator (128^3,
    then larger while every stage stays under the per-stage cap) and the 128^3 sphere-union
    configuration.  The value is voxels/s on the largest sample that finished under the cap."""
    import oracle
    import porousmediaanalysis_amd as pma

    d3q7 = args.stencil == "d3q7"
    # OpenMP threads as the host allows this process (OMP_NUM_THREADS, else every core)
    threads = args.cpu_threads or int(os.environ.get("OMP_NUM_THREADS", 0) or os.cpu_count() or 1)
    samples = {}
    best = None
    for n in args.cpu_sizes:
        key = f"gaussian_{n}"
        if best is not None and max(samples[best]["stages_s"].values()) * (n / samples[best]["size"]) ** 3 > args.cpu_cap:
            samples[key] = {"size": n, "skipped": f"> cap: a stage would exceed {args.cpu_cap:g} s"}
            continue
        with pma.Context(0) as ctx:
            ctx.generate_gaussian((n, n, n), args.sigma, args.porosity, args.seed + 1)
            solid = ctx.read(pma.VOL_SOLID)
        progress(f"cpu baseline: oracle, all six stages, {n}^3 Gaussian sample, {threads} threads")
        t = oracle_stages(oracle, solid, d3q7, threads)
        samples[key] = {"size": n, "voxels_per_s": n ** 3 / sum(t.values()),
                        "stages_s": {k: round(v, 3) for k, v in t.items()}}
        over = [k for k, v in t.items() if v > args.cpu_cap]
        if over:
            samples[key]["over_cap"] = over
            break
        best = key
    with pma.Context(0) as ctx:
        ctx.generate_spheres((128, 128, 128), 300, 6.0, 12.0, 2024)
        spheres = ctx.read(pma.VOL_SOLID)
    progress("cpu baseline: oracle, all six stages, 128^3 sphere union")
    ts = oracle_stages(oracle, spheres, d3q7, threads)
    samples["sphere_union_128"] = {"size": 128, "voxels_per_s": 128 ** 3 / sum(ts.values()),
                                   "stages_s": {k: round(v, 3) for k, v in ts.items()}}
    ref = samples[best] if best else samples["sphere_union_128"]
    return {"value": ref["voxels_per_s"], "unit": "voxels/s", "cores": threads, "kind": "port",
            "sample": f"{ref['size']}^3 voxels of the benchmark's Gaussian-field generator (sigma={args.sigma:g}, "
                      f"porosity={args.porosity:g}), all six stages: EDT and containing balls with OpenMP, "
                      f"pores, removal, watershed and geodesic serial (as in the reference), "
                      f"{sum(ref['stages_s'].values()):.1f} s; per-stage cap {args.cpu_cap:g} s",
            "samples": samples,
            "host": {"cpu_model": cpu_model(), "os_cpu_count": os.cpu_count(),
                     "omp_num_threads_env": os.environ.get("OMP_NUM_THREADS"), "threads_used": threads}}


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


def measured_traffic(kernel, size):
    """HBM bytes per launch of `kernel` measured with rocprofv3 --pmc (profiles/traffic.json)."""
    path = os.path.join(ROOT, "profiles", "traffic.json")
    if size != 1024 or not os.path.exists(path):
        return None
    with open(path) as f:
        entry = json.load(f).get(kernel)
    return round(entry["traffic_bytes"]) if entry else None


def batched(args, pma):
    """The batched workload: a per-volume loop against pma_run_batch with the tuned configuration."""
    import numpy as np

    rng = np.random.default_rng(args.seed)
    vols, kinds = [], []
    progress(f"generating {args.batch_count} subvolumes")
    with pma.Context(0) as gen:
        for i in range(args.batch_count):
            shape = tuple(int(v) for v in rng.integers(48, 193, 3))
            if i % 2 == 0:
                n = max(1, round(300 * shape[0] * shape[1] * shape[2] / 128 ** 3))
                gen.generate_spheres(shape, n, 6.0, 12.0, args.seed + i)
                kinds.append("spheres")
            else:
                gen.generate_gaussian(shape, float(rng.uniform(2, 6)), float(rng.uniform(0.5, 0.8)), args.seed + i)
                kinds.append("gaussian")
            vols.append(gen.read(pma.VOL_SOLID))
    nvox = sum(int(v.size) for v in vols)
    d3q7 = args.stencil == "d3q7"

    progress("per-volume loop on one context")
    t0 = time.perf_counter()
    with pma.Context(0) as ctx:
        ctx.set_stencil(d3q7)
        for i, v in enumerate(vols):
            ctx.load(pma.VOL_SOLID, v)
            ctx.run_all()
            if (i + 1) % 32 == 0:
                progress(f"per-volume loop: {i + 1}/{len(vols)} after {time.perf_counter() - t0:.1f} s")
    loop_s = time.perf_counter() - t0
    progress(f"per-volume loop: {loop_s:.2f} s")

    sample = vols[:: max(1, len(vols) // 24)][:24]
    t0 = time.perf_counter()
    best, tuned_vps = pma.tune_batch(sample, d3q7)
    tune_s = time.perf_counter() - t0
    progress(f"autotuner: {best.as_dict()} in {tune_s:.2f} s")

    t0 = time.perf_counter()
    pma.run_batch(vols, d3q7, best)
    batch_s = time.perf_counter() - t0
    progress(f"batch: {batch_s:.2f} s")
    t0 = time.perf_counter()
    pma.run_batch(vols, d3q7, pma.BatchConfig(8, 0.0, 0, 0))
    default_s = time.perf_counter() - t0
    progress(f"batch with the default configuration: {default_s:.2f} s")
    return {
        "metric": f"voxels/s, batched: {len(vols)} uneven subvolumes, whole pipeline each (pma_run_batch, tuned)",
        "value": nvox / batch_s,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1000 * batch_s, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "u32+f32",
        "data": "synthetic",
        "config": {"workload": f"{len(vols)} subvolumes, sides U{{48..192}}, half sphere unions, half Gaussian "
                               f"fields ({args.stencil}): EDT, containing balls, pores, empty-pore removal, "
                               f"watershed, geodesic distances with cavities",
                   "subvolumes": len(vols), "voxels": nvox, "seed": args.seed,
                   "tuned": best.as_dict()},
        "per_volume_loop_voxels_per_s": nvox / loop_s,
        "per_volume_loop_s": round(loop_s, 3),
        "batch_s": round(batch_s, 3),
        "batch_default_config_s": round(default_s, 3),
        "speedup_vs_loop": round(loop_s / batch_s, 3),
        "autotuner_s": round(tune_s, 3),
        "autotuner_sample_voxels_per_s": tuned_vps,
        "roofline": None,
        "cpu_baseline": None,
    }


def surrogate(args, pma):
    """Screening with the learned network: train on crops labelled by the exact pipeline, then
    predict the containing radius and distance-ratio fields of fresh crops, rerun the exact stages on
    a 5 % sample, and fall back to them for every sampled crop whose radius MAE exceeds the bound."""
    import numpy as np

    rng = np.random.default_rng(args.seed + 1)
    S = args.crop
    shape = (S, S, S)

    def gen(ctx, i):
        """Crop i: a sphere union (even i) or a Gaussian field with its own sigma and porosity."""
        if i % 2 == 0:
            ctx.generate_spheres(shape, max(1, round(300 * S ** 3 / 128 ** 3)), 6.0, 12.0, args.seed + 7919 * i)
        else:
            r = np.random.default_rng(args.seed + 7919 * i)
            ctx.generate_gaussian(shape, float(r.uniform(2, 6)), float(r.uniform(0.5, 0.8)), args.seed + 7919 * i)

    def exact(ctx):
        ctx.edt()
        ctx.containing_balls()
        ctx.geodesic()

    def fields(ctx):
        solid = ctx.read(pma.VOL_SOLID)
        void = solid == 0
        radius = np.sqrt(ctx.read(pma.VOL_CR2).astype(np.float32))
        dist = ctx.read(pma.VOL_SPD)
        return void, radius, dist

    # training crops are regenerated and relabelled by the exact pipeline at every step (one
    # context; the same seed gives the same crop), so memory does not grow with their number
    net = pma.Surrogate(0, args.seed)
    lab = pma.Context(0)
    t0 = time.perf_counter()
    losses = []
    steps = args.epochs * args.train_crops
    order = np.concatenate([rng.permutation(args.train_crops) for _ in range(args.epochs)])
    for it, i in enumerate(order):
        gen(lab, int(i))
        exact(lab)
        lr = 2e-3 * (0.1 ** (it / max(steps - 1, 1)))
        losses.append(net.train_step(lab, lr=lr))
        if (it + 1) % 128 == 0:
            progress(f"training step {it + 1}/{steps}: loss {np.mean(losses[-128:]):.4f}")
    train_s = time.perf_counter() - t0
    lab.close()

    progress(f"screening {args.test_crops} fresh crops")
    ctx = pma.Context(0)
    sample = set(range(0, args.test_crops, 20))  # 5 %
    t_screen = t_exact = 0.0
    rad_err, rad_max, ratio_err, ratio_max, base_err, fallbacks = [], 0.0, [], 0.0, [], 0
    for i in range(args.test_crops):
        gen(ctx, args.train_crops + i)
        t0 = time.perf_counter()
        ctx.edt()
        radius, ratio = net.predict(ctx)
        t_screen += time.perf_counter() - t0
        if i in sample:
            t0 = time.perf_counter()
            ctx.containing_balls()
            ctx.geodesic()
            t_exact += time.perf_counter() - t0
            void, r_ex, dist = fields(ctx)
            e = np.abs(radius[void] - r_ex[void])
            rad_err.append(float(e.mean()))
            rad_max = max(rad_max, float(e.max()))
            base_err.append(float(np.abs(np.sqrt(ctx.read(pma.VOL_EDT2)[void].astype(np.float32)) - r_ex[void]).mean()))
            cav = ctx.read(pma.VOL_CAVITY)
            rows = [r.split() for r in ctx.text(pma.TEXT_CAVITIES).strip().splitlines()[1:]]
            Z, X, Y = np.meshgrid(np.arange(S), np.arange(S), np.arange(S), indexing="ij")
            for row in rows:
                cid, sx, sy, sz = int(row[0]), int(row[1]), int(row[2]), int(row[3])
                sel = void & (cav == cid) & (dist < 3e38)
                eu = np.sqrt(((X - sx) ** 2 + (Y - sy) ** 2 + (Z - sz) ** 2).astype(np.float32))
                sel &= eu > 0
                if sel.any():
                    q = np.abs(ratio[sel] - dist[sel] / eu[sel])
                    ratio_err.append(float(q.mean()))
                    ratio_max = max(ratio_max, float(q.max()))
            if rad_err[-1] > args.bound:
                fallbacks += 1  # the exact fields just computed replace the estimate
        if (i + 1) % 64 == 0:
            progress(f"screened {i + 1}/{args.test_crops}")
    # the exact stages that produce the same fields, on a share of the crops, for the comparison
    t0 = time.perf_counter()
    n_ref = max(1, args.test_crops // 8)
    for i in range(n_ref):
        gen(ctx, args.train_crops + i)
        exact(ctx)
    exact_per_crop = (time.perf_counter() - t0) / n_ref
    t0 = time.perf_counter()
    for i in range(n_ref):
        gen(ctx, args.train_crops + i)
        ctx.run_all()
    plain_per_crop = (time.perf_counter() - t0) / n_ref
    ctx.close()
    net.close()
    vox = args.test_crops * S ** 3
    screened_s = t_screen + t_exact
    return {
        "metric": "voxels/s, surrogate-screened containing radius and geodesic/Euclidean ratio of fresh crops "
                  "(5 % exact reruns, fallback above the bound)",
        "value": vox / screened_s,
        "unit": "voxels/s",
        "n_gpus": 1,
        "steps": 1,
        "warmup": 0,
        "ms_per_step": round(1000 * screened_s, 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "bf16 (fp32 accumulate)",
        "data": "synthetic",
        "config": {"workload": f"U-Net 16/32/64, {args.train_crops} training crops x {args.epochs} epochs, "
                               f"{args.test_crops} fresh {S}^3 crops (half sphere unions, half Gaussian fields)",
                   "crop": S, "train_crops": args.train_crops, "epochs": args.epochs, "test_crops": args.test_crops,
                   "bound_radius_mae_voxels": args.bound, "seed": args.seed},
        "train_s": round(train_s, 3),
        "final_train_loss": float(np.mean(losses[-64:])),
        "sampled_crops": len(sample),
        "radius_mae": float(np.mean(rad_err)),
        "radius_max_error": rad_max,
        "radius_mae_without_network": float(np.mean(base_err)),
        "ratio_mae": float(np.mean(ratio_err)) if ratio_err else None,
        "ratio_max_error": ratio_max,
        "fallbacks": fallbacks,
        "exact_fields_voxels_per_s": S ** 3 / exact_per_crop,
        "plain_pipeline_voxels_per_s": S ** 3 / plain_per_crop,
        "roofline": None,
        "cpu_baseline": None,
    }


def main():
    args = parse()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (got --gpus {args.gpus}); this build has no multi-GPU path.",
              file=sys.stderr)
        sys.exit(2)
    import porousmediaanalysis_amd as pma

    if args.workload == "batched":
        print(json.dumps(batched(args, pma)))
        return
    if args.workload == "surrogate":
        print(json.dumps(surrogate(args, pma)))
        return

    S = args.size
    shape = (S, S, S)
    nvox = S ** 3
    calls = [("edt", "edt"), ("containing_balls", "containing_balls"), ("pores", "pores"),
             ("remove_empty_pores", "remove_empty_pores"), ("watershed", "watershed"), ("geodesic", "geodesic")]
    with pma.Context(0) as ctx:
        progress(f"generating {S}^3 Gaussian field sigma={args.sigma} porosity={args.porosity}")
        ctx.generate_gaussian(shape, args.sigma, args.porosity, args.seed)
        ctx.set_stencil(args.stencil == "d3q7")

        def step(tag):
            for stage, fn in calls:
                t0 = time.perf_counter()
                getattr(ctx, fn)()
                progress(f"{tag} {stage}: {time.perf_counter() - t0:.3f} s")

        for w in range(args.warmup):
            t0 = time.perf_counter()
            step(f"warmup {w + 1}/{args.warmup}")
            progress(f"warmup step {w + 1}/{args.warmup}: {time.perf_counter() - t0:.3f} s")
        stage_ms = {name: 0.0 for name in STAGES}
        stage_items = {name: 0 for name in STAGES}
        t0 = time.perf_counter()
        for s in range(args.steps):
            ts = time.perf_counter()
            step(f"step {s + 1}/{args.steps}")
            for k, name in enumerate(pma.STAGE_NAMES):
                info = ctx.stage_info(k)
                stage_ms[name] += info["device_ms"]
                stage_items[name] += info["work_items"]
            progress(f"step {s + 1}/{args.steps}: {time.perf_counter() - ts:.3f} s")
        elapsed = time.perf_counter() - t0
        # one more step, outside the timed region, with HIP events around every instrumented kernel
        pma.lib().pma_set_kernel_timing(ctx._h, 1)
        step("kernel-timing step")
        kernels = kernel_times(pma, ctx)
        pma.lib().pma_set_kernel_timing(ctx._h, 0)
        kstage_items = {name: ctx.stage_info(k)["work_items"] for k, name in enumerate(pma.STAGE_NAMES)}
        max_e, n_max = ctx.statistics()
        n_pores = ctx.text(pma.TEXT_PORES).count("\n") - 1
        n_throats = ctx.text(pma.TEXT_THROATS).count("\n") - 1
        n_throats_ws = ctx.text(pma.TEXT_THROATS_WS).count("\n") - 1

    ms_per_step = 1000.0 * elapsed / args.steps
    stages_ms = {k: round(v / args.steps, 3) for k, v in stage_ms.items()}
    stage_roofline = {}
    for k, ms in stages_ms.items():
        gbs = STAGE_BYTES_PER_VOXEL[k] * nvox / (ms * 1e-3) / 1e9 if ms > 0 else None
        stage_roofline[k] = {"bytes_per_voxel": STAGE_BYTES_PER_VOXEL[k],
                             "achieved_GBs": round(gbs, 1) if gbs else None,
                             "frac": round(gbs / HBM_PEAK_GBS, 4) if gbs else None}
    dom = max(kernels.items(), key=lambda kv: kv[1][0]) if kernels else None
    roofline = None
    if dom:
        name, (tot_ms, launches) = dom
        avg_ms = tot_ms / max(launches, 1)
        per_launch = None
        if name in KERNEL_BYTES_PER_VOXEL:
            per_launch = KERNEL_BYTES_PER_VOXEL[name] * nvox / max(launches, 1)
        elif name in STAGE_KERNEL_BYTES_PER_VOXEL:
            per_launch = STAGE_KERNEL_BYTES_PER_VOXEL[name] * nvox / max(launches, 1)
        elif name == "relax_blocks" and launches:
            per_launch = RELAX_BYTES_PER_BLOCK * kstage_items["geodesic"] / launches
        if per_launch is not None:
            achieved = per_launch / (avg_ms * 1e-3) / 1e9
            roofline = {"bound": "hbm", "kernel": name, "achieved": round(achieved, 3), "peak": HBM_PEAK_GBS,
                        "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 6),
                        "traffic": measured_traffic(name, S),
                        "bytes_per_launch": round(per_launch), "avg_launch_ms": round(avg_ms, 4),
                        "launches_per_step": launches}
        else:
            roofline = {"bound": "hbm", "kernel": name, "achieved": None, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                        "frac": None, "traffic": None, "avg_launch_ms": round(avg_ms, 4), "launches": launches,
                        "note": "no whole-volume algorithmic byte count for this kernel"}
    pipeline_bytes = sum(STAGE_BYTES_PER_VOXEL.values()) * nvox
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
                               f"({args.stencil}): EDT, containing balls, pores, empty-pore removal, watershed, "
                               f"geodesic distances with cavities",
                   "stages": list(STAGES),
                   "size": S, "sigma": args.sigma, "porosity": args.porosity,
                   "seed": args.seed, "max_edt2": max_e, "maximum_balls": n_max, "pores": n_pores,
                   "throat_keys": n_throats, "watershed_throat_keys": n_throats_ws},
        "stages_ms": stages_ms,
        "stage_roofline": stage_roofline,
        # EDT: voxels; containing balls: (ball, tile) pairs; pores: DFS replay frames;
        # geodesic: block relaxations
        "stage_work_items": {k: v // args.steps for k, v in stage_items.items()},
        # from the one extra step with events around every instrumented kernel
        "kernels_ms": {k: {"total_ms_per_step": round(v[0], 3), "launches_per_step": v[1]}
                       for k, v in sorted(kernels.items(), key=lambda kv: -kv[1][0])},
        "pipeline_algorithmic_GBs": round(pipeline_bytes / (ms_per_step * 1e-3) / 1e9, 1),
        "roofline": roofline,
        "cpu_baseline": None,
    }
    if not args.no_cpu:
        result["cpu_baseline"] = cpu_baseline(args)
    if not args.no_extras:
        # reduced samples of the batched and screening workloads, so that one run reports both
        # against the plain pipeline (the full configurations: --workload batched / surrogate)
        ex = argparse.Namespace(**vars(args))
        ex.batch_count = 64
        b = batched(ex, pma)
        result["batched"] = {k: b[k] for k in ["value", "unit", "per_volume_loop_voxels_per_s", "speedup_vs_loop",
                                               "batch_s", "per_volume_loop_s", "autotuner_s"]}
        result["batched"]["config"] = b["config"]
        ex.crop, ex.train_crops, ex.epochs, ex.test_crops = 64, 96, 4, 120
        sres = surrogate(ex, pma)
        result["surrogate"] = {k: sres[k] for k in ["value", "unit", "train_s", "radius_mae", "radius_max_error",
                                                   "radius_mae_without_network", "ratio_mae", "ratio_max_error",
                                                   "sampled_crops", "fallbacks", "exact_fields_voxels_per_s",
                                                   "plain_pipeline_voxels_per_s"]}
        result["surrogate"]["config"] = sres["config"]
    print(json.dumps(result))


if __name__ == "__main__":
    main()

"""Training-step throughput of the plain 3DGS method on one MI355X.

Workload: one Trainer.step() (render, L1 + D-SSIM loss, backward, densification
statistics, Adam with activation) at 1.1M Gaussians, 980x545, SH degree 3, FP32,
on the synthetic ring scene: 263 training poses (radius 6 +- 0.5, height N(0, 0.3),
fov 50 deg) whose targets are renders of a seed-7 teacher set of 1.1M Gaussians
with N(0, 0.01) pixel noise. The student is a seed-0 set of the same size and
distribution in the 10-unit box, trained from iteration 15,000 (past the
densification window, SH degree 3), so every timed step is a full optimizer step.

Prints one JSON line. `value` is steps/s over the timed steps. The same number of
steps then runs again with the trainer's stage clock on: HIP events recorded
between the step's launches on the library's stream give each stage's mean time
inside real steps (`stages_us`), and `roofline` is the dominant kernel's
algorithmic bytes over that time. `cpu_baseline` times the float64 CPU
restatement of the same step (oracle/, OpenMP) on a bounded sample.
"""
import argparse
import json
import math
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

N_GAUSSIANS = 1_100_000
WIDTH, HEIGHT = 980, 545
TRAIN_VIEWS, HELD_OUT = 263, 33
HBM_PEAK_GBS = 8000.0

# Algorithmic bytes per unit (useful bytes only): per Gaussian, per (splat, tile)
# pair, per pixel, per tile. Sort (the depth sort and the tile sort together) is
# charged at the reference's 6 x 8-bit pass model over the 48 significant key bits.
STAGE_BYTES = {
    # Runs inside the previous step's backward_adam (below) while that step prepared
    # this one's frame, which every timed step after the first does; the stage then
    # holds only its marker.
    "preprocess": (0, 0, 0, 0),
    "scan": (8, 0, 0, 0),
    "sort": (0, 144, 0, 0),
    "duplicate": (20, 12, 0, 0),
    "ranges": (0, 4, 0, 8),
    "rasterize": (0, 40, 24, 0),
    # The loss's partials kernel also zeroes the splat gradients (36 B/Gaussian) that the
    # rasterizer's backward adds into; the clear stage then holds only its marker.
    "loss": (36, 0, 32, 0),
    "clear_grads": (0, 0, 0, 0),
    "rasterize_backward": (36, 40, 32, 0),
    # preprocess backward + statistics + Adam in one pass: R activated 56, raw 56, SH 180,
    # splat 4 + 36 grads, moments 472, stats 16; W raw 56, SH 180, moments 472,
    # activated 56, stats 16 (the reference's separate passes move 512 + 44 + 1708);
    # plus the next step's preprocess from the same on-chip values: W splat 48, tile
    # count 4, depth key 4 and value 4, (depth, radius) 8 (its reads of the activated
    # Gaussians and SH are the ones above)
    "backward_adam": (1668, 0, 0, 0),
}
# Clocked stages summed into a reported one.
STAGE_PARTS = {"sort": ("depth_sort", "tile_sort", "list_order")}
# Tile lists binned by counts (per_tile_depth 2) run no radix sort and no duplicate
# pass: their stages are charged what they move. bin_count reads each Gaussian's
# packed rectangle (8) and tile count (4); bin_scatter reads those and its depth key
# (4) and writes a (key, splat) pair per (splat, tile) pair (8) and each tile's range
# (8) and order slot (4); list_order reads the pairs (8) and the ranges (8 a tile)
# and writes the splat lists (4 a pair).
BINNED_STAGES = {
    "bin_count": ((12, 0, 0, 0), ("depth_sort", "scan")),
    "bin_scatter": ((16, 8, 0, 12), ("duplicate",)),
    "list_order": ((0, 12, 0, 8), ("tile_sort", "ranges", "list_order")),
}
BINNED_REPLACES = ("scan", "sort", "duplicate", "ranges")
# The reference's per-step byte model (separate preprocess backward, stats and Adam).
STEP_BYTES = (2612, 240, 88)
# Stage name -> kernel name, for the PMC traffic, VALU and occupancy lines (the
# backward's kernel depends on the segment setting, see stage_kernel).
STAGE_KERNEL = {"preprocess": "preprocess_kernel", "rasterize": "rasterize_kernel",
                "rasterize_backward": "rasterize_backward_kernel", "backward_adam": "backward_adam_kernel"}


def stage_kernel(stage, tuning):
    if stage == "rasterize_backward" and tuning.get("segment"):
        return "rasterize_backward_segment_kernel"
    return STAGE_KERNEL.get(stage, stage)
# VALU issue: a wave64 instruction occupies a SIMD for 4 cycles; 1024 SIMDs at 2.4 GHz.
VALU_ISSUE_PER_S = 1024 * 2.4e9 / 4
TRAFFIC_FILE = "pmc_traffic_train_step_1.1M_980x545.txt"
METRIC = "train steps/s @1.1M Gaussians 980x545 FP32; held-out PSNR delta vs plain"


def log(*parts):
    print("[bench]", *parts, file=sys.stderr, flush=True)


def ring_poses(count):
    from diorama_amd import scene

    return scene.ring_cameras(count, WIDTH, HEIGHT, 50.0, 6.0, seed=3)


def split_views(cams):
    """263 training poses and 33 held-out ones taken every 8th."""
    from diorama_amd import scene

    return scene.split_views(cams, HELD_OUT, TRAIN_VIEWS)


def teacher_targets(ctx, cams, n):
    """rgba8 targets: the seed-7 teacher rendered on the GPU plus N(0, 0.01) noise."""
    from diorama_amd import scene

    return scene.teacher_images(ctx, cams, n, seed=7, box=5.0, noise=0.01, noise_seed=7)


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(params, sh, cam, image, sh_degree, repeats=5, budget_s=40.0):
    """The float64 CPU restatement of one optimize() step (render, loss, backward,
    stats, Adam) on every host core this process may use: one warm-up step, then the
    median of `repeats` timed steps, on the full Gaussian set when a step fits the
    budget, else on a random subset with the time scaled to the full count."""
    threads = int(os.environ.get("OMP_NUM_THREADS") or len(os.sched_getaffinity(0)))
    os.environ["OMP_NUM_THREADS"] = str(threads)
    from oracle import oracle

    n = len(params)
    rng = np.random.default_rng(0)
    target = np.zeros((WIDTH * HEIGHT, 4), dtype=np.uint8)
    for c in range(3):
        target[:, c] = (image >> (8 * c)) & 255
    lr = [1 - 0.9 ** 15000, 1 - 0.999 ** 15000, 1.6e-4 * 0.01 ** 0.5, 0.025, 0.005, 0.001, 0.0025, 0.0025 / 20]

    def inputs(sample):
        idx = np.sort(rng.choice(n, size=sample, replace=False)) if sample < n else np.arange(n)
        p = params[idx].astype(np.float64)
        s = sh[idx].astype(np.float64)
        return p, s

    def run(base):
        p, s = base[0].copy(), base[1].copy()
        m = np.zeros((len(p), 122))
        st = np.zeros((len(p), 4))
        g = oracle.activate(p)
        t0 = time.perf_counter()
        oracle.train_step(p, s, m, st, g, sh_degree, cam, [0, 0, 0], target, lr)
        return time.perf_counter() - t0

    sample = n
    base = inputs(n)
    first = run(base)  # warm-up: page-in, OpenMP thread start
    if first * (repeats + 1) > budget_s:
        sample = max(10_000, int(n * budget_s / (first * (repeats + 1))))
        base = inputs(sample)
        run(base)
    times = [run(base) for _ in range(repeats)]
    dt = float(np.median(times))
    scaled = dt * n / sample
    return {"value": 1.0 / scaled, "unit": "steps/s", "cores": oracle.threads(), "kind": "port",
            "cpu_model": cpu_model(), "host_cpus": os.cpu_count(),
            "sample": f"one optimize() step (render, loss, backward, stats, Adam) in float64 on {sample:,} of "
                      f"the {n:,} Gaussians at {WIDTH}x{HEIGHT}, OpenMP on {oracle.threads()} threads: median "
                      f"{dt:.3f} s of {repeats} steps after a warm-up step (all: "
                      + ", ".join(f"{t:.3f}" for t in times) + " s)"
                      + ("" if sample == n else f", time scaled by {n / sample:.2f} to the full count")}


def copy_calibration():
    """Three device-to-device copies of 256 MiB: a kernel whose HBM bytes are known
    (256 MiB read + 256 MiB written each), profiled to scale the PMC byte counters."""
    import torch

    a = torch.ones(64 << 20, dtype=torch.float32, device="cuda")
    b = torch.empty_like(a)
    for _ in range(3):
        b.copy_(a)
    torch.cuda.synchronize()
    print(json.dumps({"copy_calibration_bytes_each_way": a.numel() * 4}), flush=True)
    return 0


SCAN_SIZES = [1 << 20, 2 << 20, 4 << 20, 8 << 20, 16 << 20]
SORT_SIZES = [8 << 20, 30_000_000]


def scan_sort_bench(args):
    """The scan/sort microbenchmark: exclusive scans of 1M-16M geometric tile-touch
    counts and stable sorts of 8M and 30M 64-bit (Zipf tile, depth bits) keys with
    u32 values, through PrefixSum and RadixSort (highBits 16, as the tile binning
    of src/tileRasterizer.ts:279). Each launch is timed with timestamps on the
    library's stream; the input is restored by a device copy before each one."""
    import diorama_amd
    from diorama_amd import scene

    ctx = diorama_amd.Context(0)
    results = {"scan": {}, "sort": {}}
    cpu = {}
    for n in SCAN_SIZES:
        values = scene.tile_touch_counts(n, 5)
        orig = ctx.upload(values)
        data = ctx.buffer(n * 4)
        scan = diorama_amd.PrefixSum(ctx, data, n)
        times = []
        for it in range(args.warmup + args.steps):
            data.copy_from(orig)
            ctx.timestamp(0)
            scan.encode()
            ctx.timestamp(1)
            ms = ctx.elapsed_ms(0, 1)
            if it >= args.warmup:
                times.append(ms)
        out = data.download(np.uint32, n)
        ok = bool(out[-1] == np.uint32(int(values[:-1].astype(np.uint64).sum()) & 0xFFFFFFFF))
        us = float(np.mean(times)) * 1e3
        results["scan"][n] = {"us": round(us, 2), "GB/s": round(8 * n / us / 1e3, 1),
                              "hbm_frac": round(8 * n / us / 1e3 / HBM_PEAK_GBS, 4), "last_ok": ok}
        log(f"scan {n:>9d}: {us:8.1f} us  {8 * n / us / 1e3:7.1f} GB/s")
        if n == SCAN_SIZES[-1] and not args.no_cpu:
            t0 = time.perf_counter()
            np.cumsum(values, dtype=np.uint32)
            cpu["scan_16M_s"] = time.perf_counter() - t0
        scan.destroy()
        data.destroy()
        orig.destroy()
    for n in SORT_SIZES:
        keys, lo, hi = scene.tile_depth_keys(n, 7)
        payload = np.arange(n, dtype=np.uint32)
        k_orig, v_orig = ctx.upload(keys), ctx.upload(payload)
        sort = diorama_amd.RadixSort(ctx, n, 16)
        times = []
        for it in range(args.warmup + args.steps):
            sort.keys.copy_from(k_orig)
            sort.values.copy_from(v_orig)
            ctx.timestamp(0)
            sort.encode()
            ctx.timestamp(1)
            ms = ctx.elapsed_ms(0, 1)
            if it >= args.warmup:
                times.append(ms)
        v = sort.values.download(np.uint32, n)
        sorted_hi, sorted_lo = hi[v], lo[v]
        ok = bool(np.all((sorted_hi[1:] > sorted_hi[:-1]) | ((sorted_hi[1:] == sorted_hi[:-1]) &
                                                          ((sorted_lo[1:] > sorted_lo[:-1]) |
                                                           ((sorted_lo[1:] == sorted_lo[:-1]) & (v[1:] > v[:-1]))))))
        us = float(np.mean(times)) * 1e3
        results["sort"][n] = {"us": round(us, 2), "Mkeys/s": round(n / us, 1),
                              "GB/s_6pass_model": round(144 * n / us / 1e3, 1),
                              "hbm_frac_6pass_model": round(144 * n / us / 1e3 / HBM_PEAK_GBS, 4),
                              "GB/s_one_pass_bound": round(24 * n / us / 1e3, 1), "stable_sorted": ok}
        log(f"sort {n:>9d}: {us:8.1f} us  {n / us:7.1f} Mkeys/s  {144 * n / us / 1e3:7.1f} GB/s (6-pass model)")
        if n == SORT_SIZES[0] and not args.no_cpu:
            t0 = time.perf_counter()
            np.lexsort((lo, hi & 0xFFFF))
            cpu["sort_8M_s"] = time.perf_counter() - t0
        sort.destroy()
        k_orig.destroy()
        v_orig.destroy()
    ctx.close()
    big = results["sort"][SORT_SIZES[-1]]
    line = {
        "metric": "64-bit radix sort Mkeys/s @30M (Zipf tile, depth) keys",
        "value": big["Mkeys/s"],
        "unit": "Mkeys/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "higher_is_better": True,
        "dtype": "u32/u64",
        "data": "synthetic: Geometric(mean 6) counts capped at 256; Zipf(1.1) tiles over 62x35, depth U[0.2, 50]",
        "config": {"workload": "scan 1M-16M u32 and stable sort of 8M/30M 64-bit keys (highBits 16) with u32 values"},
        "results": results,
        "roofline": {"kernel": "radix sort (onesweep passes)", "bound": "hbm", "achieved": big["GB/s_6pass_model"],
                     "peak": HBM_PEAK_GBS, "unit": "GB/s", "frac": big["hbm_frac_6pass_model"], "traffic": None},
        "cpu_baseline": None if args.no_cpu else {
            "value": round(SORT_SIZES[0] / cpu["sort_8M_s"] / 1e6, 2), "unit": "Mkeys/s", "cores": 1, "kind": "port",
            "sample": f"numpy lexsort of the 8M keys: {cpu['sort_8M_s']:.2f} s; numpy cumsum of 16M counts: "
                      f"{cpu['scan_16M_s'] * 1e3:.1f} ms"},
    }
    print(json.dumps(line), flush=True)
    return 0


BATCH_SIZES = [(320, 240), (640, 360), (980, 545), (1280, 720), (1600, 900)]
BATCH_VIEWS = 16


def batch_poses():
    """The 263 training poses of the ring scene, each at a resolution drawn uniformly
    (seeded) from BATCH_SIZES, fov 50 degrees."""
    from diorama_amd import scene

    rng = np.random.default_rng(11)
    base = ring_poses(TRAIN_VIEWS + HELD_OUT)
    train, _ = split_views(base)
    out = []
    for cam in train:
        w, h = BATCH_SIZES[int(rng.integers(len(BATCH_SIZES)))]
        proj = scene.perspective(math.radians(50.0), w / h, 0.01, 100.0)
        out.append(scene.camera(cam["view"], proj, w, h))
    return out


def batch_bench(args):
    """Batched multi-view training: 16 views of uneven resolutions per step through the
    batch trainer (one captured graph per step, gradients summed into one Adam step),
    against the same views one per step (a batch of 1) on the same code path. The
    figure compared is training pixels per second."""
    import diorama_amd
    from diorama_amd import scene

    ctx = diorama_amd.Context(0)
    cams = batch_poses()
    t0 = time.perf_counter()
    images = teacher_targets(ctx, cams, args.gaussians)
    log(f"teacher targets: {len(images)} views of mixed sizes in {time.perf_counter() - t0:.1f} s")
    params, sh = scene.random_raw_gaussians(args.gaussians, 0, box=5.0)
    extent = 1.1 * max(math.sqrt(sum(v * v for v in _eye(c))) for c in cams)
    settings = diorama_amd.default_settings(extent)
    results = {}
    for batch in (BATCH_VIEWS, 1):
        trainer = diorama_amd.Trainer(ctx, params, sh, cams, images, settings, batch=batch)
        trainer.iteration = 15_000
        trainer.sh_degree = 3
        steps = args.steps if batch > 1 else args.steps * BATCH_VIEWS // 4
        for _ in range(args.warmup):
            trainer.step()
        tuned = None if args.no_autotune else trainer.autotune(steps=4)
        if tuned:
            log(f"batch {batch} autotune: " + ", ".join(f"{k['splat_batch']}/{k['digit_bits']}: {ms:.3f} ms"
                                                        for k, ms in tuned))
        ctx.synchronize()
        pixels = 0
        t0 = time.perf_counter()
        for _ in range(steps):
            trainer.step()
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        # pixels per step: the mean view size times the views per step
        mean_px = float(np.mean([c["viewport"][0] * c["viewport"][1] for c in cams]))
        pixels = mean_px * batch * steps
        graph_ms = trainer.time_stage("batch_step", 5)
        results[batch] = {"steps_per_s": steps / elapsed, "views_per_s": steps * batch / elapsed,
                          "Mpx_per_s": pixels / elapsed / 1e6, "ms_per_step": elapsed * 1e3 / steps,
                          "graph_ms": graph_ms, "splat_batch": trainer.tuning["splat_batch"],
                          "digit_bits": trainer.tuning["digit_bits"]}
        log(f"batch {batch:2d}: {elapsed * 1e3 / steps:8.2f} ms/step, {steps * batch / elapsed:8.1f} views/s, "
            f"{pixels / elapsed / 1e6:8.1f} Mpx/s (graph replay {graph_ms:.2f} ms)")
        trainer.destroy()
    ctx.close()
    b, one = results[BATCH_VIEWS], results[1]
    line = {
        "metric": "batched multi-view training views/s @1.1M Gaussians, 16 views of uneven sizes per step",
        "value": round(b["views_per_s"], 2),
        "unit": "views/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(b["ms_per_step"], 3),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seed-7 teacher renders + N(0,0.01) noise on 263 ring poses at sizes drawn from "
                "320x240..1600x900; seed-0 student",
        "config": {"workload": "batch trainer, 16 views per step, one graph replay + one Adam step per batch",
                   "gaussians": args.gaussians, "sizes": BATCH_SIZES, "views_per_step": BATCH_VIEWS},
        "batch16": {k: round(v, 3) for k, v in b.items()},
        "batch1": {k: round(v, 3) for k, v in one.items()},
        "speedup_px_per_s_vs_batch1": round(b["Mpx_per_s"] / one["Mpx_per_s"], 3),
    }
    print(json.dumps(line), flush=True)
    return 0


def scaled_settings(settings, factor):
    """The schedule of `settings` for batch steps of `factor` views: every step count
    divided by `factor`, the learning rates multiplied by factor^(3/4) (16 times fewer
    Adam steps, each on the summed gradient of 16 views). On the shortened protocol
    of the default line (11,500 views, 105k budget, one box) the batched + learned
    form's held-out PSNR against the plain rule's was +0.08 dB with every rate x4
    (sqrt 16), +0.36 dB x6, +0.44 dB x8 (this rule), +0.26 dB x11.3; runs of one
    setting differ by up to about 0.2 dB."""
    out = dict(settings)
    for key in ("position_lr_steps", "sh_degree_interval", "densify_from", "densify_until", "densify_interval",
                "opacity_reset_interval"):
        out[key] = max(1, int(round(settings[key] / factor)))
    scale = factor ** 0.75
    out["position_lr"] = [v * scale for v in settings["position_lr"]]
    for key in ("opacity_lr", "scale_lr", "rotation_lr", "color_lr", "sh_lr"):
        out[key] = settings[key] * scale
    return out


def learned_bench(args):
    """Learned densification on the ring scene, scored on the 33 held-out views.

    A: the plain rule from 100k Gaussians for --train-steps steps under the config's
       1.1M budget, logging samples (inputs and the rule's action, kept when the loss
       over the Gaussian's footprint fell over the next 100 steps) from step 500 to
       --log-until; the MLP then trains on them (2,000 Adam steps of 4,096 samples);
    B: the same run densified by the MLP (the rule where it is unsure), 1.1M budget;
    C, D: the plain rule and the MLP under --binding-budget, which this synthetic
       scene's growth does reach (the plain rule alone stops near 116k);
    E: B batched, 16 views per step with the schedule divided by 16."""
    import diorama_amd
    from diorama_amd import scene

    ctx = diorama_amd.Context(0)
    cams = ring_poses(TRAIN_VIEWS + HELD_OUT)
    train, held = split_views(cams)
    t0 = time.perf_counter()
    images = teacher_targets(ctx, train + held, N_GAUSSIANS)
    train_images, held_images = images[:len(train)], images[len(train):]
    log(f"teacher targets: {len(images)} views in {time.perf_counter() - t0:.1f} s")
    params, sh = scene.random_raw_gaussians(args.initial, 0, box=5.0)
    extent = 1.1 * max(math.sqrt(sum(v * v for v in _eye(c))) for c in train)
    settings = diorama_amd.default_settings(extent)
    budget = N_GAUSSIANS
    steps = args.train_steps
    per_densify = max(1000, 5_000_000 // max(1, (min(args.log_until, settings["densify_until"]) - 500) // 100))

    def run(name, trainer, n_steps, on_step=None):
        t0 = time.perf_counter()
        for k in range(1, n_steps + 1):
            trainer.step()
            if on_step:
                on_step(k)
            if k % 5000 == 0:
                log(f"{name}: step {k}/{n_steps}, {trainer.count} Gaussians, {time.perf_counter() - t0:.1f} s")
        ctx.synchronize()
        elapsed = time.perf_counter() - t0
        psnr = trainer.evaluate(held, held_images, with_ssim=False)["psnr"]
        log(f"{name}: {n_steps} steps in {elapsed:.1f} s, {trainer.count} Gaussians, held-out PSNR {psnr:.3f} dB")
        return {"steps": n_steps, "seconds": round(elapsed, 2), "steps_per_s": round(n_steps / elapsed, 2),
                "views_per_s": round(n_steps * max(1, trainer.batch_size) / elapsed, 2), "gaussians": trainer.count,
                "largest_count": max([trainer.count] + [d["after"] for d in trainer.densifications]),
                "held_out_psnr": round(psnr, 4), "densifications": len(trainer.densifications)}

    out = {}
    a = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
    a.set_densify_mode("log", budget=budget, samples_per_densify=per_densify)

    def stop_logging(k):
        if k == args.log_until:
            a.set_densify_mode("plain", budget=budget)

    out["plain"] = run("plain", a, steps, stop_logging)
    feats, labels, count = a.samples()
    classes = np.bincount(labels.download(np.uint32, count), minlength=4).tolist() if count else [0, 0, 0, 0]
    net = diorama_amd.DensifyMLP(ctx, seed=1)
    if count:  # standardise the inputs over the logged samples
        x = feats.download(np.float32, count * 12).reshape(count, 12).astype(np.float64)
        good = x[np.isfinite(x).all(axis=1)]
        net.set_input_norm(good.mean(axis=0), 1.0 / np.maximum(good.std(axis=0), 1e-6))
    # square-root inverse class frequency: the rule keeps almost every Gaussian
    weights = [min(30.0, math.sqrt(max(count, 1) / (4.0 * max(c, 1)))) for c in classes]
    net.set_class_weights(weights)
    t0 = time.perf_counter()
    mlp_loss = net.train(feats, labels, count, steps=2000, batch=4096, lr=2e-3, seed=7) if count else float("nan")
    mlp_s = time.perf_counter() - t0
    log(f"mlp: {count} samples {classes}, 2000 Adam steps in {mlp_s:.2f} s, loss {mlp_loss:.4f}")
    confusion = None
    if count:
        logits = ctx.buffer(count * 16)
        net.forward(feats, count, logits)
        pred = np.argmax(logits.download(np.float32, count * 4).reshape(count, 4), axis=1)
        lab = labels.download(np.uint32, count)
        confusion = np.zeros((4, 4), dtype=np.int64)
        np.add.at(confusion, (lab, pred), 1)
        log(f"mlp confusion (rows: label, cols: predicted): {confusion.tolist()}")
        logits.destroy()
    a.destroy()
    b = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
    b.set_densify_mode("learned", net, budget=budget)
    out["learned"] = run("learned", b, steps)
    b.destroy()
    if args.binding_budget:
        c = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
        c.set_densify_mode("plain", budget=args.binding_budget)
        out["plain_binding"] = run("plain, binding budget", c, steps)
        c.destroy()
        d = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
        d.set_densify_mode("learned", net, budget=args.binding_budget)
        out["learned_binding"] = run("learned, binding budget", d, steps)
        d.destroy()
    if not args.no_batch:
        e = diorama_amd.Trainer(ctx, params, sh, train, train_images, scaled_settings(settings, BATCH_VIEWS),
                                batch=BATCH_VIEWS)
        e.set_densify_mode("learned", net, budget=budget)
        out["batched_learned"] = run("batched+learned", e, steps // BATCH_VIEWS)
        e.destroy()
    net.destroy()
    ctx.close()
    delta = out["learned"]["held_out_psnr"] - out["plain"]["held_out_psnr"]
    line = {
        "metric": "held-out PSNR delta vs plain (learned densification, 1.1M budget)",
        "value": round(delta, 4),
        "unit": "dB",
        "n_gpus": 1,
        "steps": steps,
        "warmup": 0,
        "higher_is_better": True,
        "dtype": "fp32 (MLP: bf16 weights, fp32 accumulation)",
        "data": "synthetic: seed-7 1.1M teacher renders + N(0,0.01) noise, 263 train / 33 held-out ring poses "
                "at 980x545; seed-0 student",
        "config": {"workload": "plain rule vs learned densification vs batched+learned", "initial": args.initial,
                   "budget": budget, "binding_budget": args.binding_budget, "train_steps": steps,
                   "log_until": args.log_until, "labels": "rule action kept if the footprint loss fell over 100 steps",
                   "samples": count, "sample_classes": classes, "class_weights": [round(w, 3) for w in weights],
                   "mlp_loss": round(mlp_loss, 4),
                   "mlp_confusion": confusion.tolist() if confusion is not None else None,
                   "mlp_train_s": round(mlp_s, 3)},
        "runs": out,
        "budget_binds": out["plain"]["largest_count"] >= budget,
    }
    if "plain_binding" in out:
        line["binding_budget_psnr_delta"] = round(out["learned_binding"]["held_out_psnr"] -
                                                  out["plain_binding"]["held_out_psnr"], 4)
    if "batched_learned" in out:
        line["batched_learned_psnr_delta"] = round(out["batched_learned"]["held_out_psnr"] -
                                                   out["plain"]["held_out_psnr"], 4)
        line["batched_learned_views_per_s_ratio"] = round(out["batched_learned"]["views_per_s"] /
                                                          out["plain"]["views_per_s"], 3)
    print(json.dumps(line), flush=True)
    return 0


def short_psnr_delta(ctx, train, train_images, held, held_images, initial=100_000, steps=11_500, log_until=7000,
                     budget=105_000):
    """The second half of the metric on a shortened config-4 protocol, for the
    default bench line: the plain rule (logging samples) and the MLP trained on them
    each grow `initial` Gaussians for `steps` steps under a `budget` that binds on
    this scene, then both are scored on the held-out views. The steps end 2,500 past
    the opacity reset of step 9000 and before the next one. The MLP's decisions also
    drive a batch trainer (16 views a step, the schedule divided by 16, the same
    number of views), whose held-out PSNR and views/s are set against the plain
    rule's: views/s of the plain rule over its steps after logging stops, of the
    batched + learned form over its whole run. The full 30k-step comparison is
    --workload learned."""
    import diorama_amd
    from diorama_amd import scene

    params, sh = scene.random_raw_gaussians(initial, 0, box=5.0)
    extent = 1.1 * max(math.sqrt(sum(v * v for v in _eye(c))) for c in train)
    settings = diorama_amd.default_settings(extent)
    t0 = time.perf_counter()
    a = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
    a.set_densify_mode("log", budget=budget, samples_per_densify=200_000)
    plain_t0 = None
    for k in range(1, steps + 1):
        a.step()
        if k == log_until:
            a.set_densify_mode("plain", budget=budget)
            ctx.synchronize()
            plain_t0 = time.perf_counter()
    ctx.synchronize()
    plain_views_s = (steps - log_until) / (time.perf_counter() - plain_t0)
    plain = a.evaluate(held, held_images, with_ssim=False)["psnr"]
    plain_count = a.count
    feats, labels, count = a.samples()
    classes = np.bincount(labels.download(np.uint32, count), minlength=4) if count else np.zeros(4)
    net = diorama_amd.DensifyMLP(ctx, seed=1)
    if count:
        x = feats.download(np.float32, count * 12).reshape(count, 12).astype(np.float64)
        good = x[np.isfinite(x).all(axis=1)]
        net.set_input_norm(good.mean(axis=0), 1.0 / np.maximum(good.std(axis=0), 1e-6))
        net.set_class_weights([min(30.0, math.sqrt(count / (4.0 * max(c, 1)))) for c in classes])
        net.train(feats, labels, count, steps=2000, batch=4096, lr=2e-3, seed=7)
    a.destroy()
    b = diorama_amd.Trainer(ctx, params, sh, train, train_images, settings)
    b.set_densify_mode("learned", net, budget=budget)
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(steps):
        b.step()
    ctx.synchronize()
    learned_views_s = steps / (time.perf_counter() - t1)
    learned = b.evaluate(held, held_images, with_ssim=False)["psnr"]
    learned_count = b.count
    b.destroy()
    e = diorama_amd.Trainer(ctx, params, sh, train, train_images, scaled_settings(settings, BATCH_VIEWS),
                            batch=BATCH_VIEWS)
    e.set_densify_mode("learned", net, budget=budget)
    batch_steps = steps // BATCH_VIEWS
    ctx.synchronize()
    t1 = time.perf_counter()
    for _ in range(batch_steps):
        e.step()
    ctx.synchronize()
    batched_views_s = batch_steps * BATCH_VIEWS / (time.perf_counter() - t1)
    batched = e.evaluate(held, held_images, with_ssim=False)["psnr"]
    batched_count = e.count
    e.destroy()
    net.destroy()
    log(f"held-out PSNR at a {budget:,} budget after {steps} views: plain {plain:.3f} dB ({plain_count:,}, "
        f"{plain_views_s:.0f} views/s), learned {learned:.3f} dB ({learned_count:,}, {learned_views_s:.0f} views/s), "
        f"batched+learned {batched:.3f} dB ({batched_count:,}, {batched_views_s:.0f} views/s), "
        f"{time.perf_counter() - t0:.1f} s")
    return {"value": round(learned - plain, 4), "unit": "dB", "plain_db": round(plain, 4),
            "learned_db": round(learned, 4), "batched_learned_db": round(batched, 4),
            "batched_learned_delta_db": round(batched - plain, 4),
            "views_per_s": {"plain": round(plain_views_s, 1), "learned": round(learned_views_s, 1),
                            "batched_learned": round(batched_views_s, 1)},
            "batched_learned_views_per_s_ratio": round(batched_views_s / plain_views_s, 3),
            "gaussians": [plain_count, learned_count, batched_count], "budget": budget,
            "initial": initial, "steps": steps, "views_per_batch_step": BATCH_VIEWS, "samples": count,
            "held_out_views": len(held),
            "protocol": "shortened config 4 (the full 30k-step run: --workload learned)"}


def growth_bench(args):
    """Why the plain rule grows the ring scene only a little: per densification window
    (100 steps from step 500), the distribution of the mean view-space gradient
    (NDC units, the statistic src/trainer.ts:258-337 thresholds at 2e-4) over the
    Gaussians seen, and how many pass the threshold."""
    import diorama_amd
    from diorama_amd import scene

    ctx = diorama_amd.Context(0)
    cams = ring_poses(TRAIN_VIEWS + HELD_OUT)
    train, _ = split_views(cams)
    images = teacher_targets(ctx, train, N_GAUSSIANS)
    params, sh = scene.random_raw_gaussians(args.initial, 0, box=5.0)
    extent = 1.1 * max(math.sqrt(sum(v * v for v in _eye(c))) for c in train)
    settings = diorama_amd.default_settings(extent)
    threshold = settings["densify_grad_threshold"]
    t = diorama_amd.Trainer(ctx, params, sh, train, images, settings)
    windows = []
    for step in range(1, args.train_steps + 1):
        if step > settings["densify_from"] and step % settings["densify_interval"] == 0 and step < settings["densify_until"]:
            st = t.buffers["stats"].download(np.float32).reshape(-1, 4)[:t.count]
            seen = st[:, 1] > 0
            g = st[seen, 0] / st[seen, 1]
            windows.append({"step": step, "gaussians": t.count, "seen": int(seen.sum()),
                            "views_mean": round(float(st[seen, 1].mean()), 1),
                            "grad_percentiles_50_90_99_99.9": [float(f"{v:.3g}") for v in
                                                               np.percentile(g, [50, 90, 99, 99.9])],
                            "at_or_above_threshold": int((g >= threshold).sum())})
            log(json.dumps(windows[-1]))
        t.step()
    t.destroy()
    ctx.close()
    print(json.dumps({"metric": "Gaussians at or above the densify threshold per window", "unit": "Gaussians",
                      "value": int(np.median([w["at_or_above_threshold"] for w in windows[1:]])) if len(windows) > 1
                      else None, "threshold": threshold, "config": {"workload": "plain rule growth statistics",
                                                                    "initial": args.initial,
                                                                    "steps": args.train_steps},
                      "windows": windows}), flush=True)
    return 0


def measured_valu(kernel):
    """VALU instructions per launch of `kernel` from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "valu" and parts[1] == kernel:
                return int(parts[2])
    return None


def measured_occupancy(kernel):
    """Achieved waves per SIMD of `kernel` from the committed PMC summary, or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 3 and parts[0] == "occupancy" and parts[1] == kernel:
                return float(parts[2])
    return None


def measured_traffic(kernel):
    """HBM bytes per launch of `kernel` from the committed PMC summary (FETCH_SIZE and
    WRITE_SIZE already scaled by the copy calibration), or None."""
    path = os.path.join(ROOT, "profiles", TRAFFIC_FILE)
    if not os.path.exists(path):
        return None
    with open(path) as f:
        for line in f:
            parts = line.split()
            if len(parts) >= 4 and parts[0] == "traffic" and parts[1] == kernel:
                return int(parts[2]) + int(parts[3])
    return None


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=50)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--gaussians", type=int, default=N_GAUSSIANS)
    ap.add_argument("--no-cpu", action="store_true", help="skip the CPU baseline")
    ap.add_argument("--no-learned", action="store_true",
                    help="train: skip the shortened learned-vs-plain held-out PSNR comparison")
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch settings")
    ap.add_argument("--tuning", default=None,
                    help="train: tile_size,splat_batch,digit_bits[,persistent[,segment[,per_tile_depth]]] to use "
                         "instead of autotuning (e.g. 32,128,8 or 16,64,8,1 or 16,64,8,0,256 or 16,64,8,0,512,1)")
    ap.add_argument("--workload", choices=["train", "scan-sort", "batch", "learned", "growth"], default="train",
                    help="train: the training step (default); scan-sort: the scan/sort microbenchmark; "
                         "batch: the multi-view batch trainer; learned: learned vs plain densification; "
                         "growth: the plain rule's gradient statistics per densification window")
    ap.add_argument("--train-steps", type=int, default=30000, help="learned: training steps per run")
    ap.add_argument("--log-until", type=int, default=7000, help="learned: last step that logs samples")
    ap.add_argument("--initial", type=int, default=100_000, help="learned: initial Gaussians")
    ap.add_argument("--no-batch", action="store_true", help="learned: skip the batched run")
    ap.add_argument("--binding-budget", type=int, default=105_000,
                    help="learned: the second, binding Gaussian budget (0: skip those runs)")
    ap.add_argument("--copy-calibration", action="store_true",
                    help="only run three plain 256 MiB device copies (the FETCH_SIZE/WRITE_SIZE calibration)")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only; --gpus {args.gpus} is not supported in this build "
              "(multi-GPU sharding is a design, not an implementation)", file=sys.stderr)
        return 2
    if args.copy_calibration:
        return copy_calibration()
    if args.workload == "scan-sort":
        return scan_sort_bench(args)
    if args.workload == "batch":
        return batch_bench(args)
    if args.workload == "learned":
        return learned_bench(args)
    if args.workload == "growth":
        return growth_bench(args)

    import diorama_amd
    from diorama_amd import scene

    ctx = diorama_amd.Context(0)
    cams = ring_poses(TRAIN_VIEWS + HELD_OUT)
    train, held = split_views(cams)
    t0 = time.perf_counter()
    images = teacher_targets(ctx, train + held, args.gaussians)
    images, held_images = images[:len(train)], images[len(train):]
    log(f"teacher targets: {len(images) + len(held_images)} views in {time.perf_counter() - t0:.1f} s")
    params, sh = scene.random_raw_gaussians(args.gaussians, 0, box=5.0)
    extent = 1.1 * max(math.sqrt(sum(v * v for v in _eye(c))) for c in train)
    settings = diorama_amd.default_settings(extent)
    trainer = diorama_amd.Trainer(ctx, params, sh, train, images, settings)
    trainer.iteration = 15_000
    trainer.sh_degree = 3
    for _ in range(args.warmup):
        trainer.step()
    tuned = None
    if args.tuning:
        v = [int(x) for x in args.tuning.split(",")]
        trainer.tuning = {"tile_size": v[0], "splat_batch": v[1], "digit_bits": v[2],
                          "persistent": v[3] if len(v) > 3 else 0, "segment": v[4] if len(v) > 4 else 0,
                          "per_tile_depth": v[5] if len(v) > 5 else 0}
    elif not args.no_autotune:
        tuned = trainer.autotune(steps=3)
        log("autotune: " + ", ".join(f"tile {k['tile_size']}{'p' if k['persistent'] else ''} batch "
                                     f"{k['splat_batch']} digits {k['digit_bits']} segment {k['segment']} "
                                     f"per-tile depth {k['per_tile_depth']}: {ms:.3f} ms" for k, ms in tuned))
    ctx.synchronize()
    log(f"warmup {args.warmup} steps done; tuning {trainer.tuning}")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trainer.step()
    ctx.synchronize()
    elapsed = time.perf_counter() - t0
    steps_per_s = args.steps / elapsed
    log(f"timed {args.steps} steps: {elapsed * 1000 / args.steps:.3f} ms/step")

    # The same steps again with the stage clock on, for the per-stage times inside
    # real steps; P is the mean of those steps' own pair counts.
    trainer.stage_clock(True)
    t0 = time.perf_counter()
    for _ in range(args.steps):
        trainer.step()
    ctx.synchronize()
    clocked_elapsed = time.perf_counter() - t0
    clock = trainer.stage_totals()
    trainer.stage_clock(False)
    totals, clocked = clock["ms"], clock["steps"]
    if clocked != args.steps:
        raise RuntimeError(f"the stage clock counted {clocked} of {args.steps} steps")
    log(f"clocked {clocked} steps: {clocked_elapsed * 1000 / clocked:.3f} ms/step")
    pairs = int(round(clock["pairs"] / clocked))
    n = trainer.count
    x = WIDTH * HEIGHT
    n_tiles = math.ceil(WIDTH / 16) * math.ceil(HEIGHT / 16)
    stages = {}
    model = {name: (b, STAGE_PARTS.get(name, (name,))) for name, b in STAGE_BYTES.items()}
    if trainer.tuning.get("per_tile_depth") == 2:
        model = {name: v for name, v in model.items() if name not in BINNED_REPLACES}
        model.update(BINNED_STAGES)
        order = list(STAGE_BYTES)
        model = dict(sorted(model.items(), key=lambda kv: order.index("ranges") + 0.5 if kv[0] in BINNED_STAGES
                            else order.index(kv[0])))
    for name, ((bg, bp, bx, bt), parts) in model.items():
        ms = sum(totals[p] for p in parts) / clocked
        nbytes = bg * n + bp * pairs + bx * x + bt * n_tiles
        stages[name] = {"ms": ms, "bytes": nbytes, "GB/s": nbytes / (ms * 1e-3) / 1e9}
        log(f"{name:20s} {ms * 1000:9.1f} us  {stages[name]['GB/s']:8.1f} GB/s algorithmic")
    step_gpu_ms = sum(totals.values()) / clocked
    dominant = max(STAGE_KERNEL, key=lambda k: stages[k]["ms"])
    d = stages[dominant]
    tuning_now = trainer.tuning
    roofline = {"kernel": dominant, "bound": "hbm", "achieved": round(d["GB/s"], 1), "peak": HBM_PEAK_GBS,
                "unit": "GB/s", "frac": round(d["GB/s"] / HBM_PEAK_GBS, 4),
                "traffic": measured_traffic(stage_kernel(dominant, tuning_now)),
                "algorithmic_bytes": d["bytes"]}
    valu = measured_valu(stage_kernel(dominant, tuning_now))
    if valu:
        # The rasterizer passes are bound by VALU issue, not HBM: the share of the
        # issue limit their PMC-counted instructions take at the measured time.
        roofline["valu_issue_frac"] = round(valu / (d["ms"] * 1e-3) / VALU_ISSUE_PER_S, 4)
    step_bytes = STEP_BYTES[0] * n + STEP_BYTES[1] * pairs + STEP_BYTES[2] * x

    cpu = None
    if not args.no_cpu:
        t0 = time.perf_counter()
        cpu = cpu_baseline(params, sh, train[0], images[0], 3)
        log(f"cpu baseline: {cpu['value']:.4f} steps/s on {cpu['cores']} threads of {cpu['cpu_model']} "
            f"({time.perf_counter() - t0:.1f} s)")
    trainer_tuning = trainer.tuning
    trainer.destroy()
    delta = None if args.no_learned else short_psnr_delta(ctx, train, images, held, held_images)
    autotune_table = None if tuned is None else [[k["tile_size"], k["splat_batch"], k["digit_bits"], k["persistent"],
                                                  k["segment"], k["per_tile_depth"], round(ms, 4)]
                                                 for k, ms in tuned]
    ctx.close()
    line = {
        "metric": METRIC,
        "value": round(steps_per_s, 2),
        "unit": "steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(elapsed * 1000 / args.steps, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "fp32",
        "data": "synthetic: seed-7 teacher renders + N(0,0.01) noise on 263 ring poses; seed-0 student",
        "config": {"workload": "plain 3DGS training step (Trainer.step: render, L1+D-SSIM, backward, stats, Adam)",
                   "gaussians": n, "width": WIDTH, "height": HEIGHT, "sh_degree": 3, "train_views": len(train),
                   "tuning": trainer_tuning, "autotune_ms": autotune_table,
                   "pairs": pairs, "views_per_s": round(steps_per_s, 2),
                   "step_bytes_model": step_bytes,
                   "step_GBps_model": round(step_bytes / (elapsed / args.steps) / 1e9, 1),
                   "parallelism": "single GPU"},
        "stages_us": {k: round(v["ms"] * 1000, 1) for k, v in stages.items()},
        "stage_clock": {"steps": clocked, "ms_per_step": round(clocked_elapsed * 1000 / clocked, 4),
                        "gpu_ms_per_step": round(step_gpu_ms, 4),
                        "parts_us": {k: round(totals[k] / clocked * 1000, 1) for k in totals}},
        "stages_hbm_frac": {k: round(v["GB/s"] / HBM_PEAK_GBS, 4) for k, v in stages.items()},
        "roofline": roofline,
        # Achieved occupancy of the rasterizer passes, from the committed PMC summary
        # (SQ_WAVE_CYCLES over GRBM_GUI_ACTIVE per CU; 8 waves per SIMD at most).
        "rasterizer_waves_per_simd": {k: measured_occupancy(stage_kernel(k, trainer_tuning))
                                      for k in ("rasterize", "rasterize_backward")},
        "cpu_baseline": cpu,
        "held_out_psnr_delta": delta,
    }
    if cpu:
        line["config"]["speedup_vs_cpu"] = round(steps_per_s / cpu["value"], 1)
    print(json.dumps(line), flush=True)
    return 0


def _eye(cam):
    """Camera centre of a view matrix [R | t]: -Rᵀ t."""
    v = cam["view"]
    return [-(v[c * 4] * v[12] + v[c * 4 + 1] * v[13] + v[c * 4 + 2] * v[14]) for c in range(3)]


if __name__ == "__main__":
    sys.exit(main())

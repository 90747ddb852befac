"""Benchmark of the PD-Flow scene-flow path on one MI355X.

Default workload (--workload learned): 256 VGA (480x640) RGB-D pairs per step, batched with the
learned warm start -- both pyramids of every pair, one network pass over their 60x80 levels, the
solves from there with the calibrated schedule (every level's iterations divided by 2.5, the finest
level's budget chosen on training pairs, pdflow/calibrate.py), then the acceptance rule computed from
the learned solve alone (finest-level energy per pixel, warped residual against the zero-motion
residual, bounds calibrated on the plain solver's training results), up to three stages of 16 more
finest-level iterations for the pairs outside the bounds, and the plain re-solve of those still
outside -- all inside the timed region, with two batches in flight so that one batch's rule,
continuation and fallback overlap the next batch's solve (--pipeline).  The line also carries the
rate of the rule as the benchmark configuration words it (the plain solve of every pair inside the
step, its energy and residual as the references), the plain single-pair and plain batched rates of
the same pairs, end-point errors against the plain solve and the rendered ground truth, and how
often the calibrated rule agrees with the plain-referenced one.

--workload plain: the drop-in path on one VGA pair per step (reference schedule 25..100).
Inputs are seeded synthetic pairs already resident in HBM when the timed region starts.  One
JSON line is printed, including the roofline of the primal-dual iteration launches (148 bytes per
pixel per launch over their HIP-event-timed duration) and a CPU baseline.
"""
import argparse
import json
import os
import sys
import time

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, os.path.join(ROOT, "pd-flow_amd"))

import numpy as np  # noqa: E402

METRIC = "VGA scene-flow pairs/s: batched+learned-init vs plain, same EPE; HBM % peak"
HBM_PEAK_GBS = 8000.0
# One primal-dual iteration launch (any number of iterations k) must read the 11 per-level
# constant planes and the 13 state planes and write the 13 state planes once per pixel: 37 FP32
# words = 148 bytes per pixel per launch, the algorithmic HBM bytes of the roofline.  Dividing the
# same bytes by pixel-iterations instead gives an "effective" rate that temporal blocking lifts
# above the HBM peak; it is reported under its own name.
BYTES_PER_PIXEL_LAUNCH = 148
PROFILES = os.path.join(ROOT, "profiles")
# the sources of the iteration kernels (the kernels and the per-pixel arithmetic they inline)
ITERATE_SRCS = [os.path.join(ROOT, "pd-flow_amd", "csrc", f) for f in ("pdflow_iterate.hip", "pdflow_pixel.hpp")]
RULE_JSON = os.path.join(ROOT, "pd-flow_amd", "pdflow", "weights", "warmstart_qvga_rule.json")


def source_sha16(paths):
    import hashlib
    h = hashlib.sha256()
    for path in paths:
        with open(path, "rb") as f:
            h.update(f.read())
    return h.hexdigest()[:16]


def measured_traffic(workload_key):
    """HBM bytes per launch of the iteration kernel for this workload, from the committed
    rocprofv3 --pmc summary (profiles/*_hbm_traffic.txt): its 'traffic:' line must name this
    workload and the iteration kernels' current source hash, else the reading is stale and None
    is returned with the reason"""
    sha = source_sha16(ITERATE_SRCS)
    if not os.path.isdir(PROFILES):
        return None, "no profiles/ directory"
    for name in sorted(os.listdir(PROFILES)):
        if not name.endswith("_hbm_traffic.txt"):
            continue
        with open(os.path.join(PROFILES, name)) as f:
            for line in f:
                if not line.startswith("traffic:"):
                    continue
                kv = dict(t.split("=", 1) for t in line[len("traffic:"):].split())
                if kv.get("workload") == workload_key:
                    if kv.get("iterate_source_sha16") != sha:
                        return None, f"{name} was measured on another build of the iteration kernels"
                    return float(kv["bytes_per_launch"]), f"profiles/{name}"
    return None, "no traffic profile of this workload"


def measured_l2_hit_rate(traffic_source):
    """the L2 hit rate of the same launches, from the 'L2 hit rate over the' line of the traffic
    profile that gave the traffic (None without one)"""
    if not traffic_source or not traffic_source.startswith("profiles/"):
        return None
    with open(os.path.join(ROOT, traffic_source)) as f:
        for line in f:
            if line.startswith("L2 hit rate over the"):
                return float(line.rsplit(":", 1)[1])
    return None


def roofline(kernel, timing, workload_key=None):
    """roofline of the dominant kernel from HIP-event timing (launches, ms, pixel-iterations,
    pixels covered) on the stream the launches run on"""
    launches, kms, pix_iters, pixels = timing
    sec = kms * 1e-3
    achieved = BYTES_PER_PIXEL_LAUNCH * pixels / sec / 1e9
    traffic, source = measured_traffic(workload_key) if workload_key else (None, "not profiled")
    return {"bound": "hbm", "kernel": kernel, "achieved": round(achieved, 1), "peak": HBM_PEAK_GBS,
            "unit": "GB/s", "frac": round(achieved / HBM_PEAK_GBS, 4), "traffic": traffic,
            "traffic_source": source,
            "l2_hit_rate": measured_l2_hit_rate(source),
            "algorithmic_bytes_per_launch": round(BYTES_PER_PIXEL_LAUNCH * pixels / launches),
            "avg_launch_us": round(kms / launches * 1e3, 3), "launches": launches,
            "bytes_per_pixel_per_launch": BYTES_PER_PIXEL_LAUNCH,
            "effective_gbs_over_pixel_iterations": round(BYTES_PER_PIXEL_LAUNCH * pix_iters / sec / 1e9, 1)}


def log(msg):
    print(msg, file=sys.stderr, flush=True)


def cpu_baseline(rows, cam_mode, seed, budget_s):
    """The CPU oracle (a C port of the reference per-pixel arithmetic and driver) timed on this
    host on a bounded sample of the same workload."""
    sys.path.insert(0, os.path.join(ROOT, "oracle"))
    from oracle import Oracle
    from pdflow import synth_pair_float
    threads = min(16, os.cpu_count() or 1)
    I1, Z1, I2, Z2 = synth_pair_float(seed, 480 // cam_mode, 640 // cam_mode)
    o = Oracle(rows, cam_mode, threads=threads)

    def timed(budget):
        n, t0 = 0, time.perf_counter()
        while True:
            o.solve_pair(I1, Z1, I2, Z2)
            n += 1
            el = time.perf_counter() - t0
            if el >= budget or n >= 200:
                return n, el

    n, el = timed(budget_s)
    o.set_threads(1)
    n1, el1 = timed(budget_s / 2)
    model = "unknown"
    try:
        with open("/proc/cpuinfo") as f:
            model = next((ln.split(":", 1)[1].strip() for ln in f if ln.startswith("model name")), model)
    except OSError:
        pass
    return {"value": n / el, "unit": "pairs/s", "cores": threads, "kind": "port",
            "sample": f"{n} pairs of {480 // cam_mode}x{640 // cam_mode} (rows={rows}) in {el:.1f} s, "
                      f"C oracle -O3 OpenMP over pixels, {threads} threads",
            "value_1_thread": n1 / el1,
            "sample_1_thread": f"{n1} pairs in {el1:.1f} s on 1 thread",
            "cpu_model": model, "host_cpus": os.cpu_count()}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=20)
    ap.add_argument("--warmup", type=int, default=3)
    ap.add_argument("--rows", type=int, default=480)
    ap.add_argument("--seed", type=int, default=1)
    ap.add_argument("--no-graph", action="store_true")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=10.0)
    ap.add_argument("--tb-k", type=int, default=None,
                    help="iterations per launch of the temporally blocked kernel (1 = off)")
    ap.add_argument("--tb-min", type=int, default=None, help="smallest level (pixels) it runs on")
    ap.add_argument("--tb-shape", type=int, default=None,
                    help="batched/learned: region shape of the temporally blocked kernel (0: 64x32, "
                         "1: 32x32)")
    ap.add_argument("--workload", choices=("plain", "batched", "mixed", "stream", "learned"), default="learned",
                    help="plain: one pair per step; batched: --batch-pairs pairs of --rows per "
                         "step; mixed: 128 pairs (50%% 480x640, 30%% 240x320, 20%% 120x160); "
                         "learned: --batch-pairs pairs of --rows per step, network warm start at "
                         "60x80, calibrated schedule and acceptance rule (--rule)")
    ap.add_argument("--batch-pairs", type=int, default=256)
    ap.add_argument("--rule", default=RULE_JSON,
                    help="learned: the calibrated schedule and acceptance bounds (pdflow/calibrate.py)")
    ap.add_argument("--weights", default=os.path.join(ROOT, "pd-flow_amd", "pdflow", "weights", "warmstart_qvga.bin"))
    ap.add_argument("--autotune", default=None, metavar="CACHE",
                    help="batched/learned: pick the kernel configuration with pdf_batch_autotune "
                         "(cached in the file CACHE) before timing")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="BYTES",
                    help="first run a plain device copy of BYTES three times (the calibration point "
                         "of HBM byte counters under rocprofv3 --pmc)")
    ap.add_argument("--schedule", default=None, metavar="N,N,...",
                    help="learned: the whole per-level iteration schedule (coarse to fine) in place of "
                         "the calibrated one")
    ap.add_argument("--continue-iters", default="16,16,16", metavar="N[,N...]",
                    help="learned: pairs outside the rule's bounds first get N more finest-level "
                         "iterations from where they stand (pdf_batch_continue), stage by stage for "
                         "a list, each stage re-checking the rule; the plain path then re-solves "
                         "those still outside (0: no continuation)")
    ap.add_argument("--no-serving", action="store_true",
                    help="learned: skip the serving rate (inputs in and flows out over PCIe every step)")
    ap.add_argument("--pipeline", type=int, choices=(0, 1), default=1,
                    help="learned: 1 = two batches in flight on their own streams, so that one batch's "
                         "continuation and fallback overlap the next batch's solve; 0 = one batch, "
                         "steps back to back")
    args = ap.parse_args()

    if args.gpus != 1:
        print(f"bench.py: this build runs on one GPU only; --gpus {args.gpus} is not supported "
              "(pairs are independent, a multi-GPU run would shard them with no exchange).",
              file=sys.stderr)
        sys.exit(2)

    from pdflow import Solver, device_count, level_table, synth_pair_float
    from pdflow.lib import DeviceBuffer
    if device_count() < 1:
        print("bench.py: no GPU visible", file=sys.stderr)
        sys.exit(1)

    if args.copy_calibration:
        from pdflow import hip_library
        nb = args.copy_calibration // 16 * 16
        src, dst = DeviceBuffer(nb), DeviceBuffer(nb)
        for _ in range(3):
            if hip_library().pdf_device_copy(dst.ptr, src.ptr, nb):
                raise RuntimeError("pdf_device_copy failed")
        log(f"bench: copy calibration, 3 copies of {nb} bytes")
        src.free()
        dst.free()

    if args.rows not in (15, 30, 60, 120, 240, 480):
        print("bench.py: --rows must be one of 15, 30, 60, 120, 240, 480", file=sys.stderr)
        sys.exit(2)
    if args.workload == "stream":
        return run_stream(args)
    if args.workload == "learned":
        return run_learned(args)
    if args.workload != "plain":
        return run_batched(args)
    rows, cam_mode = args.rows, 480 // args.rows   # input frames at the solve resolution
    table = level_table(rows, cam_mode)
    iters = [t[4] for t in table]
    H, W = 480 // cam_mode, 640 // cam_mode
    P = H * W
    I1, Z1, I2, Z2 = synth_pair_float(args.seed, H, W)
    bufs = []
    for a in (I1, Z1, I2, Z2):
        b = DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    s = Solver(rows, cam_mode)
    lib, ctx = s.lib, s.ctx
    if args.tb_k is not None or args.tb_min is not None:
        s.set_iteration_kernel(args.tb_k if args.tb_k is not None else 6,
                               args.tb_min if args.tb_min is not None else 120 * 160)
    arr = (np.ctypeslib.ctypes.c_uint * len(iters))(*iters)
    use_graph = 0 if args.no_graph else 1

    def step():
        rc = lib.pdf_solve_pair_device(ctx, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, arr, use_graph)
        if rc:
            raise RuntimeError(lib.pdf_last_error(ctx).decode())

    log(f"bench: plain pair {H}x{W} rows={rows} iters={iters} graph={use_graph}")
    for _ in range(args.warmup):
        step()
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        step()
    s.synchronize()
    el = time.perf_counter() - t0
    pairs_s = args.steps / el
    log(f"bench: {args.steps} steps in {el:.3f} s -> {pairs_s:.1f} pairs/s")

    # roofline of the dominant kernel (fused primal-dual iteration): HIP events around every
    # launch on the solver's stream, over an eager pass of the same steps
    s.enable_kernel_timing(True)
    kt_steps = max(1, min(args.steps, 20))
    for _ in range(kt_steps):
        rc = lib.pdf_solve_pair_device(ctx, bufs[0].ptr, bufs[1].ptr, bufs[2].ptr, bufs[3].ptr, arr, 0)
        if rc:
            raise RuntimeError(lib.pdf_last_error(ctx).decode())
    timing = s.kernel_timing()
    s.enable_kernel_timing(False)
    rl = roofline("primal-dual iteration launches (k_pd_tb<K>, k_pd_level, k_pd_fused)", timing)
    log(f"bench: iteration launches {timing[0]}, avg {rl['avg_launch_us']:.2f} us, {rl['achieved']:.0f} GB/s")

    # flow sanity: the solve produced finite values
    dx, dy, dz = s.motion_field()
    assert np.isfinite(dx).all() and np.isfinite(dy).all() and np.isfinite(dz).all()

    out = {
        "metric": METRIC,
        "value": round(pairs_s, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded RGB-D generator, seed {args.seed})",
        "config": {
            "workload": f"plain PD-Flow single pair {H}x{W}, rows={rows}, {len(iters)} levels, "
                        f"iterations {iters}, drop-in path" + (", hipGraph" if use_graph else ""),
            "rows": rows, "cols": rows * 4 // 3, "cam_mode": cam_mode, "pairs_per_step": 1,
            "pixel_iterations_per_pair": int(sum(r * c * n for (_, r, c, _, n) in table)),
        },
        "roofline": rl,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(rows, cam_mode, args.seed, args.cpu_seconds)
    print(json.dumps(out), flush=True)


def run_stream(args):
    """Streaming: one new frame per step (one pyramid per frame, the previous one reused),
    then the solve of (previous frame, new frame)."""
    from pdflow import Solver, level_table, synth_sequence_float
    from pdflow.lib import DeviceBuffer
    rows, cam = args.rows, 480 // args.rows
    H, W = 480 // cam, 640 // cam
    frames = []
    for t in range(8):
        I, Z = synth_sequence_float(args.seed, H, W, t)
        bI, bZ = DeviceBuffer(I.nbytes), DeviceBuffer(Z.nbytes)
        bI.upload(I)
        bZ.upload(Z)
        frames.append((bI, bZ))
    s = Solver(rows, cam)
    graph = not args.no_graph
    n = 0
    for _ in range(args.warmup + 1):
        s.stream_frame_device(*frames[n % 8], graph=graph)
        n += 1
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(args.steps):
        s.stream_frame_device(*frames[n % 8], graph=graph)
        n += 1
    s.synchronize()
    el = time.perf_counter() - t0
    rate = args.steps / el
    log(f"bench: streaming {H}x{W}: {args.steps} frames in {el:.3f} s -> {rate:.1f} flows/s")
    table = level_table(rows, cam)
    out = {"metric": METRIC, "value": round(rate, 3), "unit": "pairs/s", "n_gpus": 1,
           "steps": args.steps, "warmup": args.warmup, "ms_per_step": round(el / args.steps * 1e3, 4),
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": f"synthetic (seeded moving scene {args.seed}, 8 frames cycled)",
           "config": {"workload": f"streaming plain PD-Flow {H}x{W}, rows={rows}: one new frame "
                                  "and one pyramid per step, then the solve" + (", hipGraph" if graph else ""),
                      "pixel_iterations_per_pair": int(sum(r * c * k for (_, r, c, _, k) in table))}}
    print(json.dumps(out), flush=True)


def run_batched(args):
    """Batched ragged solve: one step = one batched solve of all pairs (shared launches)."""
    from pdflow import Batch, level_table, synth_pair_float
    if args.workload == "mixed":
        rows_list = [480] * 64 + [240] * 38 + [120] * 26
        name = "batched plain PD-Flow, 128 mixed pairs (64x 480x640, 38x 240x320, 26x 120x160)"
    else:
        rows_list = [args.rows] * args.batch_pairs
        name = f"batched plain PD-Flow, {args.batch_pairs} pairs of {args.rows}x{args.rows * 4 // 3}"
    b = Batch(rows_list)
    if args.tb_k is not None or args.tb_min is not None:
        b.set_iteration_kernel(args.tb_k if args.tb_k is not None else 4,
                               args.tb_min if args.tb_min is not None else 4800)
    if args.tb_shape is not None:
        b.set_tb_shape(args.tb_shape)
    distinct = {}
    for i, r in enumerate(rows_list):
        key = (r, i % 16)
        if key not in distinct:
            cam = 480 // r
            distinct[key] = synth_pair_float(args.seed + 1000 * r + i % 16, 480 // cam, 640 // cam)
        b.set_inputs(i, *distinct[key])
    log(f"bench: {name}")
    if args.autotune:
        log(f"bench: autotuned (shape, K, ms) = {b.autotune(args.autotune)}")
    for _ in range(max(1, args.warmup)):
        b.solve(graph=not args.no_graph)
    b.synchronize()
    log(f"bench: {b.launch_count()} launches per batched solve")
    t0 = time.perf_counter()
    for _ in range(args.steps):
        b.solve(graph=not args.no_graph)
    b.synchronize()
    el = time.perf_counter() - t0
    pairs_s = len(rows_list) * args.steps / el
    log(f"bench: {args.steps} batched solves in {el:.3f} s -> {pairs_s:.1f} pairs/s")
    b.enable_kernel_timing(True)
    b.solve(graph=False)
    timing = b.kernel_timing()
    b.enable_kernel_timing(False)
    rl = roofline("kb_pd_tb / kb_pd_level (batched primal-dual iterations)", timing)
    log(f"bench: PD launches {timing[0]}, {timing[1]:.3f} ms, {rl['achieved']:.0f} GB/s")
    for i in (0, len(rows_list) - 1):
        for a in b.flow(i):
            assert np.isfinite(a).all()
    pix_pair = {r: int(sum(rr * c * n for (_, rr, c, _, n) in level_table(r, 480 // r))) for r in set(rows_list)}
    out = {
        "metric": METRIC,
        "value": round(pairs_s, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": f"synthetic (seeded RGB-D generator, 16 seeds per size from {args.seed})",
        "config": {"workload": name + (", hipGraph" if not args.no_graph else ""),
                   "pairs_per_step": len(rows_list),
                   "pixel_iterations_per_step": int(sum(pix_pair[r] for r in rows_list))},
        "roofline": rl,
    }
    print(json.dumps(out), flush=True)


def plain_single_pair_rate(rows, seed, steps=20):
    """pairs/s of the plain single-pair path (drop-in, hipGraph) on one pair of `rows`"""
    from pdflow import Solver, synth_pair_float
    from pdflow.lib import DeviceBuffer
    cam = 480 // rows
    planes = synth_pair_float(seed, 480 // cam, 640 // cam)
    bufs = []
    for a in planes:
        b = DeviceBuffer(a.nbytes)
        b.upload(a)
        bufs.append(b)
    s = Solver(rows, cam)
    for _ in range(3):
        s.solve_pair_device(bufs, graph=True)
    s.synchronize()
    t0 = time.perf_counter()
    for _ in range(steps):
        s.solve_pair_device(bufs, graph=True)
    s.synchronize()
    return steps / (time.perf_counter() - t0)


def run_learned(args):
    """Batched + learned warm start: one step = one batched solve of all pairs (pyramids, one
    network pass over their 60x80 levels, injection, solves with the calibrated schedule), then the
    calibrated acceptance rule, continuation of the pairs outside it and the plain re-solve of
    those still outside."""
    from pdflow import Batch, Net, synth_pairs_float
    rows, n = args.rows, args.batch_pairs
    cam = 480 // rows
    for path in (args.weights, args.rule):
        if not os.path.exists(path):
            raise SystemExit(f"bench.py: {path} is missing")
    with open(args.rule) as f:
        rule = json.load(f)
    log(f"bench: rendering {n} pairs of {480 // cam}x{640 // cam}")
    *planes, gt = synth_pairs_float(args.seed, n, 480 // cam, 640 // cam, with_gt=True)
    net = Net(max_batch=n, h=60, w=80)
    net.load(args.weights)
    stages = [int(x) for x in str(args.continue_iters).split(",") if int(x) > 0]
    sched = [int(x) for x in args.schedule.split(",")] if args.schedule else list(rule["schedule"])
    if len(sched) != 6 and rows == 480:
        raise SystemExit("bench.py: --schedule needs 6 levels")
    emax, rmax = rule["energy_per_pixel_max"], rule["residual_ratio_max"]
    P = rows * (rows * 4 // 3)

    def make(iters=None):
        b = Batch([rows] * n, iters=None if iters is None else [iters] * n)
        if args.tb_shape is not None:
            b.set_tb_shape(args.tb_shape)
        for i in range(n):
            b.set_inputs(i, *(p[i] for p in planes))
        return b

    learned = make(sched)
    learned.set_learned(net)
    if args.autotune:
        log(f"bench: autotuned (shape, K, ms) = {learned.autotune(args.autotune)}")
    # a second batch of the same pairs with its own network instance (activations are per
    # network) and stream, for the pipelined steps
    batches = [learned]
    if args.pipeline:
        net2 = Net(max_batch=n, h=60, w=80)
        net2.load(args.weights)
        learned2 = make(sched)
        learned2.set_learned(net2)
        if args.autotune:
            learned2.autotune(args.autotune)
        batches.append(learned2)

    def outside(e, r, r0):
        return ~(e <= emax * P) | ~(r <= rmax * r0)

    def finish(b):
        """the calibrated rule after the learned solve of batch b: pairs outside the bounds get the
        continuation stages (re-checked alone after each), the plain path re-solves those still
        outside; returns (continued, fell back) flags"""
        e, r, r0 = b.quality()
        bad = np.nonzero(outside(e, r, r0))[0]
        cont = np.zeros(n, np.uint8)
        for extra in stages:
            if not len(bad):
                break
            cont[bad] = 1
            b.continue_finest(bad, extra)
            e2, r2, r02 = b.quality(bad)
            bad = bad[outside(e2, r2, r02)]
        b.resolve_plain(bad)
        flags = np.zeros(n, np.uint8)
        flags[bad] = 1
        return cont, flags

    def run_steps(k, batches, finish):
        """k steps, each a whole batch (solve, rule, continuation, fallback).  Pipelined: the next
        batch's solve is queued on its own stream before this batch's finish, whose host
        synchronisations wait on this batch's stream only.  That solve waits on the GPU for this
        batch's solve (wait_for right behind it), so the solves run one at a time at full width and
        only the finish overlaps the next solve."""
        graph = not args.no_graph
        nb = len(batches)
        batches[0].solve(graph=graph)
        flags = None
        for s in range(k):
            cur = batches[s % nb]
            if nb > 1 and s + 1 < k:
                batches[(s + 1) % nb].wait_for(cur)
                batches[(s + 1) % nb].solve(graph=graph)
            flags = finish(cur)[1]
            if nb == 1 and s + 1 < k:
                cur.solve(graph=graph)
        for b in batches:
            b.synchronize()
        return flags

    flags = run_steps(max(1, args.warmup), batches, finish)
    t0 = time.perf_counter()
    flags = run_steps(args.steps, batches, finish)
    el = time.perf_counter() - t0
    pairs_s = n * args.steps / el
    log(f"bench: {args.steps} learned steps in {el:.3f} s -> {pairs_s:.1f} pairs/s "
        f"(fallback {int(flags.sum())}/{n})")

    # serving: each step also copies its pairs' inputs in from page-locked host memory before the
    # solve and their flows out after the rule, on the batch's stream, so that one batch's copies
    # overlap the other batch's kernels (the PCIe-inclusive rate; `value` keeps the inputs resident)
    serving = None
    if not args.no_serving:
        from pdflow import PinnedBuffer
        src = PinnedBuffer(n * 4 * P)
        src.array[:] = np.stack([np.ascontiguousarray(p, np.float32).reshape(n, -1) for p in planes], axis=1).ravel()
        dsts = [PinnedBuffer(n * 3 * P) for _ in batches]

        def run_serving(k):
            """run_steps with the copies: a batch's inputs go in before its solve, its flows come
            out after its finish"""
            nb = len(batches)

            def start(b, after=None):
                b.upload_inputs_async(src.array)
                if after is not None:
                    b.wait_for(after)
                b.solve(graph=not args.no_graph)

            start(batches[0])
            for s in range(k):
                cur = batches[s % nb]
                if nb > 1 and s + 1 < k:
                    start(batches[(s + 1) % nb], after=cur)
                finish(cur)
                cur.download_flows_async(dsts[s % nb].array)
                if nb == 1 and s + 1 < k:
                    start(cur)
            for b in batches:
                b.synchronize()

        run_serving(max(1, args.warmup))
        t0 = time.perf_counter()
        run_serving(args.steps)
        serving = n * args.steps / (time.perf_counter() - t0)
        # the flows that came back are the batch's own
        last = batches[(args.steps - 1) % len(batches)]
        got = dsts[(args.steps - 1) % len(batches)].array
        for i in (0, n - 1):
            for k, w in enumerate(last.flow(i)):
                if not np.array_equal(got[(3 * i + k) * P:(3 * i + k + 1) * P], w):
                    raise SystemExit("bench.py: a flow copied out asynchronously differs from the batch's")
        log(f"bench: serving (inputs in and flows out over PCIe every step) {serving:.1f} pairs/s")
        for buf in [src] + dsts:
            buf.free()

    # the rule as the configuration words it: every step also solves every pair with the plain
    # path, whose energy and residual are the references; pairs more than 5% above them get the
    # same continuation stages, then the plain re-solve
    plain = make()
    steps_ref = max(1, args.steps // 2)

    def finish_ref(b):
        plain.solve(graph=not args.no_graph)
        pe, pr, _ = plain.quality()
        e, r, _ = b.quality()
        bad = np.nonzero(~(e <= 1.05 * pe) | ~(r <= 1.05 * pr))[0]
        for extra in stages:
            if not len(bad):
                break
            b.continue_finest(bad, extra)
            e2, r2, _ = b.quality(bad)
            bad = bad[~(e2 <= 1.05 * pe[bad]) | ~(r2 <= 1.05 * pr[bad])]
        b.resolve_plain(bad)
        flags = np.zeros(n, np.uint8)
        flags[bad] = 1
        return None, flags

    run_steps(1, [learned], finish_ref)
    t0 = time.perf_counter()
    run_steps(steps_ref, [learned], finish_ref)
    plain.synchronize()
    rule_as_worded = n * steps_ref / (time.perf_counter() - t0)
    log(f"bench: rule as worded (plain solves inside the step) {rule_as_worded:.1f} pairs/s")

    # the learned solve alone (no rule), and the plain batched solve, for the same pairs
    t0 = time.perf_counter()
    for _ in range(args.steps):
        learned.solve(graph=not args.no_graph)
    learned.synchronize()
    learned_only = n * args.steps / (time.perf_counter() - t0)
    t0 = time.perf_counter()
    for _ in range(max(1, args.steps // 4)):
        plain.solve(graph=True)
    plain.synchronize()
    plain_batched = n * max(1, args.steps // 4) / (time.perf_counter() - t0)
    plain_single = plain_single_pair_rate(rows, args.seed)
    log(f"bench: learned alone {learned_only:.1f}, plain batched {plain_batched:.1f}, "
        f"plain single pair {plain_single:.1f} pairs/s")

    # quality on these pairs: end-point errors against the plain solve and the rendered truth, and
    # how the calibrated rule's decisions compare with the plain-referenced rule's
    def gt_epe(b):
        out = []
        for i in range(n):
            ok = planes[1][i] > 0
            d = np.sqrt(sum((f.astype(np.float64) - g) ** 2 for f, g in zip(b.flow(i), gt[i])))[ok]
            out.append(d.mean())
        return np.asarray(out)

    pe, pr, _ = plain.quality()
    learned.solve(graph=True)
    learned.synchronize()
    le, lr_, l0 = learned.quality()
    self_ok = ~outside(le, lr_, l0)
    plain_ok = (le <= 1.05 * pe) & (lr_ <= 1.05 * pr)
    epe_l = learned.epe(plain)
    gt_learned = gt_epe(learned)
    cont, flags = finish(learned)
    epe_f = learned.epe(plain)
    gt_plain, gt_final = gt_epe(plain), gt_epe(learned)
    ablation = make(sched)
    ablation.solve(graph=True)
    ablation.synchronize()
    epe_a = ablation.epe(plain)
    gt_abl = gt_epe(ablation)
    del ablation

    def stats(a):
        return {"mean": round(float(np.mean(a)), 6), "median": round(float(np.median(a)), 6),
                "p90": round(float(np.percentile(a, 90)), 6), "max": round(float(np.max(a)), 6)}

    # roofline of the dominant kernel, the temporally blocked iteration kb_pd_tb<K>: HIP events
    # around each of its launches on the batch's stream, over eager passes of the learned solve
    learned.enable_kernel_timing(True, tb_only=True)
    for _ in range(3):
        learned.solve(graph=False)
    timing = learned.kernel_timing()
    learned.enable_kernel_timing(False)
    key = f"learned-{n}x{rows}-{','.join(map(str, sched))}-shape{1 if args.tb_shape is None else args.tb_shape}"
    rl = roofline("kb_pd_tb<K, 32, 32> (batched, K = 2 to 5 primal-dual iterations per launch)", timing, key)
    ho = rule.get("heldout", {})
    out = {
        "metric": METRIC,
        "value": round(pairs_s, 3),
        "unit": "pairs/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": round(el / args.steps * 1e3, 4),
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32 (network bf16 x bf16 -> f32)",
        "data": f"synthetic (seeded RGB-D generator, seeds {args.seed}..{args.seed + n - 1})",
        "config": {"workload": f"batched + learned warm start, {n} pairs of {rows}x{rows * 4 // 3}: network at "
                               f"the 60x80 level, iterations {sched} (reference / {rule['factor']}, finest level "
                               f"calibrated), acceptance rule computed from the learned solve alone (finest-level "
                               f"energy <= {emax:.3f} per pixel and warped residual <= {rmax:.4f} x the zero-motion "
                               f"residual, bounds from the plain solver's training results), pairs outside it "
                               + (f"continued by {'+'.join(map(str, stages))} finest-level iterations, then "
                                  if stages else "")
                               + "re-solved by the plain path; rule, continuation and re-solves inside the timed "
                               "region" + (", hipGraph" if not args.no_graph else "")
                               + (", two batches in flight" if args.pipeline else ""),
                   "pairs_per_step": n, "batches_in_flight": len(batches),
                   "weights": os.path.relpath(args.weights, ROOT), "rule": os.path.relpath(args.rule, ROOT)},
        "vs_plain": {"plain_single_pair_pairs_s": round(plain_single, 3),
                     "plain_batched_pairs_s": round(plain_batched, 3),
                     "learned_without_rule_pairs_s": round(learned_only, 3),
                     "rule_as_worded_pairs_s": round(rule_as_worded, 3),
                     "serving_pairs_s": None if serving is None else round(serving, 3),
                     "serving_pcie_bytes_per_pair": {"to_device": 16 * P, "to_host": 12 * P},
                     "speedup_vs_plain_single_pair": round(pairs_s / plain_single, 3),
                     "speedup_vs_plain_batched": round(pairs_s / plain_batched, 3)},
        "rule": {"accepted_fraction": round(float(self_ok.mean()), 4),
                 "plain_referenced_rule_accepts": round(float(plain_ok.mean()), 4),
                 "agreement_with_plain_referenced_rule": round(float((self_ok == plain_ok).mean()), 4),
                 "accepted_but_plain_rule_rejects": round(float((self_ok & ~plain_ok).mean()), 4),
                 "heldout_calibration_check": {k: ho.get(k) for k in ("pairs", "agree", "rule_accepts",
                                                                      "plain_rule_accepts",
                                                                      "accepted_by_rule_rejected_by_plain_rule",
                                                                      "corrupted_network")},
                 "continued_fraction": round(float(cont.mean()), 4), "continue_iters": stages,
                 "fallback_fraction": round(float(flags.mean()), 4)},
        "quality_m": {"epe_vs_plain_learned": stats(epe_l), "epe_vs_plain_final": stats(epe_f),
                      "epe_vs_plain_ablation": stats(epe_a), "epe_vs_truth_plain": stats(gt_plain),
                      "epe_vs_truth_learned": stats(gt_learned), "epe_vs_truth_final": stats(gt_final),
                      "epe_vs_truth_ablation": stats(gt_abl),
                      "energy_ratio_to_plain": stats(le / pe), "residual_ratio_to_plain": stats(lr_ / pr)},
        "roofline": rl,
    }
    if not args.no_cpu_baseline:
        log("bench: CPU baseline ...")
        out["cpu_baseline"] = cpu_baseline(rows, cam, args.seed, args.cpu_seconds)
        out["cpu_baseline"]["sample"] += " (plain path: the CPU oracle has no learned path)"
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

#!/usr/bin/env python3
# SPDX-License-Identifier: MIT
"""Benchmark of the rigid AVBD step on one MI355X.

Workload: a 32768-body pile (half boxes with edges U[0.2, 0.6] m, half
analytic spheres with radius U[0.1, 0.3] m, densities log-uniform in
[200, 5000] kg/m^3, friction U[0.3, 0.8], the reference's default body
options: margin 5 mm, damping 0.05, speed limits 100 m/s and 100 rad/s)
dropped from a jittered 32^3 grid into a static five-wall bin; dt = 1/60,
legacy AVBD profile with 10 iterations (pose phase and velocity phase),
FP32, 600 steps.  The timed window is the last K steps of the 600-step run
(the settled pile), after max(W, 600 - K) untimed steps (defaults: K = 60,
W = 5, so steps 540..600 are timed).

value    = AVBD body-iterations per second of wall time over the K steps
           (whole steps: predictor, broadphase, narrowphase, contact cache,
           prepare, solve);
roofline = the dominant kernel, the dataflow colour sweep k_sweep (one
           launch per iteration; --per-color-sweep: k_stage_color, one launch
           per colour per iteration): its algorithmic bytes per launch over
           its average launch duration, both from the launches of every
           4th timed step, each bracketed by HIP events on the world's
           stream;
cpu_baseline = the oracle (same algorithm, serial, one thread) on a bounded
           sample of the same workload, body-iterations per second of its
           solve phase.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
for sub in ("parallel-mater_amd", "tests"):
    path = os.path.join(ROOT, sub)
    if path not in sys.path:
        sys.path.insert(0, path)

METRIC = "AVBD body-iters/s @32k bodies; sim-steps/s to 1e-4 residual, MLP vs plain"
HBM_PEAK_GBS = 8000.0


KERNEL_TIMING_STRIDE = 4


RUN_STEPS = 600   # the pile workload's run length


def traffic_from_profiles(window, kernel):
    """HBM bytes per colour-sweep launch measured by rocprofv3 --pmc
    (FETCH_SIZE + WRITE_SIZE, scaled by a copy-kernel calibration) over the
    same timed window of the same command and kernel, when recorded; None
    otherwise."""
    path = os.path.join(ROOT, "profiles", "color_sweep_traffic.json")
    try:
        with open(path) as f:
            rec = json.load(f)
    except (OSError, ValueError):
        return None
    if list(rec.get("window", [])) != list(window) or rec.get("kernel") != kernel:
        return None
    return rec.get("bytes_per_launch")


def summarize_pmc(fetch_dir, write_dir, cal_fetch_dir, cal_write_dir, bench_json, out_path):
    """Per-dispatch HBM traffic of the colour sweep from rocprofv3 --pmc runs
    of this script (FETCH_SIZE, then WRITE_SIZE, each restricted to the sweep
    kernel and k_bp_fill) over the bench line's timed window, each counter
    scaled by what it reads for a device copy of known size (the
    --copy-calibration runs).  k_bp_fill runs once per step and marks the
    steps.  Writes the record that the bench line's `traffic` reads."""
    import csv
    import glob

    def dispatches(root):
        rows = {}
        for path in glob.glob(os.path.join(root, "**", "*counter_collection.csv"), recursive=True):
            with open(path) as f:
                for r in csv.DictReader(f):
                    key = int(r["Dispatch_Id"])
                    name = r["Kernel_Name"]
                    rows.setdefault(key, [name, 0.0])[1] += float(r["Counter_Value"])
        return [rows[k] for k in sorted(rows)]

    with open(bench_json) as f:
        line = json.loads(f.read().strip().splitlines()[-1])
    kernel = line["roofline"]["kernel"].split(" (")[0]
    first, last = line["untimed_steps"], line["untimed_steps"] + line["steps"]

    def window_mean(root):
        # without k_bp_fill in the run there are no step marks: every
        # dispatch of the kernel counts (warm-up steps included)
        rows = dispatches(root)
        marked = any("k_bp_fill" in name for name, _ in rows)
        step, values = -1, []
        for name, value in rows:
            if "k_bp_fill" in name:
                step += 1
            elif kernel in name and (not marked or first <= step < last):
                values.append(value)
        return sum(values) / max(1, len(values)), len(values)

    def copy_kib(root):
        # the calibration's largest dispatch: one 256 MiB copy (the fill
        # before it writes as much and reads nothing)
        return max(v for _, v in dispatches(root))

    fetch, count = window_mean(fetch_dir)
    write, _ = window_mean(write_dir)
    copy_bytes = float(256 << 20)
    fetch_scale = copy_bytes / (copy_kib(cal_fetch_dir) * 1024.0)
    write_scale = copy_bytes / (copy_kib(cal_write_dir) * 1024.0)
    rec = {
        "kernel": kernel,
        "workload": line["config"]["workload"],
        "window": [first, last],
        "fetch_kib_per_dispatch": fetch,
        "write_kib_per_dispatch": write,
        "fetch_scale": fetch_scale,
        "write_scale": write_scale,
        "bytes_per_launch": (fetch * fetch_scale + write * write_scale) * 1024.0,
        "algorithmic_bytes_per_launch": line["roofline"]["algorithmic_bytes_per_launch"],
        "dispatches": count,
        "steps_marked": any("k_bp_fill" in name for name, _ in dispatches(fetch_dir)),
    }
    with open(out_path, "w") as f:
        json.dump(rec, f, indent=1)
    return rec


def cpu_model():
    try:
        with open("/proc/cpuinfo") as f:
            for line in f:
                if line.startswith("model name"):
                    return line.split(":", 1)[1].strip()
    except OSError:
        pass
    return "unknown"


def cpu_baseline(scene, states, seconds, profile, iterations):
    import _oracle as O
    w = scene.build(O.OracleWorld(profile=profile, base_iterations=iterations, lanes=0))
    w.set_states(states)
    w.timing(reset=True)
    steps = 0
    start = time.process_time()
    while steps == 0 or time.process_time() - start < seconds:
        w.step(1)
        steps += 1
    t = w.timing()
    return {
        "value": t["body_iterations"] / (t["solve_ms"] / 1e3),
        "unit": "body-iters/s",
        "cores": 1,
        "cpu": cpu_model(),
        "kind": "port",
        "sample": f"{steps} steps of the same pile from the GPU's state at the timed window start "
                  f"(cold contact cache), serial oracle, solve phase only; "
                  f"{t['step_ms'] / steps:.1f} ms per whole step",
        "steps_per_s": steps / (t["step_ms"] / 1e3),
    }


def warmstart_bench(args, tolerance=1e-4, cap=64, samples_wanted=1_000_000):
    """The learned warm start against plain AVBD, both iterating each phase
    to an RMS constraint residual of 1e-4 (the budget capping at 64): log
    cold-contact samples from plain runs of four seeded 32768-body piles
    (seeds 11-14), train the network on 10^6 of them on the GPU, then run the
    bench pile (seed 2) with each warm start (plain; the predicted penalty;
    with --warm-lambda also the predicted penalty and, where a contact starts
    in violation, lambda): warm-up steps, then timed steps, each step checked
    for the fallback. value = sim-steps/s with the predicted penalty."""
    import pmamd
    import scenarios as S
    t0 = time.perf_counter()
    logs = []
    per_seed = samples_wanted // 4
    for seed in (11, 12, 13, 14):
        # every cold contact of the 240 steps, then a uniform sample of them
        # (falling, impact and settling phases alike)
        w = S.pile(side=args.side, seed=seed).build(pmamd.World(profile=0, base_iterations=args.iterations))
        w.set_contact_log(8_000_000)
        w.step(240)
        log = w.contact_log()
        w.close()
        keep = np.random.default_rng(seed).choice(len(log), size=min(per_seed, len(log)), replace=False)
        logs.append(log[np.sort(keep)])
        print(f"# bench: seed {seed}: {len(log)} cold-contact samples, {len(logs[-1])} kept", file=sys.stderr,
              flush=True)
    samples = np.concatenate(logs)
    t1 = time.perf_counter()
    params, loss = pmamd.warmstart_train(samples, epochs=6, learning_rate=1e-3, seed=1)
    t2 = time.perf_counter()
    var = float(np.mean((samples[:, 16:] - samples[:, 16:].mean(0)) ** 2))
    print(f"# bench: trained on {len(samples)} samples in {t2 - t1:.1f} s, loss {loss[-1]:.4g} "
          f"(target variance {var:.4g})", file=sys.stderr, flush=True)
    runs = {}
    scene = S.pile(side=args.side, seed=2, analytic_spheres=True)
    modes = ("plain", "learned", "learned_lambda") if args.warm_lambda else ("plain", "learned")
    for mode in modes:
        w = scene.build(pmamd.World(profile=0, base_iterations=cap, residual_tolerance=tolerance))
        if mode != "plain":
            w.set_warmstart(params, penalty=True, lam=mode == "learned_lambda")
        w.step(args.warmup)
        pmamd.lib().pmhDeviceSynchronize()
        its = []
        start = w.warmstart_stats()
        ts = time.perf_counter()
        for _ in range(args.steps):
            w.step(1)
            its.append(w.residuals()[1].copy())
        pmamd.lib().pmhDeviceSynchronize()
        wall = time.perf_counter() - ts
        a = np.array(its, dtype=np.float64)
        st = w.warmstart_stats()
        runs[mode] = {"steps_per_s": args.steps / wall, "pose_iterations": float(a[:, 0].mean()),
                      "velocity_iterations": float(a[:, 1].mean()),
                      "capped_steps": int((a[:, 0] >= cap).sum()),
                      "fallback_rate": (st[1] - start[1]) / max(1, st[0] - start[0]) if mode != "plain" else None}
        print(f"# bench: {mode}: {runs[mode]}", file=sys.stderr, flush=True)
        w.close()
    name, cus = pmamd.device_info()
    p, l_ = runs["plain"], runs["learned"]
    return {
        "metric": METRIC, "value": l_["steps_per_s"], "unit": "sim-steps/s to 1e-4 residual",
        "n_gpus": 1, "steps": args.steps, "warmup": args.warmup,
        "ms_per_step": 1e3 / l_["steps_per_s"], "higher_is_better": True, "scaling": "weak",
        "vs_baseline": None, "dtype": "f32 (network bf16 on MFMA, fp32 accumulation)", "data": "synthetic",
        "config": {"workload": f"learned warm start vs plain AVBD on the {args.side ** 3}-body pile (seed 2), "
                               f"each phase iterated to an RMS constraint residual of {tolerance} (cap {cap}); "
                               f"network trained on {len(samples)} cold-contact samples logged from plain runs of "
                               f"four other seeded piles",
                   "device": f"{name} ({cus} CUs)"},
        "plain": p, "learned": l_, "learned_lambda": runs.get("learned_lambda"),
        "iteration_ratio": p["pose_iterations"] / l_["pose_iterations"],
        "steps_per_s_ratio": l_["steps_per_s"] / p["steps_per_s"],
        "training": {"samples": int(len(samples)), "epochs": 6, "seconds": t2 - t1,
                     "loss_per_epoch": [float(x) for x in loss], "target_variance": var,
                     "logging_seconds": t1 - t0},
        "roofline": None, "cpu_baseline": None,
    }


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=60)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--workload", choices=["pile", "chains", "batch", "warmstart"], default="pile",
                    help="pile: 32^3 boxes and spheres in a bin; chains: capsule chains with ball and hinge joints; "
                         "batch: independent scenes of 8-512 bodies in one world")
    ap.add_argument("--scenes", type=int, default=4096, help="scene count of the batch workload")
    ap.add_argument("--side", type=int, default=32)
    ap.add_argument("--chains", type=int, default=4096, help="chain count of the chains workload (64 links each)")
    ap.add_argument("--iterations", type=int, default=10)
    ap.add_argument("--profile", type=int, default=0, help="0 legacy (10 iterations), 1 realtime v1")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--warm-lambda", action="store_true",
                    help="warmstart workload: also run the predicted-lambda mode")
    ap.add_argument("--no-cpu-baseline", action="store_true")
    ap.add_argument("--no-warmstart", action="store_true",
                    help="pile: skip the learned-vs-plain warm-start half of the metric")
    ap.add_argument("--cooperative", action="store_true", help="whole solve in one cooperative launch (grid barriers)")
    ap.add_argument("--per-color-sweep", action="store_true",
                    help="one colour-sweep launch per colour per iteration instead of the dataflow sweep")
    ap.add_argument("--lanes", type=int, default=0, help="lanes per body in the colour sweep (8, 16, 32, 64; 0 = library default)")
    ap.add_argument("--copy-calibration", type=int, default=0, metavar="MIB",
                    help="only run a device-to-device copy of MIB MiB three times (scale for rocprofv3 FETCH_SIZE/WRITE_SIZE)")
    ap.add_argument("--summarize-pmc", nargs=5, metavar=("FETCH_DIR", "WRITE_DIR", "CAL_FETCH_DIR", "CAL_WRITE_DIR",
                                                         "BENCH_JSON"),
                    help="no GPU: write profiles/color_sweep_traffic.json from rocprofv3 --pmc output directories "
                         "(FETCH_SIZE and WRITE_SIZE of a bench run and of --copy-calibration 256) and that run's "
                         "bench line")
    ap.add_argument("--pmc-out", default=os.path.join(ROOT, "profiles", "color_sweep_traffic.json"),
                    help="where --summarize-pmc writes its record (the bench line's traffic reads the default)")
    args = ap.parse_args()
    if args.summarize_pmc:
        rec = summarize_pmc(*args.summarize_pmc, args.pmc_out)
        print(json.dumps(rec))
        return
    if args.copy_calibration:
        import torch
        n = args.copy_calibration * (1 << 20) // 4
        src = torch.ones(n, dtype=torch.float32, device="cuda")
        dst = torch.empty_like(src)
        for _ in range(3):
            dst.copy_(src)
        torch.cuda.synchronize()
        print(json.dumps({"copy_bytes_per_launch": n * 4}))
        return
    if args.workload == "warmstart":
        if args.gpus != 1:
            print("bench.py runs on one GPU only.")
            sys.exit(2)
        print(json.dumps(warmstart_bench(args)), flush=True)
        return
    if args.gpus != 1:
        print(f"bench.py runs on one GPU only (asked for {args.gpus}); "
              "multi-GPU sharding of this path is a design note, not a measured configuration.")
        sys.exit(2)

    import pmamd
    import scenarios as S

    if args.workload == "batch":
        scene = S.batch(count=args.scenes, seed=7)
        dynamic = scene.n - args.scenes
        workload = (f"{args.scenes} independent scenes in one world, bodies per scene log-uniform in 8-512 "
                    f"(half boxes, half analytic spheres, pile materials), drops and stacks on a ground box each")
    elif args.workload == "chains":
        scene = S.chains(count=args.chains, links=64)
        dynamic = args.chains * 64
        workload = (f"{args.chains} chains of 64 capsule links (r 0.05 m, length 0.3 m) joined by ball and hinge "
                    f"joints in turn, every 8th link 100x heavier, each chain held by a ball joint to a static "
                    f"anchor, swinging into each other onto a ground box")
    else:
        scene = S.pile(side=args.side, seed=2, analytic_spheres=True)
        dynamic = args.side ** 3
        workload = (f"pile of {args.side ** 3} bodies (half boxes 0.2-0.6 m, half analytic spheres "
                    f"r 0.1-0.3 m) dropped into a static 5-wall bin")
    # Chains touch few bodies each: smaller pools (per body on average).
    pools = {"chains": dict(pairs_per_body=48, points_per_body=32),
             "batch": dict(pairs_per_body=40, points_per_body=40)}.get(args.workload, {})
    world = scene.build(pmamd.World(profile=args.profile, base_iterations=args.iterations,
                                    cooperative=args.cooperative, solve_lanes=args.lanes,
                                    per_color_sweep=args.per_color_sweep,
                                    kernel_timing_stride=KERNEL_TIMING_STRIDE, **pools))
    print(f"# bench: {scene.n} bodies, warmup {args.warmup}, timed {args.steps}", file=sys.stderr, flush=True)
    tuned = None
    if args.workload in ("batch", "chains") and not args.lanes:
        # Lane tile of the colour sweep picked by timing each on this world
        # (part of the warm-up).
        world.step(max(1, args.warmup // 2))
        tuned = pmamd.autotune_lanes(world, steps=4)
        print(f"# bench: autotuned lane tile {tuned[0]} ({tuned[1]})", file=sys.stderr, flush=True)
    def advance(steps):
        try:
            world.step(steps)
        except pmamd.PmhError:
            print(f"# bench: step failed, stats {world.stats()}", file=sys.stderr, flush=True)
            raise

    # The pile is timed over the last K steps of its 600-step run (settled);
    # the other workloads after W warm-up steps.
    settle = max(args.warmup, RUN_STEPS - args.steps) if args.workload == "pile" else args.warmup
    if settle:
        advance(settle)
    pmamd.lib().pmhDeviceSynchronize()
    start_states = world.states()
    world.totals(reset=True)
    world.solve_timing(reset=True)
    world.kernel_timing(reset=True)
    pmamd.lib().pmhDeviceSynchronize()
    latencies = []
    t0 = time.perf_counter()
    if args.workload == "batch":
        # One batch step per call: its latency as a caller sees it.
        for _ in range(args.steps):
            t1 = time.perf_counter()
            advance(1)
            pmamd.lib().pmhDeviceSynchronize()
            latencies.append((time.perf_counter() - t1) * 1e3)
    else:
        advance(args.steps)
    pmamd.lib().pmhDeviceSynchronize()
    wall = time.perf_counter() - t0
    totals = world.totals()
    solve_ms, launches = world.solve_timing()
    kt = world.kernel_timing()
    stats = world.stats()
    phases = world.phase_times()
    body_iters = totals["body_iterations"]
    if args.cooperative or kt["launches"] == 0:
        # One cooperative launch per step: the unit is the whole solve.
        unit_launches = max(1, totals["steps"])
        per_launch_bytes = totals["algorithmic_bytes"] / unit_launches
        per_launch_s = solve_ms / 1e3 / unit_launches
        kernel = "k_solve_all (one cooperative launch per step)"
    else:
        per_launch_bytes = kt["algorithmic_bytes"] / kt["launches"]
        per_launch_s = kt["milliseconds"] / 1e3 / kt["launches"]
        lanes_used = args.lanes or (tuned[0] if tuned else 32)
        joints = 'true' if scene.joints is not None else 'false'
        if args.per_color_sweep:
            kernel = f"k_stage_color<{lanes_used}, {joints}> (colour sweep, one launch per colour per iteration)"
        else:
            kernel = f"k_sweep<{lanes_used}, {joints}> (dataflow colour sweep, one launch per iteration)"
    achieved = per_launch_bytes / per_launch_s / 1e9
    window = (settle, settle + args.steps)
    traffic = (traffic_from_profiles(window, kernel.split(" (")[0])
               if args.workload == "pile" and args.side == 32 else None)
    name, cus = pmamd.device_info()
    line = {
        "metric": METRIC,
        "value": body_iters / wall,
        "unit": "body-iters/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "untimed_steps": settle,
        "ms_per_step": wall * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic",
        "config": {
            "workload": f"{workload}, dt 1/60, "
                        f"{'legacy AVBD, ' + str(args.iterations) + ' iterations' if args.profile == 0 else 'realtime v1, 8 sweeps'}, "
                        f"timed steps {settle}..{settle + args.steps}"
                        + (f" of a {RUN_STEPS}-step run (settled pile)" if args.workload == "pile" else ""),
            "bodies": scene.n,
            "dynamic_bodies": dynamic,
            "joints": 0 if scene.joints is None else len(scene.joints),
            "iterations": args.iterations if args.profile == 0 else 8,
            "profile": "legacy" if args.profile == 0 else "realtime-v1",
            "solve_lanes": args.lanes or (tuned[0] if tuned else 32),
            "device": f"{name} ({cus} CUs)",
        },
        "solve_phase_body_iters_per_s": body_iters / (solve_ms / 1e3),
        "solve_phase_gbs": totals["algorithmic_bytes"] / (solve_ms / 1e3) / 1e9,
        "steps_per_s": args.steps / wall,
        "solve_ms_per_step": solve_ms / args.steps,
        "velocity_phases": totals["velocity_phases"],
        "solve_launches_per_step": launches / max(1, args.steps),
        "last_step": {"phase_ms": {k: round(v, 4) for k, v in phases.items()}, **stats},
        **({"scenes": args.scenes, "scene_steps_per_s": args.scenes * args.steps / wall,
            "batch_latency_ms": {"p50": float(np.percentile(latencies, 50)),
                                 "p99": float(np.percentile(latencies, 99))},
            "autotune_ms_per_step": tuned[1] if tuned else None}
           if args.workload == "batch" else {}),
        "roofline": {
            "bound": "hbm",
            "achieved": achieved,
            "peak": HBM_PEAK_GBS,
            "unit": "GB/s",
            "frac": achieved / HBM_PEAK_GBS,
            "traffic": traffic,
            "kernel": kernel,
            "algorithmic_bytes_per_launch": per_launch_bytes,
            "avg_launch_ms": per_launch_s * 1e3,
            "timed_launches": kt["launches"],
            "timed_every_nth_step": KERNEL_TIMING_STRIDE,
        },
    }
    if not args.no_cpu_baseline:
        print("# bench: cpu baseline (oracle, one thread)", file=sys.stderr, flush=True)
        if args.workload == "batch":
            # The oracle steps a sample of 8 of the scenes, one after the
            # other, from rest.
            sample = S.Batch(scene.shapes, scene.scenes[:8], "sample")
            total = {"steps": 0, "body_iterations": 0, "solve_ms": 0.0, "step_ms": 0.0}
            import _oracle as O
            start = time.process_time()
            for s_ in range(8):
                w = sample.build_one(O.OracleWorld(profile=args.profile, base_iterations=args.iterations, lanes=0), s_)
                w.timing(reset=True)
                for _ in range(30):
                    w.step(1)
                t = w.timing()
                for k in total:
                    total[k] += t[k] if k != "steps" else 30
                if time.process_time() - start > args.cpu_seconds:
                    break
            line["cpu_baseline"] = {
                "value": total["body_iterations"] / (total["solve_ms"] / 1e3), "unit": "body-iters/s",
                "cores": 1, "kind": "port",
                "sample": f"the first {s_ + 1} scenes of the batch, 30 steps each from rest, one after the other, "
                          f"serial oracle, solve phase only",
                "scene_steps_per_s": total["steps"] / (total["step_ms"] / 1e3)}
        elif args.workload == "chains":
            # The oracle runs a 16-chain sample of the same construction
            # from rest (the full workload takes minutes per oracle step).
            sample = S.chains(count=16, links=64)
            line["cpu_baseline"] = cpu_baseline(sample, sample.states, args.cpu_seconds, args.profile,
                                                args.iterations)
            line["cpu_baseline"]["sample"] = ("16 chains of the same construction from rest, serial oracle, "
                                              "solve phase only; " + line["cpu_baseline"]["sample"].split("; ")[-1])
        else:
            line["cpu_baseline"] = cpu_baseline(scene, start_states, args.cpu_seconds, args.profile,
                                                args.iterations)
    else:
        line["cpu_baseline"] = None
    if args.workload == "pile" and not args.no_warmstart:
        # The metric's second half: sim-steps/s to an RMS residual of 1e-4,
        # learned warm start against plain (the warmstart workload, compact).
        print("# bench: learned warm start vs plain, to RMS 1e-4", file=sys.stderr, flush=True)
        ws = warmstart_bench(args)
        line["warmstart"] = {k: ws[k] for k in ("unit", "plain", "learned", "iteration_ratio", "steps_per_s_ratio")}
        line["warmstart"]["workload"] = ws["config"]["workload"]
        line["warmstart"]["training_samples"] = ws["training"]["samples"]
    print(json.dumps(line), flush=True)


if __name__ == "__main__":
    main()

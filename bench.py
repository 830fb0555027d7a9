"""Benchmark of the physics hot path on one MI355X.

Workload (default): the rubble-pile stress scene, plain step on one scene: 50,000 compound bodies of 1-64 bricks
(geometric p = 0.3, truncated at 64), five materials, packed in a 40 x 20 x 40 m box, v ~ N(0, 3 m/s) per axis,
w ~ N(0, 2 rad/s), FP32, 60 Hz, 12 sub-steps, synthetic seeded input. A step is one World::step of the whole
scene; the metric counts every live brick each step (brick-steps/s).

Prints one JSON line (see the keys below). Progress goes to stderr.
"""
import argparse
import json
import os
import sys
import time

import numpy as np

ROOT = os.path.dirname(os.path.abspath(__file__))
sys.path.insert(0, ROOT)

HBM_PEAK_GBS = 8000.0
# algorithmic bytes per manifold solve in one colour pass (box path): the hot contact record read, impulses written,
# both bodies' v and w read and written
SOLVE_BYTES_PER_MANIFOLD = 392


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(bodies, shape_body, shapes, prm, seconds=15.0, max_steps=20):
    """The oracle (the reference's box step restated in C, reference order, one thread) on the same scene: as many
    steps as fit in about `seconds`, starting from the same state."""
    from oracle import pyoracle
    o = pyoracle.BoxOracle(bodies, shape_body, shapes, prm, pyoracle.ORDER_REFERENCE)
    nbricks = len(shapes)
    t0 = time.perf_counter()
    n = 0
    while n < max_steps:
        o.step(np.float32(1.0 / 60.0))
        n += 1
        if time.perf_counter() - t0 > seconds:
            break
    dt = time.perf_counter() - t0
    return {"value": nbricks * n / dt, "unit": "brick-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle box step (reference order), same scene, first {n} steps from the same state, {dt:.1f} s"}


def main():
    ap = argparse.ArgumentParser()
    ap.add_argument("--gpus", type=int, default=1)
    ap.add_argument("--steps", type=int, default=30)
    ap.add_argument("--warmup", type=int, default=5)
    ap.add_argument("--bodies", type=int, default=50000)
    ap.add_argument("--substeps", type=int, default=12)
    ap.add_argument("--colours", type=int, default=None, help="colour cap (default: the library's)")
    ap.add_argument("--tail-min", type=int, default=None)
    ap.add_argument("--no-cpu", action="store_true")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (scenes shard by scene, see DESIGN.md)",
              file=sys.stderr)
        sys.exit(2)

    import shatterforge_amd as sf
    from shatterforge_amd import scenes
    from shatterforge_amd import world
    from shatterforge_amd.world import params

    if sf.lib().sf_device_count() < 1:
        raise SystemExit("no GPU visible")
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    t = time.perf_counter()
    bodies, shape_body, shapes = scenes.rubble(args.bodies, seed=1)
    nbricks = len(shapes)
    log(f"scene: {len(bodies)} bodies, {nbricks} bricks ({time.perf_counter() - t:.1f} s)")
    prm = params(substeps=args.substeps, sleep=True)
    if args.colours is not None:
        prm[world.P["COL_LIMIT"]] = args.colours
    if args.tail_min is not None:
        prm[world.P["TAIL_MIN"]] = args.tail_min
    w = sf.World.from_scene(bodies, shape_body, shapes, prm)
    for k in range(args.warmup):
        w.step(1.0 / 60.0)
    log(f"warmup {args.warmup} steps done: {w.stats()}")
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        w.step(1.0 / 60.0)
    sync()
    el = time.perf_counter() - t0
    st = w.stats()
    log(f"timed {args.steps} steps: {el * 1e3 / args.steps:.2f} ms/step, stats {st}")
    # the dominant kernel, timed live with HIP events on the library's stream over a second run of the same length
    w.set_kernel_timing(True)
    for k in range(args.steps):
        w.step(1.0 / 60.0)
    ksum_ms, klaunch, kitems = w.kernel_stats()
    w.set_kernel_timing(False)
    achieved = SOLVE_BYTES_PER_MANIFOLD * kitems / (ksum_ms * 1e-3) / 1e9 if ksum_ms > 0 else 0.0
    log(f"solve-bucket kernel: {klaunch} launches, {kitems} manifold solves, avg {ksum_ms / max(klaunch, 1) * 1e3:.1f} us")
    w.set_timing(True)
    w.step(1.0 / 60.0)
    stages = w.stage_ms()
    w.set_timing(False)
    log(f"stages (ms): {stages}")

    value = nbricks * args.steps / el
    out = {
        "metric": "brick-steps/s (batched vs plain GPU step) + p50/p99 what-if query latency",
        "value": value,
        "unit": "brick-steps/s",
        "n_gpus": 1,
        "steps": args.steps,
        "warmup": args.warmup,
        "ms_per_step": el * 1e3 / args.steps,
        "higher_is_better": True,
        "scaling": "weak",
        "vs_baseline": None,
        "dtype": "f32",
        "data": "synthetic (seeded rubble pile)",
        "config": {"workload": "rubble-pile stress, plain step on one scene", "bodies": len(bodies), "bricks": nbricks,
                   "substeps": args.substeps, "dt": "1/60", "mode": "plain",
                   "colour_limit": int(prm[world.P["COL_LIMIT"]]), "tail_min": int(prm[world.P["TAIL_MIN"]])},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": None,
                     "kernel": "k_solve_bucket", "launches": klaunch, "avg_us": ksum_ms / max(klaunch, 1) * 1e3,
                     "bytes_per_unit": SOLVE_BYTES_PER_MANIFOLD, "units": kitems},
        "stages_ms": stages,
        "stats": st,
    }
    if not args.no_cpu:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(bodies, shape_body, shapes, prm, args.cpu_seconds)
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

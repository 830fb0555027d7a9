"""Launch-shape tuners for the box step and the batched brick frame.

The knobs it tunes change how the step's work is spread over the GPU, never what is computed: threads per
workgroup of the colour-bucket solve (one manifold per thread; within a colour no two manifolds share a moving body,
so any grouping gives the same results), the colour-bucket size below which the small colours are solved one after
another in one workgroup instead of one launch each, and whether the sub-steps replay as a captured HIP graph.

Every candidate is timed on one world built from the caller's own scene. The candidates take turns, a few steps each,
on that one trajectory (it does not depend on the knobs), in the opposite order every other round so that a drift of
the step cost along the trajectory does not favour the candidates that come last; each visit times only its last
step, so a graph capture on switching is not counted. The table (per candidate, the median step time)
is returned and can be written as JSON and applied to a world later.

The brick frame's tuner works on the caller's own batched world (its scenes as they are): the contact solve's route
(one workgroup per scene up to a contact count, else colour-by-colour launches over all scenes) and the structure
sweeps' route (potential in LDS from packed rows up to the device's limit, or every structure from the full rows in
global memory). Neither changes the results.
"""
import itertools
import json
import time

import numpy as np

CANDIDATES = {"solve_nt": (64, 128, 256), "tail_min": (256, 1024, 4096), "graph": (True, False)}


def apply(world, best):
    """Set a tuned configuration (a row of the table) on a World."""
    from .world import P
    prm = world.params()
    prm[P["SOLVE_NT"]] = best["solve_nt"]
    prm[P["TAIL_MIN"]] = best["tail_min"]
    world.set_params(prm)
    world.set_graph(bool(best["graph"]))


def tune_box(bodies, shape_body, shapes, prm, rounds=4, visit_steps=3, warmup=2, dt=1.0 / 60.0, path=None):
    """Time every candidate on the scene (bodies, shape_body, shapes) with parameters prm; returns the table
    {"candidates": [...], "best": {...}, ...} and writes it to path as JSON when a path is given."""
    import shatterforge_amd as sf
    from .world import P
    names = list(CANDIDATES)
    combos = [dict(zip(names, v)) for v in itertools.product(*(CANDIDATES[k] for k in names))]
    w = sf.World.from_scene(bodies, shape_body, shapes, prm)
    try:
        for _ in range(warmup):
            w.step(dt)
        times = [[] for _ in combos]
        base = w.params()
        for rnd in range(rounds):
            order = range(len(combos)) if rnd % 2 == 0 else range(len(combos) - 1, -1, -1)
            for k in order:
                c = combos[k]
                p = base.copy()
                p[P["SOLVE_NT"]] = c["solve_nt"]
                p[P["TAIL_MIN"]] = c["tail_min"]
                w.set_params(p)
                w.set_graph(c["graph"])
                for _ in range(visit_steps - 1):
                    w.step(dt)
                t0 = time.perf_counter()
                w.step(dt)  # (a step ends with the stream synchronised)
                times[k].append(time.perf_counter() - t0)
        rows = [dict(c, ms=float(np.median(t) * 1e3)) for c, t in zip(combos, times)]
        default = next(r for r in rows if r["solve_nt"] == 64 and r["tail_min"] == 1024 and r["graph"])
        best = min(rows, key=lambda r: r["ms"])
        table = {"workload": {"bodies": int(len(bodies)), "shapes": int(len(shapes)),
                              "substeps": int(prm[P["SUBSTEPS"]]), "colour_limit": int(prm[P["COL_LIMIT"]])},
                 "rounds": rounds, "timed_steps_per_candidate": rounds, "candidates": rows, "best": best,
                 "default": default}
    finally:
        w.close()
    if path:
        with open(path, "w") as f:
            json.dump(table, f, indent=1)
    return table


def load(path):
    """A table written by tune_box."""
    with open(path) as f:
        return json.load(f)


LATTICE_CANDIDATES = {"fused": (0, 32768, 1 << 30), "sor_lds": (-1, 0)}
LATTICE_DEFAULT = {"fused": 32768, "sor_lds": -1}


def apply_lattice(lat, best):
    """Set a tuned configuration (a row of tune_lattice's table) on a Lattice."""
    lat.set_fused(int(best["fused"]))
    lat.set_sor_lds(int(best["sor_lds"]))


def tune_lattice(lat, rounds=8, dt=1.0 / 60.0, path=None, apply=True):
    """Time every candidate on the Lattice `lat` as it is: each round steps one frame per candidate (in the opposite
    order every other round), so every candidate samples frames spread over the same stretch of the trajectory (a
    frame's cost follows the fracture going on, which one frame per candidate would mistake for the candidate's); the
    median over the rounds decides. The scenes advance rounds x candidates frames, the same frames whatever the
    candidates. Returns the table and leaves the best applied (the default when apply is false). The worker threads
    of the host bookkeeping (set_host_threads) are not tuned: they matter only on frames with fracture work, and a
    change rebuilds the thread pool."""
    names = list(LATTICE_CANDIDATES)
    combos = [dict(zip(names, v)) for v in itertools.product(*(LATTICE_CANDIDATES[k] for k in names))]
    times = [[] for _ in combos]
    for rnd in range(rounds):
        order = range(len(combos)) if rnd % 2 == 0 else range(len(combos) - 1, -1, -1)
        for k in order:
            apply_lattice(lat, combos[k])
            t0 = time.perf_counter()
            lat.step(dt)
            times[k].append(time.perf_counter() - t0)
    rows = [dict(c, ms=float(np.median(t) * 1e3)) for c, t in zip(combos, times)]
    best = min(rows, key=lambda r: r["ms"])
    table = {"rounds": rounds, "frames": rounds * len(combos), "candidates": rows, "best": best,
             "default": dict(LATTICE_DEFAULT)}
    apply_lattice(lat, best if apply else LATTICE_DEFAULT)
    if path:
        with open(path, "w") as f:
            json.dump(table, f, indent=1)
    return table

"""Benchmark of the physics hot path on one MI355X.

Workload (default): the rubble-pile stress scene, plain step on one scene: 50,000 compound bodies of 1-64 bricks
(geometric p = 0.3, truncated at 64), five materials (with their friction and restitution), packed in a 40 x 20 x 40 m
box, v ~ N(0, 3 m/s) per axis, w ~ N(0, 2 rad/s), FP32, 60 Hz, 12 sub-steps, the GPU world's solver (mass-ratio limit
20, six shock-propagation level passes, two supports-first colouring rounds), synthetic seeded input. A step is one
World::step of the whole scene; the metric counts every live brick each step (brick-steps/s).

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
# HBM bytes per k_solve_bucket dispatch on this workload, measured by FETCH_SIZE / WRITE_SIZE counter runs calibrated by
# a copy of known size: the json block of this profile (with the commit and launch shape it was measured at)
TRAFFIC_PROFILE = os.path.join(ROOT, "profiles", "k_solve_bucket_rubble50k_traffic.md")
# algorithmic bytes per manifold solve in one colour pass (box path): the hot contact record read, impulses written,
# both bodies' v and w read and written
SOLVE_BYTES_PER_MANIFOLD = 392


def solve_traffic(solve_nt):
    """the counter-measured HBM bytes per coloured manifold solve of k_solve_bucket, if the profile was taken at this
    launch shape (the profile's json block: bytes per manifold solve by workgroup size)"""
    try:
        text = open(TRAFFIC_PROFILE).read()
        block = text.split("```json", 1)[1].split("```", 1)[0]
        rec = json.loads(block)
    except (OSError, IndexError, ValueError):
        return None, None
    per = rec.get("bytes_per_manifold_solve", {}).get(str(int(solve_nt)))
    return (float(per) if per is not None else None), rec


def log(*a):
    print(*a, file=sys.stderr, flush=True)


def cpu_baseline(bodies, shape_body, shapes, prm, seconds=15.0, max_steps=20, sync=lambda: None):
    """The oracle (the reference's box step restated in C, reference order, one thread) on the same scene: as many
    steps as fit in about `seconds`, starting from the same state; then the GPU on the same steps (a fresh world
    from the same state, the same launch shapes), so the two rates cover the same window."""
    import shatterforge_amd as sf
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
    w = sf.World.from_scene(bodies, shape_body, shapes, prm)
    sync()
    t1 = time.perf_counter()
    for _ in range(n):
        w.step(1.0 / 60.0)
    sync()
    gdt = time.perf_counter() - t1
    w.close()
    gpu = nbricks * n / gdt
    return {"value": nbricks * n / dt, "unit": "brick-steps/s", "cores": 1, "kind": "port",
            "sample": f"oracle box step (reference order), same scene and parameters, steps 1-{n} from the same state, "
                      f"{dt:.1f} s",
            "gpu_same_steps": gpu, "gpu_over_cpu_same_steps": gpu / (nbricks * n / dt)}


def pile_anchor(n=20000, substeps=4, steps=3):
    """the harness check against the reference's published single-thread figure: pile(20000) at 4 sub-steps took
    118 ms per step on a 2-core 2.8 GHz Xeon VM (docs/PHASE0.md:12); the oracle in the reference order, steps 2-4"""
    from oracle import pyoracle
    from shatterforge_amd import scenes
    from shatterforge_amd.world import params
    b, sb, s = scenes.pile(n)
    o = pyoracle.BoxOracle(b, sb, s, params(substeps=substeps, sleep=False), pyoracle.ORDER_REFERENCE)
    o.step(np.float32(1.0 / 60.0))
    t0 = time.perf_counter()
    for _ in range(steps):
        o.step(np.float32(1.0 / 60.0))
    ms = (time.perf_counter() - t0) * 1e3 / steps
    log(f"pile({n}) oracle, {substeps} sub-steps: {ms:.1f} ms/step (reference: 118 ms)")
    return {"bricks": n, "substeps": substeps, "ms_per_step": ms, "reference_ms_per_step": 118.0,
            "reference_machine": "2-core 2.8 GHz Xeon VM, one thread (docs/PHASE0.md:3,12)", "steps_timed": steps}


def tower_cpu(frames, seeds_warm):
    """the C1 tower on the oracle in the reference's sequential order (one thread), the same frames as the GPU's plain
    tower line: the load field by the reference's warm-up, the same frames before the charges, then `frames` timed"""
    from oracle import pyoracle
    from shatterforge_amd import scenes
    mat, xyz, infill = scenes.tower()
    xyz = np.array(xyz, np.int32)
    xyz[:, 0] += 320 - 6
    xyz[:, 2] += 320 - 6
    O = pyoracle.LatticeOracle(order=pyoracle.ORDER_SEQ, brick_cap=len(mat) + 64, body_cap=8192)
    O.add_structure(mat, xyz, infill)
    O.finalize()
    O.warm_all()
    for _ in range(seeds_warm):
        O.step()
    for cx, cz in [(1, 1), (1, 11), (11, 1), (11, 11)]:
        O.explode((cx - 6) * 0.5, 0.6, (cz - 6) * 0.5, 2.2, 1.4)
    t0 = time.perf_counter()
    for _ in range(frames):
        O.step()
    el = time.perf_counter() - t0
    st = O.stats()
    return {"value": len(mat) * frames / el, "unit": "brick-steps/s", "cores": 1, "kind": "port",
            "ms_per_frame": el * 1e3 / frames, "pieces": int(st[2]), "standing": int(st[0]),
            "sample": f"oracle brick frame (sequential order), C1 tower, the {frames} frames after the charges"}


def _lattice_mix(n_scenes, seed=11):
    """what-if-sized scenes: frame buildings of 2k-40k bricks (log-uniform), 1-8 charges at column bases"""
    from shatterforge_amd import serve
    return [serve.make_query(k, seed * 1000 + k) for k in range(n_scenes)]


def _frame_split(L, frames):
    """step `frames` frames, summing the library's own per-frame split (wall ms of the host phases and the waits on
    the device: sf_lat_phase_ms / sf_lat_subphase_ms); returns (elapsed s, mean split per frame)"""
    acc, sub = {}, {}
    t0 = time.perf_counter()
    for _ in range(frames):
        L.step(1.0 / 60.0)
        for k, v in L.phase_ms().items():
            acc[k] = acc.get(k, 0.0) + v
        for k, v in L.subphase_ms().items():
            sub[k] = sub.get(k, 0.0) + v
    el = time.perf_counter() - t0
    split = {k: v / frames for k, v in acc.items()}
    split.update({k: v / frames for k, v in sub.items()})
    return el, split


def bench_lattice(frames=120, batches=(32, 256), cpu=True):
    """legacy brick frame: the plain C1 tower collapse (with the oracle on the same frames); batches of what-if scenes
    in one world; every scene of the batches also stepped in a world of its own (one launch sequence per scene) over
    the same frames, so batched and plain cover the same scenes and the same frame window"""
    from shatterforge_amd import autotune, scenes, serve
    from shatterforge_amd.lattice import Lattice
    from shatterforge_amd.load import LoadBatch
    out = {}
    # C1: 12 x 12 x 40 tower, load field converged by the reference's warm-up sweeps, 4 charges at the column bases
    mat, xyz, infill = scenes.tower()
    xyz = np.array(xyz, np.int32)
    xyz[:, 0] += 320 - 6
    xyz[:, 2] += 320 - 6
    L = Lattice(len(mat) + 64, 8192)
    s = L.add_scene(len(mat) + 64, 8192, structs=2)
    L.add_structure(s, mat, xyz, infill)
    L.finalize(s)
    L.warm_all(s)
    nwarm = 0
    while L.stats(s)[14] > 0 or nwarm < 30:
        L.step(1.0 / 60.0)
        nwarm += 1
        if nwarm > 200:
            break
    for cx, cz in [(1, 1), (1, 11), (11, 1), (11, 11)]:  # the corners of the four 2 x 2 columns
        L.explode(s, (cx - 6) * 0.5, 0.6, (cz - 6) * 0.5, 2.2, 1.4)
    el, split = _frame_split(L, frames)
    st = L.stats(s)
    out["plain_tower"] = {"bricks": int(len(mat)), "frames": frames, "ms_per_frame": el * 1e3 / frames,
                          "brick_steps_per_s": len(mat) * frames / el, "warmup_frames": nwarm,
                          "pieces": int(st[2]), "standing": int(st[0]), "split_ms": split}
    L.close()
    log(f"lattice plain tower: {el * 1e3 / frames:.2f} ms/frame")
    if cpu:
        c = tower_cpu(frames, nwarm)
        out["plain_tower"]["cpu_baseline"] = c
        out["plain_tower"]["gpu_over_cpu"] = out["plain_tower"]["brick_steps_per_s"] / c["value"]
        log(f"lattice tower on the oracle (sequential order, one thread): {c['ms_per_frame']:.2f} ms/frame")
    # batches of what-if scenes in one world vs the same scenes one world each
    n_max = max(batches)
    qs_all = _lattice_mix(n_max)
    def setup(world, slots, qq):
        lb = LoadBatch()
        for q in qq:
            lb.add(q["mat"], q["xyz"], q["infill"])
        lb.solve(rel_tol=1e-6)
        for k, (sl, q) in enumerate(zip(slots, qq)):
            st_ = world.add_structure(sl, q["mat"], q["xyz"], q["infill"])
            world.finalize(sl)
            world.set_field(sl, st_, lb.field(k).astype(np.float32))
            for ch in q["charges"]:
                world.explode(sl, *ch)
        lb.close()
    # plain: each scene in its own world, one untimed frame, then the timed frames
    plain_t = []
    for q in qs_all:
        P = Lattice(q["bricks"] + 64, 4096)
        ps = P.add_scene(q["bricks"] + 64, 4096, structs=2, seed=q["rng_seed"])
        setup(P, [ps], [q])
        P.step(1.0 / 60.0)
        t0 = time.perf_counter()
        for f in range(frames):
            P.step(1.0 / 60.0)
        plain_t.append(time.perf_counter() - t0)
        P.close()
    rows = []
    for n_scenes in batches:
        qs = qs_all[:n_scenes]
        caps = [q["bricks"] + 64 for q in qs]
        def batch():
            W = Lattice(sum(caps), 4096 * n_scenes, scenes=n_scenes)
            sl = [W.add_scene(caps[k], 4096, structs=2, seed=qs[k]["rng_seed"]) for k in range(n_scenes)]
            setup(W, sl, qs)
            W.step(1.0 / 60.0)
            return W
        # the launch-shape tuner on a copy of the batch (the knobs change no result), its best applied to the
        # measured world, which then steps the same frames as the plain worlds
        T = batch()
        tune = autotune.tune_lattice(T, rounds=8, apply=False)
        T.close()
        B = batch()
        autotune.apply_lattice(B, tune["best"])
        el, split = _frame_split(B, frames)
        nbr = sum(q["bricks"] for q in qs)
        plain = nbr * frames / sum(plain_t[:n_scenes])
        rows.append({"scenes": n_scenes, "bricks": nbr, "frames": frames, "ms_per_frame": el * 1e3 / frames,
                     "brick_steps_per_s": nbr * frames / el, "plain_same_scenes_brick_steps_per_s": plain,
                     "batched_over_plain": nbr * frames / el / plain, "split_ms": split,
                     "tuned": {"best": tune["best"], "default_ms": next(
                         (r["ms"] for r in tune["candidates"] if r["fused"] == 32768 and r["sor_lds"] == -1), None),
                               "candidates": len(tune["candidates"]), "frames": tune["frames"]}})
        B.close()
        log(f"lattice batched {n_scenes} scenes ({nbr} bricks): {el * 1e3 / frames:.2f} ms/frame, "
            f"{nbr * frames / el:.3e} brick-steps/s; the same scenes one world each {plain:.3e}")
    best = max(rows, key=lambda r: r["brick_steps_per_s"])
    out["batched"] = dict(best, by_batch=rows,
                          plain_note="every scene of the batch stepped alone in its own world over the same frames "
                                     "(frame 1 untimed, frames 2 to 1 + frames timed)")
    log(f"lattice batched/plain {best['batched_over_plain']:.1f} at {best['scenes']} scenes")
    return out


def bench_whatif(n_queries=64, rate=200.0, frames=300, slots=64):
    """the served what-if stream: Poisson arrivals, continuous batching, latency from arrival to answer"""
    from shatterforge_amd import serve
    qs = [serve.make_query(k, 7000 + k) for k in range(n_queries)]
    arr = serve.poisson_arrivals(rate, n_queries, seed=5)
    srv = serve.WhatIfServer(slots=slots, slot_bricks=max(q["bricks"] for q in qs) + 64, slot_bodies=4096, frames=frames)
    t0 = time.perf_counter()
    done = srv.run_stream(qs, arr)
    el = time.perf_counter() - t0
    lat = np.array([q["latency"] for q in done])
    res = {"queries": len(done), "rate_per_s": rate, "frames_per_query": frames, "slots": slots,
           "p50_ms": float(np.percentile(lat, 50) * 1e3), "p99_ms": float(np.percentile(lat, 99) * 1e3),
           "answered_per_s": len(done) / el, "brick_steps_per_s": srv.brick_steps / el,
           "mean_bricks": float(np.mean([q["bricks"] for q in qs])), "batched_frames": srv.frames_run,
           "server_ms": srv.timing()}
    srv.close()
    log(f"what-if: p50 {res['p50_ms']:.0f} ms p99 {res['p99_ms']:.0f} ms, {res['answered_per_s']:.1f} q/s")
    return res


def bench_load_batch(n_struct=1024, seed=4):
    """load-path precision + batching: n_struct seeded frame buildings of 2k-40k bricks (log-uniform, 1-8 ground
    columns removed), FP32 sweeps with FP64 residual and refinement to max|b - A phi| / max|b| <= 1e-6, all
    structures in one persistent launch"""
    from shatterforge_amd import serve
    from shatterforge_amd.load import LoadBatch
    rng = np.random.default_rng(seed)
    t0 = time.perf_counter()
    L = LoadBatch()
    nbr = 0
    for k in range(n_struct):
        n = int(np.exp(rng.uniform(np.log(2000), np.log(40000))))
        mat, xyz, infill = serve.building_near(n, int(rng.integers(1, 1 << 30)))
        L.add(mat, xyz, infill)
        nbr += len(mat)
    t_build = time.perf_counter() - t0
    t0 = time.perf_counter()
    conv = L.solve(rel_tol=1e-6)
    t_solve = time.perf_counter() - t0
    tm = L.timing()
    ms, sw32, sw64 = tm["ms"], tm["brick_sweeps32"], tm["brick_sweeps64"]
    res = np.array([L.info(k)["resid"] for k in range(n_struct)])
    L.close()
    # algorithmic bytes: 16 per brick per FP32 sweep (potential read and written, source, packed bonds) and 24 per
    # brick per FP64 residual/refinement pass
    alg = 16.0 * sw32 + 24.0 * sw64
    out = {"structures": n_struct, "bricks": nbr, "converged": int(conv), "max_resid64": float(res.max()),
           "device_ms": ms, "wall_solve_s": t_solve, "host_topology_s": t_build,
           "brick_sweeps32": int(sw32), "brick_passes64": int(sw64),
           "achieved_GBs": alg / (ms * 1e-3) / 1e9 if ms > 0 else 0.0}
    out["roofline_frac"] = out["achieved_GBs"] / HBM_PEAK_GBS
    log(f"load batch: {n_struct} structures, {nbr} bricks, {ms:.1f} ms device, {conv} converged, max resid {res.max():.2e}")
    return out


def bench_site(frames=1200, seed=1):
    """site-scale collapse: 37 structures (~380k bricks) in one scene, load fields from the converged solve, eight
    charge sets of four 2.2 m charges at column bases, `frames` frames"""
    from shatterforge_amd import scenes
    from shatterforge_amd.lattice import Lattice
    from shatterforge_amd.load import LoadBatch
    st = scenes.site(seed)
    nbr = sum(len(m) for m, _, _ in st)
    L = Lattice(nbr + 1024, 32768)
    s = L.add_scene(nbr + 1024, 32768, structs=len(st))
    for mat, xyz, infill in st:
        L.add_structure(s, mat, xyz, infill)
    L.finalize(s)
    lb = LoadBatch()
    for mat, xyz, infill in st:
        lb.add(mat, xyz, infill)
    lb.solve(rel_tol=1e-6)
    for k in range(len(st)):
        L.set_field(s, k, lb.field(k).astype(np.float32))
    lb.close()
    for ch in scenes.site_charges(st, seed):
        for c in ch:
            L.explode(s, *c)
    t0 = time.perf_counter()
    for f in range(frames):
        L.step(1.0 / 60.0)
        if f % 200 == 0:
            stt = L.stats(s)
            log(f"site frame {f}: {int(stt[1])} bricks in {int(stt[2])} pieces, {int(stt[4])} contacts")
    el = time.perf_counter() - t0
    stt = L.stats(s)
    ms = L.matstats(s).reshape(10, 5)[1:6]
    out = {"structures": len(st), "bricks": nbr, "frames": frames, "ms_per_frame": el * 1e3 / frames,
           "brick_steps_per_s": nbr * frames / el, "standing": int(stt[0]), "pieces": int(stt[2]),
           "failed_joints": L.failures(s).tolist(), "fragments": ms.tolist()}
    L.close()
    log(f"site: {el * 1e3 / frames:.2f} ms/frame")
    return out


def bench_particles(n=100_000, bodies=5000, steps=90, seed=2):
    """particle debris: n crumbs (r 0.1 m) raining onto a rubble pile of `bodies` compound bodies (GPU world solver),
    `steps` steps; the particle stage's device time per step (HIP events on the world's stream) and the moving
    particle-sub-steps it performed"""
    import shatterforge_amd as sf
    from shatterforge_amd import scenes, world
    from shatterforge_amd.world import params
    b, sb, s = scenes.rubble(bodies, box=(20.0, 10.0, 20.0), seed=seed)
    w = sf.World.from_scene(b, sb, s, params(**world.GPU_WORLD))
    w.particles_capacity(n)
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0] = rng.uniform(-10.0, 10.0, n)
    rec[:, 1] = rng.uniform(12.0, 20.0, n)
    rec[:, 2] = rng.uniform(-10.0, 10.0, n)
    rec[:, 6] = 0.1
    rec[:, 7] = rng.choice([0.5, 0.6], n)
    w.add_particles(rec)
    w.set_timing(True)
    pms, moving, t0 = 0.0, 0, time.perf_counter()
    for _ in range(steps):
        w.step(1.0 / 60.0)
        pms += w.particle_ms()
        moving += int(w.particle_stats()[1])
    el = time.perf_counter() - t0
    st = w.particle_stats()
    w.close()
    psub = int(params()[world.P["PSUB"]])
    res = {"particles": n, "bodies": bodies, "steps": steps, "ms_per_step": el * 1e3 / steps,
           "particle_stage_ms_per_step": pms / steps, "moving_per_step": moving / steps, "asleep_at_end": int(st[2]),
           "moving_particle_substeps_per_s": moving * psub / (pms * 1e-3) if pms > 0 else 0.0}
    log(f"particles: {n} on {bodies} bodies, stage {pms / steps:.3f} ms/step, {moving / steps:.0f} moving, "
        f"whole step {el * 1e3 / steps:.2f} ms")
    return res


def bench_chunks(n=20000, steps=30, warm=30, cpu=True, seed=5):
    """convex chunks: n bodies, each one cell of the chunk library (random pattern and cell, random orientation),
    dropped in a loose grid onto the ground and stepped with the GPU world's solver and contact reduction; the GPU's
    ms per step over steps warm+1 .. warm+steps, and the oracle (coloured order, one thread) on the same scene and
    the same steps (at most 20 s of them)"""
    import math
    import shatterforge_amd as sf
    from shatterforge_amd import hulls, scenes, world
    from shatterforge_amd.world import params
    rng = np.random.default_rng(seed)
    side = int(math.ceil((n / 6) ** 0.5))
    items = []
    for k in range(n):
        pat = int(rng.integers(0, hulls.pattern_count()))
        ids = hulls.pattern(pat)
        h = ids[int(rng.integers(0, len(ids)))]
        hh = hulls.cell(h)["box_h"] * 0.25 + 0.04
        gx, gz, gy = k % side, (k // side) % side, k // (side * side)
        p = ((gx - side / 2) * 0.9, 0.5 + gy * 0.9, (gz - side / 2) * 0.9)
        ax = rng.random(3) - 0.5
        ax = ax / np.linalg.norm(ax)
        ang = rng.random() * 6.0
        q = np.array([ax[0] * math.sin(ang / 2), ax[1] * math.sin(ang / 2), ax[2] * math.sin(ang / 2), math.cos(ang / 2)], np.float32)
        items.append(scenes.body([scenes.shape(p, hh, q=q, mat=1 + k % 5, hull=h)], v=(0, -1, 0)))
    b, sb, s = scenes.bodies(items)
    prm = params(**world.GPU_WORLD)
    w = sf.World.from_scene(b, sb, s, prm)
    for _ in range(warm):
        w.step(1.0 / 60.0)
    w.stats()   # (a read-back: the warm-up has finished)
    t0 = time.perf_counter()
    for _ in range(steps):
        w.step(1.0 / 60.0)
    st = w.stats()
    ms = (time.perf_counter() - t0) * 1e3 / steps
    w.close()
    res = {"bodies": n, "steps": steps, "warmup": warm, "ms_per_step": ms, "chunk_steps_per_s": n / (ms * 1e-3),
           "manifolds": st["manifolds"], "reduced": st["reduced"], "awake": st["awake"]}
    log(f"chunks: {n} bodies, steps {warm + 1}-{warm + steps}: {ms:.3f} ms/step, {st['manifolds']} manifolds")
    if cpu:
        from oracle import pyoracle
        o = pyoracle.BoxOracle(b, sb, s, prm, pyoracle.ORDER_COLOURED, hulls=hulls.tables())
        for _ in range(warm):
            o.step(1.0 / 60.0)
        k, t1 = 0, time.perf_counter()
        while k < steps and time.perf_counter() - t1 < 20.0:
            o.step(1.0 / 60.0)
            k += 1
        cms = (time.perf_counter() - t1) * 1e3 / k
        res["cpu_baseline"] = {"kind": "port", "cores": 1, "ms_per_step": cms, "steps": k,
                               "sample": f"oracle box step with chunks (coloured order), the same scene, steps {warm + 1}-{warm + k}",
                               "gpu_over_cpu": cms / ms}
        log(f"chunks cpu: {cms:.1f} ms/step ({cms / ms:.1f}x)")
    return res


def _bricks_of(bricks):
    from shatterforge_amd import structure
    b = structure.Bricks()
    for mn, size, m, paint, fl in bricks:
        i = b.add(mn, size, m, paint)
        if fl != 1:
            b.set_flags(i, fl)
    b.build_joints()
    return b


def bench_structure(cpu=True, batch=256, seed=2, reps=3):
    """structural analysis (the spring-network solve): a 12-storey frame tower of 3 x 3 bays analysed alone, then a
    batch of `batch` seeded frame buildings (4-16 storeys, 2-4 bays, some ground columns left out) in one launch, a
    workgroup per building; the oracle (reference summation order, one thread) on the tower and on a sample of the
    batch"""
    from shatterforge_amd import scenes, structure
    tower = _bricks_of(scenes.frame_tower(12, 3))
    case = structure.gravity_case(tower)
    net = structure.Net()
    walls, splits = [], []
    for _ in range(reps):
        net.clear()
        net.add(tower, *case)
        t0 = time.perf_counter()
        net.solve()
        walls.append((time.perf_counter() - t0) * 1e3)
        splits.append(net.ms().tolist())
    k = int(np.argsort(walls)[len(walls) // 2])
    r = net.result(0)
    res = {"tower": {"storeys": 12, "bays": 3, "bricks": len(tower), "cells": r.nodes, "links": r.links,
                     "levels": r.levels, "iterations": r.iterations, "max_ratio": float(r.max_ratio),
                     "ms": walls[k], "host_build_ms": splits[k][0], "upload_ms": splits[k][1],
                     "device_solve_ms": splits[k][2], "device_checks_ms": splits[k][3]}}
    log(f"structure: tower {r.nodes} cells, {r.iterations} iterations, {walls[k]:.1f} ms (device {splits[k][2]:.1f})")
    towers = scenes.frame_towers(batch, seed=seed)
    stores = [_bricks_of(t) for t, _ in towers]
    cases = [(b, *structure.gravity_case(b)) for b in stores]
    walls, splits = [], []
    for _ in range(2):
        net.clear()
        for c in cases:
            net.add(*c)
        t0 = time.perf_counter()
        net.solve()
        walls.append((time.perf_counter() - t0) * 1e3)
        splits.append(net.ms().tolist())
    k = int(np.argmin(walls))
    cells = sum(net.result(i).nodes for i in range(batch))
    its = [net.result(i).iterations for i in range(batch)]
    res["batch"] = {"buildings": batch, "cells": cells, "iterations_mean": float(np.mean(its)),
                    "iterations_max": int(max(its)), "ms": walls[k], "analyses_per_s": batch / (walls[k] * 1e-3),
                    "host_build_ms": splits[k][0], "upload_ms": splits[k][1], "device_solve_ms": splits[k][2],
                    "device_checks_ms": splits[k][3]}
    log(f"structure: batch of {batch}, {cells} cells, {walls[k]:.0f} ms ({batch / (walls[k] * 1e-3):.0f} analyses/s, "
        f"device {splits[k][2]:.0f} ms)")
    if cpu:
        from oracle import pyoracle as po  # the checker: timed as the CPU baseline only

        def oracle_case(bricks):
            o = po.BricksOracle()
            for mn, size, m, paint, fl in bricks:
                i = o.add(mn, size, m, paint)
                if fl != 1:
                    o.set_flags(i, fl)
            o.build_joints()
            n = len(bricks)
            load = np.array([o.mass(i) * np.float32(9.81) for i in range(n)], np.float32)
            sink = np.array([bk[0][1] <= 0 for bk in bricks], np.uint8)
            return o, np.arange(n, dtype=np.uint32), load, sink

        o, set_, load, sink = oracle_case(scenes.frame_tower(12, 3))
        t0 = time.perf_counter()
        po.analyse_oracle(o, set_, load, sink, order=po.ORDER_REFERENCE)
        tower_cpu = (time.perf_counter() - t0) * 1e3
        sample = 8
        t0 = time.perf_counter()
        for t, _ in towers[:sample]:
            po.analyse_oracle(*oracle_case(t), order=po.ORDER_REFERENCE)
        per = (time.perf_counter() - t0) * 1e3 / sample
        res["cpu_baseline"] = {"kind": "port", "cores": 1, "tower_ms": tower_cpu, "batch_ms_per_analysis": per,
                               "batch_analyses_per_s": 1e3 / per, "sample": f"the tower, and the first {sample} "
                               "buildings of the batch (with their store built)",
                               "tower_speedup": tower_cpu / res["tower"]["ms"],
                               "batch_speedup": res["batch"]["analyses_per_s"] / (1e3 / per)}
        log(f"structure cpu: tower {tower_cpu:.0f} ms, {per:.0f} ms per building of the batch")
    return res


def bench_box(args, sync):
    """the plain box step on the rubble pile: the JSON line's value, roofline line and stages"""
    import shatterforge_amd as sf
    from shatterforge_amd import scenes
    from shatterforge_amd import world
    from shatterforge_amd.world import params
    t = time.perf_counter()
    bodies, shape_body, shapes = scenes.rubble(args.bodies, seed=1)
    nbricks = len(shapes)
    log(f"scene: {len(bodies)} bodies, {nbricks} bricks ({time.perf_counter() - t:.1f} s)")
    prm = params(substeps=args.substeps, sleep=True, **({} if args.phys else world.GPU_WORLD))
    if args.colours is not None:
        prm[world.P["COL_LIMIT"]] = args.colours
    if args.tail_min is not None:
        prm[world.P["TAIL_MIN"]] = args.tail_min
    tuned = None
    if args.solve_nt is not None:
        prm[world.P["SOLVE_NT"]] = args.solve_nt
    if not args.no_autotune and args.solve_nt is None:
        # launch shapes (solve workgroup size, small-colour threshold, graph replay) timed on this scene; none of
        # them changes the results
        from shatterforge_amd import autotune
        t = time.perf_counter()
        table = autotune.tune_box(bodies, shape_body, shapes, prm)
        tuned = {"best": table["best"], "default": table["default"], "candidates": len(table["candidates"]),
                 "seconds": time.perf_counter() - t}
        log(f"autotune: best {table['best']} (default {table['default']['ms']:.2f} ms) in {tuned['seconds']:.1f} s")
        prm[world.P["SOLVE_NT"]] = table["best"]["solve_nt"]
        prm[world.P["TAIL_MIN"]] = table["best"]["tail_min"]
    w = sf.World.from_scene(bodies, shape_body, shapes, prm)
    if tuned is not None:
        w.set_graph(bool(tuned["best"]["graph"]))
    # coloured manifold solves of every step of this world (both passes of every sub-step): with the counter
    # profile's per-dispatch bytes they give the bytes per manifold solve
    coloured = [0]
    def counted():
        st_ = w.stats()
        coloured[0] += (st_["manifolds"] - st_["overflow"]) * 2 * args.substeps
    for k in range(args.warmup):
        w.step(1.0 / 60.0)
        counted()
    log(f"warmup {args.warmup} steps done: {w.stats()}")
    sync()
    t0 = time.perf_counter()
    for k in range(args.steps):
        w.step(1.0 / 60.0)
        counted()  # (a host read of the step's counters: the step has already synchronised)
    sync()
    el = time.perf_counter() - t0
    st = w.stats()
    graph = dict(zip(("captures", "replays"), w.graph_stats()))
    log(f"timed {args.steps} steps: {el * 1e3 / args.steps:.2f} ms/step, stats {st}, sub-step graph {graph}")
    # the dominant kernel, timed live with HIP events on the library's stream over a second run of the same length
    w.set_kernel_timing(True)
    for k in range(args.steps):
        w.step(1.0 / 60.0)
        counted()
    ksum_ms, klaunch, kitems = w.kernel_stats()
    w.set_kernel_timing(False)
    achieved = SOLVE_BYTES_PER_MANIFOLD * kitems / (ksum_ms * 1e-3) / 1e9 if ksum_ms > 0 else 0.0
    log(f"solve-bucket kernel: {klaunch} launches, {kitems} manifold solves, avg {ksum_ms / max(klaunch, 1) * 1e3:.1f} us")
    w.set_timing(True)
    w.step(1.0 / 60.0)
    counted()
    stages = w.stage_ms()
    w.set_timing(False)
    log(f"stages (ms): {stages}")
    log(f"all {args.warmup + 2 * args.steps + 1} steps of this world: {coloured[0]} coloured manifold solves")

    value = nbricks * args.steps / el
    per_solve, trec = solve_traffic(int(prm[world.P["SOLVE_NT"]]))
    traffic = per_solve * kitems / max(klaunch, 1) if per_solve is not None else None
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
                   "colour_limit": int(prm[world.P["COL_LIMIT"]]), "tail_min": int(prm[world.P["TAIL_MIN"]]),
                   "solve_workgroup": int(prm[world.P["SOLVE_NT"]]),
                   "solver": "reference CPU model (phys.rs)" if args.phys else
                             "GPU world (mass-ratio limit 20, 6 level passes, 2 supports-first rounds)"},
        "roofline": {"bound": "hbm", "achieved": achieved, "peak": HBM_PEAK_GBS, "unit": "GB/s",
                     "frac": achieved / HBM_PEAK_GBS, "traffic": traffic,
                     "traffic_source": None if per_solve is None else
                                       f"profiles/k_solve_bucket_rubble50k_traffic.md (commit {trec.get('commit')}): "
                                       f"{per_solve:.0f} HBM bytes per manifold solve x this run's solves per launch",
                     "kernel": "k_solve_bucket", "launches": klaunch, "avg_us": ksum_ms / max(klaunch, 1) * 1e3,
                     "bytes_per_unit": SOLVE_BYTES_PER_MANIFOLD, "units": kitems},
        "stages_ms": stages,
        "stats": st,
        "substep_graph": graph,
        "autotune": tuned,
    }
    return out, (bodies, shape_body, shapes, prm)


def bench_copy(gib=1.0, reps=5):
    """a device-to-device copy of known size (the bench-only program shatterforge_amd/sfb_copy, built apart from the
    library: 16-byte loads and stores between two 1 GiB buffers), run as a child process: the reference reading for the
    HBM byte counters (FETCH_SIZE and WRITE_SIZE per byte moved), which is profiled in a run of its own"""
    import subprocess
    exe = os.path.join(ROOT, "shatterforge_amd", "sfb_copy")
    r = subprocess.run([exe, str(gib), str(reps)], capture_output=True, text=True, check=True)
    res = json.loads(r.stdout.strip().splitlines()[-1])
    log(f"copy: {res['GBs_read_plus_write']:.0f} GB/s (read + write)")
    return res


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
    ap.add_argument("--no-autotune", action="store_true", help="keep the default launch shapes for the box step")
    ap.add_argument("--solve-nt", type=int, default=None, help="box solve workgroup size (skips the autotuner)")
    ap.add_argument("--phys", action="store_true", help="the box step with the reference CPU model's solver (the GPU "
                    "world's extensions off)")
    ap.add_argument("--cpu-seconds", type=float, default=15.0)
    ap.add_argument("--no-lattice", action="store_true", help="skip the legacy brick frame and what-if measurements")
    ap.add_argument("--parts", default="box,lattice,whatif,load,site,particles,structure,chunks",
                    help="comma list of box, lattice, whatif, load, site, particles, structure, chunks, copy (a device copy of "
                         "known size by the bench-only program shatterforge_amd/sfb_copy)")
    ap.add_argument("--whatif-queries", type=int, default=64)
    ap.add_argument("--load-batch", type=int, default=0, help="also run the load-path batch of this many structures")
    ap.add_argument("--site", type=int, default=0, help="also run the site-scale collapse for this many frames")
    args = ap.parse_args()
    if args.gpus != 1:
        print(f"bench.py measures one GPU; --gpus {args.gpus} is not supported (the scenes would shard whole across "
              f"GPUs; nothing here runs on more than one)",
              file=sys.stderr)
        sys.exit(2)

    import shatterforge_amd as sf

    if sf.lib().sf_device_count() < 1:
        raise SystemExit("no GPU visible")
    try:
        import torch
        sync = torch.cuda.synchronize if torch.cuda.is_available() else (lambda: None)
    except Exception:
        sync = lambda: None  # noqa: E731

    parts = set(args.parts.split(","))
    if args.no_lattice:
        parts -= {"lattice", "whatif"}
    out = {"metric": "brick-steps/s (batched vs plain GPU step) + p50/p99 what-if query latency", "value": None,
           "unit": "brick-steps/s", "n_gpus": 1, "steps": args.steps, "warmup": args.warmup, "ms_per_step": None,
           "higher_is_better": True, "scaling": "weak", "vs_baseline": None, "dtype": "f32",
           "data": "synthetic (seeded scenes)", "config": {}}
    scene = None
    if "box" in parts:
        out, scene = bench_box(args, sync)
    if "copy" in parts:
        out["copy"] = bench_copy()
    if "lattice" in parts:
        out["lattice"] = bench_lattice(cpu=not args.no_cpu)
    if "whatif" in parts:
        out["whatif"] = bench_whatif(n_queries=args.whatif_queries)
    if "load" in parts and not args.load_batch:
        args.load_batch = 1024
    if "site" in parts and not args.site:
        args.site = 1200
    if args.load_batch:
        out["load_batch"] = bench_load_batch(args.load_batch)
    if args.site:
        out["site"] = bench_site(args.site)
    if "particles" in parts:
        out["particles"] = bench_particles()
    if "structure" in parts:
        out["structure"] = bench_structure(cpu=not args.no_cpu)
    if "chunks" in parts:
        out["chunks"] = bench_chunks(cpu=not args.no_cpu)
    if not args.no_cpu and scene is not None:
        log("cpu baseline ...")
        out["cpu_baseline"] = cpu_baseline(*scene, args.cpu_seconds, sync=sync)
        out["cpu_baseline"]["pile_anchor"] = pile_anchor()
    print(json.dumps(out), flush=True)


if __name__ == "__main__":
    main()

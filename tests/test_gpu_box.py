"""Box step on the GPU against the oracle (oracle/box_world.c in the device's colour order) through the C ABI.

The bar: candidate lists, manifold keys, colours and island labels equal exactly; contact data and body states equal
bit for bit (the kernels perform the oracle's f32 operations in the same order, with no FMA contraction).
"""
import numpy as np
import pytest

import shatterforge_amd as sf
from shatterforge_amd import scenes, world
from shatterforge_amd.world import params

pytestmark = pytest.mark.gpu
DT = np.float32(1.0 / 60.0)


def _assert_same_step(w, o, tag):
    assert np.array_equal(w.candidates(), o.candidates()), f"{tag}: candidates"
    gk, gd = w.manifolds()
    ok, od = o.manifolds()
    assert np.array_equal(gk, ok), f"{tag}: manifold keys"
    assert np.array_equal(gd, od), f"{tag}: manifold data max diff {np.abs(gd - od).max() if len(gd) else 0}"
    assert np.array_equal(w.colours(), o.colours()), f"{tag}: colours"
    assert np.array_equal(w.islands(), o.islands()), f"{tag}: islands"
    gb, ob = w.bodies(), o.bodies()
    assert np.array_equal(gb, ob), f"{tag}: bodies max diff {np.abs(gb - ob).max()}"


def test_pile_steps_bit_exact(oracle):
    print("case: pile(600) substeps 4, 8 steps", flush=True)
    b, sb, s = scenes.pile(600)
    prm = params(substeps=4, sleep=False)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for k in range(8):
        w.step(DT)
        o.step(DT)
        _assert_same_step(w, o, f"step {k}")
    st, ost = w.stats(), o.stats()
    assert st["manifolds"] == ost["manifolds"] and st["points"] == ost["points"] and st["colours"] == ost["colours"]


def test_rubble_compound_steps_bit_exact_with_sleep(oracle):
    print("case: rubble(300) compound bodies, 60 steps, sleep on", flush=True)
    b, sb, s = scenes.rubble(300, box=(7.0, 5.0, 7.0), seed=5)
    prm = params(substeps=6)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for k in range(60):
        w.step(DT)
        o.step(DT)
        gb, ob = w.bodies(), o.bodies()
        assert np.array_equal(gb, ob), f"step {k}: max diff {np.abs(gb - ob).max()}"
    _assert_same_step(w, o, "final")


@pytest.mark.parametrize("limit", [2, 5])
def test_overflow_jacobi_bit_exact(oracle, limit):
    # a tiny colour cap sends most contacts to the mass-splitting overflow pass; device == oracle bit for bit
    print(f"case: rubble(400) colour limit {limit}", flush=True)
    b, sb, s = scenes.rubble(400, box=(8.0, 6.0, 8.0), seed=21)
    prm = params(substeps=4, col_limit=limit, tail_min=64)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for k in range(15):
        w.step(DT)
        o.step(DT)
        gb, ob = w.bodies(), o.bodies()
        assert np.array_equal(gb, ob), f"step {k}: max diff {np.abs(gb - ob).max()}"
    st = w.stats()
    assert st["overflow"] > 0 and st["overflow"] == o.stats()["overflow"]
    _assert_same_step(w, o, "final")


def test_first_frame_candidates_of_the_reference_pile():
    # docs/GPU.md:89: 5,871 candidates on the first frame of the 2,000-box pile
    b, sb, s = scenes.pile(2000)
    w = sf.World.from_scene(b, sb, s, params(substeps=4, sleep=False))
    w.step(DT)
    st = w.stats()
    assert st["pairs"] + st["ground"] == 5871 == len(w.candidates())


def test_stack_gate_and_reference_exports():
    # engine/src/phys.rs:672-683 through the reference's own exports (engine/src/wasm.rs:33-47)
    world.stack_setup(20, 12, 0)
    p0 = world.pose()
    top0 = p0[191]
    maxv = 0.0
    for f in range(600):
        world.step(1.0 / 60.0)
        if f >= 540:
            maxv = max(maxv, world.max_speed())
    p = world.pose()
    drift = float(np.hypot(np.hypot(p[190], p[191] - top0), p[192]))
    assert drift < 0.01 and maxv < 0.01, (drift, maxv)
    assert world.body_count() == 20 and world.manifolds() >= 20
    assert world.body_pos(19, 1) == p[191]


def test_pile_two_seconds_within_reference_tolerance():
    # gpu/test_gpu.cjs:1049-1061 against the CPU reference heights of docs/GPU.md:158 (y 0.117-4.828 m)
    world.pile_setup(2000, 4)
    for _ in range(120):
        world.step(1.0 / 60.0)
    y = world.pose().reshape(-1, 10)[:, 1]
    assert np.isfinite(y).all()
    assert abs(y.min() - 0.117) < 0.02 and abs(y.max() - 4.828) < 0.1, (y.min(), y.max())


def test_determinism_run_twice():
    b, sb, s = scenes.rubble(2000, box=(12.0, 8.0, 12.0), seed=9)
    out = []
    for _ in range(2):
        w = sf.World.from_scene(b, sb, s, params())
        for _ in range(20):
            w.step(DT)
        out.append((w.bodies(), w.colours(), w.islands()))
        w.close()
    for a, c in zip(out[0], out[1]):
        assert np.array_equal(a, c)


def test_graph_replay_same_as_launches():
    # the sub-steps replayed as a captured graph give the same bodies as launching them one by one, through the
    # steps where the colour counts change and the graph is captured again
    print("case: rubble(2000), 40 steps, graph on vs off", flush=True)
    b, sb, s = scenes.rubble(2000, box=(14.0, 8.0, 14.0), seed=11)
    prm = params(substeps=4)
    wg = sf.World.from_scene(b, sb, s, prm)
    wd = sf.World.from_scene(b, sb, s, prm)
    wd.set_graph(False)
    for k in range(40):
        wg.step(DT)
        wd.step(DT)
        gb, db = wg.bodies(), wd.bodies()
        assert np.array_equal(gb, db), f"step {k}: max diff {np.abs(gb - db).max()}"
    assert np.array_equal(wg.colours(), wd.colours())
    caps, reps = wg.graph_stats()
    print(f"graph captures {caps}, replays {reps}", flush=True)
    assert caps >= 1 and reps >= 1 and caps + reps == 40
    assert wd.graph_stats() == (0, 0)
    wg.close()
    wd.close()


def test_autotune_table_and_tuned_world_same_results(tmp_path):
    # every launch shape the tuner may pick computes the same bodies as the defaults
    from shatterforge_amd import autotune
    print("case: rubble(1500) tuner, 1 round, then tuned vs default 12 steps", flush=True)
    b, sb, s = scenes.rubble(1500, box=(12.0, 8.0, 12.0), seed=13)
    prm = params(substeps=4)
    table = autotune.tune_box(b, sb, s, prm, rounds=1, path=str(tmp_path / "tune.json"))
    assert len(table["candidates"]) == 18 and all(r["ms"] > 0 for r in table["candidates"])
    assert autotune.load(str(tmp_path / "tune.json"))["best"] == table["best"]
    for cfg in (table["best"], dict(solve_nt=256, tail_min=4096, graph=False), dict(solve_nt=128, tail_min=256, graph=True)):
        wt = sf.World.from_scene(b, sb, s, prm)
        autotune.apply(wt, cfg)
        wd = sf.World.from_scene(b, sb, s, prm)
        for k in range(12):
            wt.step(DT)
            wd.step(DT)
        assert np.array_equal(wt.bodies(), wd.bodies()), f"{cfg}: bodies differ"
        assert np.array_equal(wt.colours(), wd.colours())
        wt.close()
        wd.close()


def test_full_size_rubble_determinism_and_complete_candidates():
    # the full rubble pile, two worlds stepped side by side: fast bodies' swept boxes can have more partners than a
    # shape's candidate slots; the list must still be complete and sorted, whatever order the cell table was filled
    # in, so both worlds agree step after step
    print("case: rubble 50,000 bodies, two worlds, 4 steps", flush=True)
    b, sb, s = scenes.rubble(50000, seed=1)
    A = sf.World.from_scene(b, sb, s, params(substeps=12, sleep=True))
    B = sf.World.from_scene(b, sb, s, params(substeps=12, sleep=True))
    for k in range(4):
        A.step(DT)
        B.step(DT)
        ca, cb = A.candidates(), B.candidates()
        assert np.array_equal(ca, cb), f"step {k}: candidate lists differ"
        assert np.array_equal(A.bodies(), B.bodies(), equal_nan=True), f"step {k}: bodies differ"
        assert A.stats()["dropped"] == 0
        # per shape, partners ascending and unique
        a0 = ca[:, 0].astype(np.int64) * (1 << 32) + (ca[:, 1].astype(np.int64) + 1)
        assert np.all(np.diff(a0) > 0), f"step {k}: candidates not in canonical order"
    A.close()
    B.close()


def test_full_size_rubble_step_matches_oracle_and_invariants(oracle):
    # the rubble workload at full size (50,000 polycube bodies): one step bit for bit, then invariants
    print("case: rubble 50k bodies, 1 step vs oracle", flush=True)
    b, sb, s = scenes.rubble(50000, seed=1)
    prm = params()
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    w.step(DT)
    o.step(DT)
    _assert_same_step(w, o, "rubble 50k step 0")
    for _ in range(4):
        pre = w.bodies()
        w.step(DT)
        keys, _ = w.manifolds()
        col = w.colours()
        K = int(prm[world.P["COL_LIMIT"]])
        # canonical candidates: sorted, unique
        c = w.candidates()
        kc = (c[:, 0].astype(np.int64) << 32) + (c[:, 1].astype(np.int64) + 1)
        assert (np.diff(kc) > 0).all()
        # colour validity on moving bodies
        awake0 = pre[:, 17] > 0.5
        ba = sb[keys[:, 0]]
        bb = np.where(keys[:, 1] >= 0, sb[np.maximum(keys[:, 1], 0)], -1)
        for side in (ba, bb):
            sel = (col < K) & (side >= 0)
            sel &= awake0[np.maximum(side, 0)]
        pairs = np.concatenate([np.stack([ba, col], 1)[(col < K) & awake0[ba]],
                                np.stack([bb, col], 1)[(col < K) & (bb >= 0) & awake0[np.maximum(bb, 0)]]])
        assert len(np.unique(pairs, axis=0)) == len(pairs)
        # island labels are roots: label[label[i]] == label[i] <= i
        isl = w.islands()
        assert (isl[isl] == isl).all() and (isl <= np.arange(len(isl))).all()
        assert np.isfinite(w.bodies()).all()
    assert w.stats()["dropped"] == 0


def _sleep_scene():
    # three 6-box stacks: A at rest, B with a top box sliding (its island stays restless until it stops), C hit by a
    # box sliding in along the ground after the others fell asleep
    items = []
    for x in (-4.0, 0.0, 4.0):
        for k in range(6):
            v = (1.0, 0.0, 0.0) if (x == 0.0 and k == 5) else (0, 0, 0)
            items.append(scenes.box((x, 0.25 + 0.5 * k, 0.0), (0.25, 0.25, 0.25), v=v))
    items.append(scenes.box((6.8, 0.25, 0.0), (0.25, 0.25, 0.25), v=(-6.0, 0.0, 0.0)))
    return scenes.boxes(items)


def test_island_sleep_bit_exact(oracle):
    # islands fall asleep one at a time; a sleepy body in an island with a restless member must stay awake
    print("case: three stacks with sleep on, 150 steps", flush=True)
    b, sb, s = _sleep_scene()
    prm = params(substeps=4)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    awake = []
    for k in range(150):
        w.step(DT)
        o.step(DT)
        gb, ob = w.bodies(), o.bodies()
        assert np.array_equal(gb, ob), f"step {k}: max diff {np.abs(gb - ob).max()}"
        awake.append(int(gb[:, 17].sum()))
        if k % 10 == 0:
            assert np.array_equal(w.islands(), o.islands()), f"step {k}: islands"
    n = len(b)
    assert max(awake) == n and min(awake) < n, awake
    assert len(set(awake)) >= 3, awake


# ---------------------------------------------------------------- the GPU world's solver extensions
def _ext_case(name):
    """Small scenes that drive each extension: the reference's landing (shock propagation, ratio limit, restitution),
    a bouncy drop on a static floor (restitution of the pair), a head-on meeting at 100:1 (ratio limit, no gravity)."""
    if name == "landing":
        hl, hh, v = 0.135, 0.5, 7.7
        items = [scenes.box((0, hl + 0.05, 0), (hl, hl, hl), density=2400.0, e=0.1, v=(0, -v, 0)),
                 scenes.box((0, 0.05 + 2 * hl + hh * 0.4 + 0.0005, 0), (hh, hh * 0.4, hh), density=2400.0, e=0.1, v=(0, -v, 0))]
        return items, params(sleep=False, **world.GPU_WORLD), 40
    if name == "bounce":
        items = [scenes.box((0, 2.25, 0), (0.25, 0.25, 0.25), e=0.3),
                 scenes.box((0, -0.5, 0), (3.0, 0.5, 3.0), density=0.0, e=0.5)]
        return items, params(sleep=False, **world.GPU_WORLD), 100
    items = [scenes.box((0, 10, 0), (0.5, 0.5, 0.5), density=12500.0, v=(2.0, 0, 0)),
             scenes.box((0.755, 10, 0), (0.25, 0.25, 0.25), density=1000.0)]
    return items, params(sleep=False, ratio=20.0, gy=0.0), 10


@pytest.mark.parametrize("name", ["landing", "bounce", "headon"])
def test_gpu_world_extensions_bit_exact_small(oracle, name):
    print(f"case: {name} with the GPU world's extensions", flush=True)
    items, prm, frames = _ext_case(name)
    b, sb, s = scenes.boxes(items)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for k in range(frames):
        w.step(DT)
        o.step(DT)
        _assert_same_step(w, o, f"{name} step {k}")


@pytest.mark.parametrize("limit", [3, 16])
def test_gpu_world_extensions_rubble_bit_exact(oracle, limit):
    # mixed materials (restitution 0.05-0.3, densities 500-7850, 1-64 bricks: mass ratios past 20), supports-first
    # colouring, level passes; a small colour cap puts restitution and the ratio limit through the split overflow pass
    print(f"case: rubble(500) GPU world extensions, colour limit {limit}, 30 steps", flush=True)
    b, sb, s = scenes.rubble(500, box=(9.0, 6.0, 9.0), seed=17)
    prm = params(substeps=4, col_limit=limit, tail_min=64, **world.GPU_WORLD)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for k in range(30):
        w.step(DT)
        o.step(DT)
        gb, ob = w.bodies(), o.bodies()
        assert np.array_equal(gb, ob), f"step {k}: max diff {np.abs(gb - ob).max()}"
    _assert_same_step(w, o, "final")
    assert w.stats()["overflow"] == o.stats()["overflow"]


def test_full_size_rubble_step_with_the_gpu_world_extensions(oracle):
    print("case: rubble 50k bodies with the GPU world's extensions, 1 step vs oracle", flush=True)
    b, sb, s = scenes.rubble(50000, seed=1)
    prm = params(**world.GPU_WORLD)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    w.step(DT)
    o.step(DT)
    _assert_same_step(w, o, "rubble 50k extensions step 0")


def test_full_size_rubble_300_steps_sane(oracle):
    # the rubble pile (50,000 bodies) over its 300 steps with the GPU world's solver: the 20 m pile collapses under
    # gravity, so kinetic energy grows, but kinetic plus potential energy never rises (no energy from nowhere: the
    # solve, restitution and the mass-ratio limit only take energy out), nothing is NaN, no candidate or manifold is
    # dropped, and no body sinks through the ground
    print("case: rubble 50k bodies, 300 steps, energy bound and invariants", flush=True)
    b, sb, s = scenes.rubble(50000, seed=1)
    w = sf.World.from_scene(b, sb, s, params(**world.GPU_WORLD))
    m = np.where(b[:, 13] > 0, 1.0 / np.maximum(b[:, 13].astype(np.float64), 1e-30), 0.0)

    def energy(B):
        v2 = (B[:, 7:10].astype(np.float64) ** 2).sum(1)
        w3 = B[:, 10:13].astype(np.float64)
        # rotational energy in the body frame: R^T w, principal inverse inertia B[:, 14:17]
        q = B[:, 3:7].astype(np.float64)
        x, y, z, ww = q[:, 0], q[:, 1], q[:, 2], q[:, 3]
        R = np.stack([np.stack([1 - 2 * (y * y + z * z), 2 * (x * y + z * ww), 2 * (x * z - y * ww)], 1),
                      np.stack([2 * (x * y - z * ww), 1 - 2 * (x * x + z * z), 2 * (y * z + x * ww)], 1),
                      np.stack([2 * (x * z + y * ww), 2 * (y * z - x * ww), 1 - 2 * (x * x + y * y)], 1)], 2)
        wl = np.einsum("nij,ni->nj", R, w3)
        ii = B[:, 14:17].astype(np.float64)
        rot = 0.5 * (wl ** 2 / np.where(ii > 0, ii, np.inf)).sum(1)
        return float((0.5 * m * v2 + rot + m * 9.81 * B[:, 1].astype(np.float64)).sum())

    e0 = energy(w.bodies())
    for k in range(300):
        w.step(DT)
        if k % 30 == 29:
            B = w.bodies()
            st = w.stats()
            assert not np.isnan(B).any(), f"step {k}: NaN"
            assert st["dropped"] == 0, f"step {k}: dropped"
            assert B[:, 1].min() > 0.15, f"step {k}: a body at y {B[:, 1].min()}"
            e = energy(B)
            assert e <= e0 * (1.0 + 1e-3), f"step {k}: energy {e:.4e} above the start's {e0:.4e}"
    assert e < e0


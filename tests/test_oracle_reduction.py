"""Contact reduction in the box oracle (gpu/gpu_world.js :834-920), pinned by the reference's own GPU checks
(gpu/test_gpu.cjs redpairs :1029-1047, compound :652-674; docs/GPU.md Contact reduction) and by geometry.

  * redpairs: 80 slabs of 6 x 6 boxes on a floor that is one static body: every body pair with more than 32
    manifolds is reduced (the slab against the floor and against the ground), keeps at most 27, and keeps its
    four corner boxes (each the only one furthest out along a diagonal direction);
  * pairs at or under the threshold, and a world with reduction off, are untouched: the same step bit for bit;
  * compound: a 20 x 20 x 2 building of bricks, one body, landing at 5.9 m/s on a floor of bricks rebounds by less
    than 0.8 m/s, and reduction takes its hundreds of floor manifolds out of the overflow pass.
"""
import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd.world import params, GPU_WORLD

DT = np.float32(1.0 / 60.0)


def _redpairs():
    floor = [scenes.shape((-15 + i * 0.5 + 0.25, -0.25, -12 + k * 0.5 + 0.25), (0.25, 0.25, 0.25), mat=1)
             for i in range(60) for k in range(48)]
    items = [scenes.body(floor, fixed=True)]
    corners = []
    for i in range(10):
        for k in range(8):
            sh = [scenes.shape((-15 + i * 3 + a * 0.5 + 0.25, 0.125, -12 + k * 3 + c * 0.5 + 0.25), (0.24, 0.125, 0.24), mat=1)
                  for a in range(6) for c in range(6)]
            corners.append({0, 5, 30, 35})  # shape indices within the slab: (a, c) = (0,0), (0,5), (5,0), (5,5)
            items.append(scenes.body(sh))
    return scenes.bodies(items), corners


def _pairs(keys, shape_body):
    """body pair of each manifold (the ground as -1)"""
    a = shape_body[keys[:, 0]]
    b = np.where(keys[:, 1] >= 0, shape_body[np.maximum(keys[:, 1], 0)], -1)
    return [(int(x), int(y)) for x, y in zip(a, b)]


@pytest.mark.parametrize("order", ["reference", "coloured"])
def test_redpairs_every_big_pair_reduced_corners_kept(oracle, order):
    (b, sb, s), corners = _redpairs()
    o = oracle.BoxOracle(b, sb, s, params(substeps=8, sleep=False, **GPU_WORLD),
                         oracle.ORDER_COLOURED if order == "coloured" else oracle.ORDER_REFERENCE)
    for _ in range(3):
        o.step(DT)
    keys, _ = o.manifolds()
    pairs = _pairs(keys, sb)
    col = o.colours() if order == "coloured" else None
    by = {}
    for m, p in enumerate(pairs):
        by.setdefault(p, []).append(m)
    big = {p: ms for p, ms in by.items() if len(ms) > 32}
    assert len(big) == 160, len(big)   # each slab with the floor and with the ground
    st = o.stats()
    assert st["reduced"] > 0
    if col is not None:
        total_red = 0
        for p, ms in big.items():
            red = [m for m in ms if col[m] == -2]
            kept = [m for m in ms if col[m] != -2]
            total_red += len(red)
            assert red and 4 <= len(kept) <= 27, (p, len(red), len(kept))
            slab = max(p)
            first = int(np.nonzero(sb == slab)[0][0])
            kept_shapes = {(int(keys[m, 0]) if sb[keys[m, 0]] == slab else int(keys[m, 1])) - first for m in kept}
            assert corners[slab - 1] <= kept_shapes, (p, sorted(kept_shapes))
        assert total_red == st["reduced"]
        # pairs at or under the threshold keep everything
        assert all(col[m] != -2 for p, ms in by.items() if len(ms) <= 32 for m in ms)


def test_reduction_off_and_pairs_under_the_threshold_leave_the_step_unchanged(oracle):
    b, sb, s = scenes.rubble(300, box=(7.0, 5.0, 7.0), seed=5)
    base = params(substeps=6, sleep=False, ratio=20.0, shock=6, col_support=2)
    on = params(substeps=6, sleep=False, ratio=20.0, shock=6, col_support=2, red_k=200)
    o1 = oracle.BoxOracle(b, sb, s, base, oracle.ORDER_COLOURED)
    o2 = oracle.BoxOracle(b, sb, s, on, oracle.ORDER_COLOURED)
    for _ in range(20):
        o1.step(DT)
        o2.step(DT)
    assert o2.stats()["reduced"] == 0
    assert np.array_equal(o1.bodies(), o2.bodies())


def _compound(n=20):
    sh = [scenes.shape((-n / 2 * 0.5 + i * 0.5 + 0.25, 0.75 + j * 0.5, -n / 2 * 0.5 + k * 0.5 + 0.25), (0.25, 0.25, 0.25), mat=1)
          for i in range(n) for k in range(n) for j in range(2)]
    m = n // 2 + 2
    fl = [scenes.shape((-m / 2 + i + 0.5, 0.25, -m / 2 + k + 0.5), (0.5, 0.25, 0.5), mat=1) for i in range(m) for k in range(m)]
    return scenes.bodies([scenes.body(fl, fixed=True), scenes.body(sh, v=(0, -5.9, 0))])


def test_compound_landing_rebounds_less_than_concrete_allows(oracle):
    b, sb, s = _compound()
    runs = {}
    for red in (0, 32):
        o = oracle.BoxOracle(b, sb, s, params(substeps=12, sleep=False, **{**GPU_WORLD, "red_k": red}), oracle.ORDER_COLOURED)
        vmax, hit, ovf, nred = -1e9, False, 0, 0
        for f in range(60):
            o.step(DT)
            vy = float(o.bodies()[1, 8])
            if not hit and vy > -5.5:
                hit = True
            if hit:
                vmax = max(vmax, vy)
            st = o.stats()
            ovf = max(ovf, st["overflow"])
            nred = max(nred, st["reduced"])
        runs[red] = (vmax, ovf, nred)
    vmax, ovf, nred = runs[32]
    # gpu/test_gpu.cjs compound: concrete (e 0.1) gives 0.59 m/s of rebound; under 0.8 passes
    assert hit and vmax < 0.8, runs
    # 400 floor manifolds come down to at most 27 (the 26 directions and the deepest): at most 11 past 16 colours
    assert nred >= 350 and ovf <= 11 and runs[0][1] >= 350, runs

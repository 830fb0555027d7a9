"""The GPU world's solver extensions in the box oracle (gpu/gpu_world.js :1152-1203 mass-ratio limit and shock
propagation, :1333-1363 restitution, :954-985 supports-first colouring), pinned by closed forms and the reference's
own checks (docs/GPU.md:155-160 and gpu/test_gpu.cjs bounce :79-96, incline :105-125, landing :579-599).

Each check is one a plausible mistake fails: a restitution taken as the smaller of the pair, a sign slip in the
approach speed, a mass-ratio factor applied to the wrong side, a hold on the upper body instead of the lower, a level
that climbs more than once per pass.
"""
import math

import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd.world import params, GPU_WORLD

DT = np.float32(1.0 / 60.0)
ORDERS = ["reference", "coloured"]


def _order(oracle, name):
    return oracle.ORDER_REFERENCE if name == "reference" else oracle.ORDER_COLOURED


def _world(oracle, items, prm, order):
    b, sb, s = scenes.boxes(items)
    return oracle.BoxOracle(b, sb, s, prm, _order(oracle, order))


def _drop(oracle, e_box, e_floor, order, frames=150):
    """A 0.5 m cube dropped with its bottom face 2 m above a floor (a static slab whose top is at y = 0); the highest
    the bottom face comes back up after the first bounce."""
    cube = scenes.box((0, 2.25, 0), (0.25, 0.25, 0.25), e=e_box)
    items = [cube]
    if e_floor is not None:
        items.append(scenes.box((0, -0.5, 0), (3.0, 0.5, 3.0), density=0.0, e=e_floor))
    w = _world(oracle, items, params(sleep=False), order)
    peak, landed, prev = 0.0, False, -1.0
    for _ in range(frames):
        w.step(DT)
        b = w.bodies()[0]
        if not landed and prev < 0.0 < b[8]:
            landed = True
        if landed:
            peak = max(peak, b[1] - 0.25)
        prev = b[8]
    return peak


@pytest.mark.parametrize("order", ORDERS)
@pytest.mark.parametrize("e", [0.5, 0.3, 0.05])
def test_bounce_rebound_is_e_squared_times_the_drop(oracle, order, e):
    # docs/GPU.md:159: rebound e^2 h from 2 m (the reference GPU: 0.490 / 0.176 / 0.005 m)
    peak = _drop(oracle, e, None, order)
    want = e * e * 2.0
    assert abs(peak - want) <= 0.05 * want + 0.001, (peak, want)


def test_no_restitution_no_bounce_and_the_pair_takes_the_larger(oracle):
    assert _drop(oracle, 0.0, None, "coloured") < 0.002
    # a bouncy cube on a dead floor and a dead cube on a bouncy floor bounce alike: e = max(e_a, e_b)
    a = _drop(oracle, 0.5, 0.0, "coloured")
    b = _drop(oracle, 0.0, 0.5, "coloured")
    c = _drop(oracle, 0.3, 0.5, "coloured")
    assert abs(a - 0.5) <= 0.03 and abs(b - 0.5) <= 0.03 and abs(c - 0.5) <= 0.03, (a, b, c)


@pytest.mark.parametrize("order", ORDERS)
def test_incline_holds_at_25_and_slides_the_closed_form_at_40(oracle, order):
    # docs/GPU.md:160: friction 0.6 holds at 25 degrees; at 40 the block slides g (sin - mu cos) (1.5^2 - 0.5^2) / 2
    # between 0.5 s and 1.5 s (reference GPU: 0.000 m; 1.798 m)
    for deg in (25, 40):
        th = math.radians(deg)
        q = (0.0, 0.0, math.sin(th / 2), math.cos(th / 2))
        n = np.array([-math.sin(th), math.cos(th), 0.0])
        t = np.array([math.cos(th), math.sin(th), 0.0])
        c = np.array([0.0, 4.0, 0.0])
        slab = scenes.box(tuple(c), (6.0, 0.25, 2.0), q=q, density=0.0)
        blk = scenes.box(tuple(c + 0.5 * n), (0.25, 0.25, 0.25), q=q)
        w = _world(oracle, [blk, slab], params(sleep=False), order)
        along = []
        for _ in range(90):
            w.step(DT)
            along.append(float(w.bodies()[0, 0:3] @ t))
        moved = along[29] - along[89]  # down the slope, 0.5 s to 1.5 s
        a = max(0.0, 9.81 * (math.sin(th) - 0.6 * math.cos(th)))
        if a == 0.0:
            assert abs(moved) < 0.001, moved
        else:
            assert abs(moved - a) < 0.01 * a, (moved, a)


def _landing(oracle, hh, v, feats, order="coloured"):
    """gpu/test_gpu.cjs:579-599: a concrete cube of 0.27 m (47 kg) just above the ground, a concrete slab (hh) on it, both
    falling at v; the fastest each moves up over 40 frames."""
    hl = 0.135
    light = scenes.box((0, hl + 0.05, 0), (hl, hl, hl), density=2400.0, e=0.1, v=(0, -v, 0))
    heavy = scenes.box((0, 0.05 + 2 * hl + hh * 0.4 + 0.0005, 0), (hh, hh * 0.4, hh), density=2400.0, e=0.1, v=(0, -v, 0))
    w = _world(oracle, [light, heavy], params(sleep=False, **(GPU_WORLD if feats else {})), order)
    lmax = hmax = -1e9
    for _ in range(40):
        w.step(DT)
        b = w.bodies()
        lmax, hmax = max(lmax, b[0, 8]), max(hmax, b[1, 8])
    return hmax, lmax


@pytest.mark.parametrize("hh,v", [(0.5, 7.7), (0.5, 2.0), (1.0, 5.0), (0.3, 7.7)])
def test_landing_rebounds_within_restitution_with_shock_propagation(oracle, hh, v):
    # concrete's restitution 0.1: neither body may come back up faster than 0.1 v + 0.1 m/s (the reference's check)
    heavy, light = _landing(oracle, hh, v, True)
    assert heavy <= 0.1 * v + 0.1 and light <= 0.1 * v + 0.1, (heavy, light)


def test_landing_without_the_extensions_rebounds_as_the_reference_reports(oracle):
    # gpu/test_gpu.cjs:582-583: without shock propagation the 960 kg slab came back up at 3.65 m/s from 7.7, 0.96 from 2
    assert abs(_landing(oracle, 0.5, 7.7, False)[0] - 3.65) < 0.15
    assert abs(_landing(oracle, 0.5, 2.0, False)[0] - 0.96) < 0.1


@pytest.mark.parametrize("ratio", [0.0, 20.0])
def test_mass_ratio_limit_head_on(oracle, ratio):
    """A 1 m cube of 100x the mass of a 0.5 m cube meets it at 2 m/s with no gravity. Unlimited the two move on
    together at 2 M / (M + m) (perfectly inelastic); limited to 20:1 the heavy side responds as if 20 m, so both end near
    2 * 20 / 21, and the kinetic energy only goes down."""
    A = scenes.box((0, 10, 0), (0.5, 0.5, 0.5), density=12500.0, v=(2.0, 0, 0))
    B = scenes.box((0.755, 10, 0), (0.25, 0.25, 0.25), density=1000.0)
    w = _world(oracle, [A, B], params(sleep=False, ratio=ratio, gy=0.0), "coloured")
    for _ in range(3):
        w.step(DT)
    bd = w.bodies()
    mA, mB = 1.0 / A[0][13], 1.0 / B[0][13]
    assert abs(mA / mB - 100.0) < 1e-3
    want = 2.0 * 20.0 / 21.0 if ratio else 2.0 * 100.0 / 101.0
    assert abs(bd[0, 7] - want) < 0.005, (bd[0, 7], want)
    assert bd[1, 7] >= bd[0, 7] - 1e-4  # the light one is not left behind
    ke0 = 0.5 * mA * 4.0
    ke1 = 0.5 * mA * (bd[0, 7:10] ** 2).sum() + 0.5 * mB * (bd[1, 7:10] ** 2).sum()
    assert ke1 <= ke0


@pytest.mark.parametrize("passes", [6, 12])
def test_levels_of_a_stack_climb_one_per_pass(oracle, passes):
    # n cubes stacked on the ground: the k-th from the bottom is at level k (height 255 - k) once the passes reach it,
    # and a pass lifts a chain by exactly one level, so only the lowest `passes` get one; the order does not matter
    n = 10
    items = [scenes.box((0, 0.25 + 0.5 * k, 0), (0.25, 0.25, 0.25)) for k in range(n)]
    prm = params(sleep=False, shock=passes)
    got = []
    for order in ORDERS:
        w = _world(oracle, items, prm, order)
        w.step(DT)
        got.append(w.levels())
    want = np.array([255 - (k + 1) if k < passes else 0 for k in range(n)])
    assert np.array_equal(got[0], want), got[0]
    assert np.array_equal(got[1], want), got[1]


def test_supports_first_colouring_valid_and_supports_kept_out_of_the_overflow(oracle):
    b, sb, s = scenes.rubble(600, box=(8.0, 6.0, 8.0), seed=9)
    over = {}
    for rs in (0, 2):
        prm = params(substeps=4, col_limit=4, col_support=rs)
        w = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
        nsup = 0
        for _ in range(6):
            pre = w.bodies()
            w.step(DT)
            keys, _ = w.manifolds()
            col = w.colours()
            awake0 = pre[:, 17] > 0.5
            ba = sb[keys[:, 0]]
            bb = np.where(keys[:, 1] >= 0, sb[np.maximum(keys[:, 1], 0)], -1)
            mov_a = awake0[ba]
            mov_b = np.where(bb >= 0, awake0[np.maximum(bb, 0)], False)
            seen = set()
            for m in range(len(keys)):
                if col[m] >= 4:
                    continue
                for body, mv in ((ba[m], mov_a[m]), (bb[m], mov_b[m])):
                    if mv:
                        assert (body, col[m]) not in seen
                        seen.add((body, col[m]))
            support = ~(mov_a & mov_b)
            nsup += int(((col >= 4) & support).sum())
        over[rs] = nsup
    assert over[2] < over[0], over

"""Pins the box-core oracle (oracle/box_world.c) before it is trusted as the checker.

Every check here is against something the oracle does not restate itself: the reference's recorded numbers
(docs/GPU.md:89 candidate count, docs/GPU.md:156-158 stack drift and pile heights, engine/src/phys.rs:663-696 test
asserts), brute force, breadth-first search, and the colouring's defining invariant.
"""
import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd.world import params

DT = np.float32(1.0 / 60.0)


def test_pile_first_frame_candidates_match_the_reference(oracle):
    # docs/GPU.md:89: the first frame of the 2,000-box pile has 5,871 candidate pairs (box pairs + ground)
    b, sb, s = scenes.pile(2000)
    w = oracle.BoxOracle(b, sb, s, params(substeps=4, sleep=False), oracle.ORDER_REFERENCE)
    w.step(DT)
    st = w.stats()
    assert st["pairs"] + st["ground"] == 5871


def test_canonical_candidates_equal_brute_force(oracle):
    b, sb, s = scenes.pile(600)
    w = oracle.BoxOracle(b, sb, s, params(substeps=4, sleep=False), oracle.ORDER_COLOURED)
    bf = oracle.BoxOracle(b, sb, s, params(substeps=4, sleep=False), oracle.ORDER_COLOURED).bruteforce_pairs(DT)
    w.step(DT)
    assert np.array_equal(w.candidates(), bf)
    # a few steps later the set still equals brute force on the current pose
    for _ in range(20):
        w.step(DT)
    ref = w.bodies()
    bf2 = oracle.BoxOracle(ref, sb, s, params(substeps=4, sleep=False), oracle.ORDER_COLOURED).bruteforce_pairs(DT)
    w.step(DT)
    # (the ground candidates depend on the awake flags only; the pile stays awake without sleep)
    assert np.array_equal(w.candidates(), bf2)


@pytest.mark.parametrize("order", [0, 1])
def test_stack_of_20_gate(oracle, order):
    # engine/src/phys.rs:672-683 (drift < 10 mm, speed < 10 mm/s after 10 s); docs/GPU.md:156 records 6.93 mm and
    # 1.7 mm/s for the CPU reference without sleep
    b, sb, s = scenes.stack(20)
    w = oracle.BoxOracle(b, sb, s, params(substeps=12, sleep=False), order)
    top0 = w.bodies()[19, 1]
    maxv = 0.0
    for f in range(600):
        w.step(DT)
        if f >= 540:
            maxv = max(maxv, float(np.linalg.norm(w.bodies()[:, 7:10], axis=1).max()))
    top = w.bodies()[19, :3]
    drift = float(np.linalg.norm(top - np.array([0.0, top0, 0.0])))
    assert drift < 0.01 and maxv < 0.01
    if order == 0:
        assert round(drift * 1e3, 2) == 6.93
        assert round(maxv * 1e3, 1) == 1.7


def test_pile_heights_after_two_seconds_match_the_reference(oracle):
    # docs/GPU.md:158: pile of 2,000 after 2 s at 4 sub-steps, CPU y 0.117-4.828 m
    b, sb, s = scenes.pile(2000)
    w = oracle.BoxOracle(b, sb, s, params(substeps=4, sleep=False), oracle.ORDER_REFERENCE)
    for _ in range(120):
        w.step(DT)
    y = w.bodies()[:, 1]
    assert round(float(y.min()), 3) == 0.117
    assert round(float(y.max()), 3) == 4.828
    # the colour-ordered step (the device order) stays within the reference's own GPU-vs-CPU tolerance
    # (gpu/test_gpu.cjs:1049-1061: |min y| within 0.02 m, |max y| within 0.1 m)
    wc = oracle.BoxOracle(b, sb, s, params(substeps=4, sleep=False), oracle.ORDER_COLOURED)
    for _ in range(120):
        wc.step(DT)
    yc = wc.bodies()[:, 1]
    assert abs(yc.min() - y.min()) < 0.02 and abs(yc.max() - y.max()) < 0.1
    assert np.isfinite(wc.bodies()).all()


def test_single_box_rests(oracle):
    # engine/src/phys.rs:663-670
    body, shape = scenes.box([0.0, 1.0, 0.0], [0.25, 0.125, 0.5], q=_axis_angle([1.0, 0.3, 0.2], 0.7), density=2000.0)
    b, sb, s = scenes.boxes([(body, shape)])
    for order in (0, 1):
        w = oracle.BoxOracle(b, sb, s, params(), order)
        for _ in range(300):
            w.step(DT)
        r = w.bodies()[0]
        assert 0.1 < r[1] < 0.3 and np.linalg.norm(r[7:10]) < 0.01


def test_tilted_column_topples_and_settles(oracle):
    # engine/src/phys.rs:685-696
    items = [scenes.box([0.06 * k, 0.25 + 0.5 * k, 0.0], [0.25, 0.25, 0.25]) for k in range(10)]
    b, sb, s = scenes.boxes(items)
    w = oracle.BoxOracle(b, sb, s, params(), oracle.ORDER_COLOURED)
    for _ in range(600):
        w.step(DT)
    r = w.bodies()
    assert r[:, 1].max() < 3.0 and r[:, 1].min() > 0.2
    assert (np.linalg.norm(r[:, 7:10], axis=1) < 0.05).all()


def _axis_angle(a, ang):
    a = np.asarray(a, np.float64)
    a = a / np.linalg.norm(a)
    return [*(a * np.sin(ang / 2)), np.cos(ang / 2)]


def _bfs_labels(nb, edges):
    adj = [[] for _ in range(nb)]
    for a, b in edges:
        adj[a].append(b)
        adj[b].append(a)
    lab = -np.ones(nb, np.int64)
    for s in range(nb):
        if lab[s] >= 0:
            continue
        comp, q = [s], [s]
        lab[s] = s
        while q:
            u = q.pop()
            for v in adj[u]:
                if lab[v] < 0:
                    lab[v] = s
                    comp.append(v)
                    q.append(v)
        lab[comp] = min(comp)
    return lab


def test_colouring_valid_complete_deterministic_and_islands_equal_bfs(oracle):
    b, sb, s = scenes.rubble(400, box=(8.0, 6.0, 8.0), seed=3)
    prm = params(substeps=4)
    w1 = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    w2 = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    for step in range(30):
        pre = w1.bodies()
        w1.step(DT)
        w2.step(DT)
        keys, _ = w1.manifolds()
        col = w1.colours()
        assert np.array_equal(col, w2.colours())
        assert np.array_equal(w1.bodies(), w2.bodies())
        K = int(prm[20])
        # validity: no two manifolds of one colour share a body that was moving at the step start
        awake0 = pre[:, 17] > 0.5
        ba = sb[keys[:, 0]]
        bb = np.where(keys[:, 1] >= 0, sb[np.maximum(keys[:, 1], 0)], -1)
        seen = {}
        for m in range(len(keys)):
            if col[m] >= K:
                continue
            for body in (ba[m], bb[m]):
                if body >= 0 and awake0[body]:
                    assert (body, col[m]) not in seen, (m, seen.get((body, col[m])))
                    seen[(body, col[m])] = m
        assert ((col >= 0) & (col <= K)).all()
        # islands: union over manifolds between two moving bodies, labelled by the minimum index
        both = (bb >= 0) & awake0[ba] & np.where(bb >= 0, awake0[np.maximum(bb, 0)], False)
        lab = _bfs_labels(len(b), list(zip(ba[both], bb[both])))
        assert np.array_equal(w1.islands(), lab)

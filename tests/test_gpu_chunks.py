"""Convex chunks and contact reduction on the GPU against the oracle in the device's colour order, through the C ABI.

The bar: candidates, manifold keys and data, colours (-2 for a reduced manifold), islands and body states equal bit for
bit, every step. Scenes: a heap of chunks from the library's patterns with boxes and a mixed compound, chunks dropped on
fixed chunks, chunks thrown at the ground (two ground patches per chunk), the reference's redpairs scene (160 reduced
pairs), and the 800-brick building landing on a floor of bricks.
"""
import math

import numpy as np
import pytest

import shatterforge_amd as sf
from shatterforge_amd import hulls, scenes, world
from shatterforge_amd.world import params

pytestmark = pytest.mark.gpu
DT = np.float32(1.0 / 60.0)


def _qa(axis, ang):
    a = np.asarray(axis, np.float64)
    a = a / np.linalg.norm(a)
    s = math.sin(ang / 2)
    return np.array([a[0] * s, a[1] * s, a[2] * s, math.cos(ang / 2)], np.float32)


def _same(w, o, tag):
    assert np.array_equal(w.candidates(), o.candidates()), f"{tag}: candidates"
    gk, gd = w.manifolds()
    ok, od = o.manifolds()
    assert np.array_equal(gk, ok), f"{tag}: manifold keys"
    assert np.array_equal(gd, od), f"{tag}: manifold data max diff {np.abs(gd - od).max() if len(gd) else 0}"
    assert np.array_equal(w.colours(), o.colours()), f"{tag}: colours"
    assert np.array_equal(w.islands(), o.islands()), f"{tag}: islands"
    gb, ob = w.bodies(), o.bodies()
    assert np.array_equal(gb, ob), f"{tag}: bodies max diff {np.abs(gb - ob).max()}"


def _run(oracle, scene, prm, steps, tag, every=1):
    b, sb, s = scene
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED, hulls=hulls.tables())
    for k in range(steps):
        w.step(DT)
        o.step(DT)
        if (k + 1) % every == 0 or k + 1 == steps:
            _same(w, o, f"{tag} step {k}")
    return w, o


def _heap(seed=3, n=120):
    """chunks of the library's patterns (and a few boxes), a brick broken into a body of chunks, one body of a box and
    two chunks; dropped on a jittered grid"""
    rng = np.random.default_rng(seed)
    items = []
    for k in range(n):
        pat = int(rng.integers(0, hulls.pattern_count()))
        ids = hulls.pattern(pat)
        h = ids[int(rng.integers(0, len(ids)))] if k % 7 else 0
        hh = (hulls.cell(h)["box_h"] * 0.25 + 0.04) if h else np.array([0.2, 0.15, 0.25])
        gx, gz, gy = k % 6, (k // 6) % 6, k // 36
        p = ((gx - 2.5) * 0.8 + (rng.random() - 0.5) * 0.1, 0.5 + gy * 0.8, (gz - 2.5) * 0.8 + (rng.random() - 0.5) * 0.1)
        items.append(scenes.body([scenes.shape(p, hh, q=_qa(rng.random(3) - 0.5, rng.random() * 6), mat=1 + k % 5, hull=h)],
                                 v=(0, -2, 0)))
    items.append(scenes.body(list(hulls.shapes((0.0, 4.5, 0.0), _qa((1, 0, 0), 0.3), (0.5, 0.25, 0.5), 4, mat=2))))
    items.append(scenes.body([scenes.shape((2.0, 5.0, 0.0), (0.25, 0.25, 0.25), mat=3),
                              scenes.shape((2.5, 5.0, 0.0), (0.2, 0.2, 0.2), mat=3, hull=hulls.pattern(2)[0]),
                              scenes.shape((2.0, 5.5, 0.0), (0.2, 0.2, 0.2), mat=3, hull=hulls.pattern(7)[1])]))
    return scenes.bodies(items)


def test_chunk_heap_bit_exact_with_sleep_and_reduction(oracle):
    print("case: chunk heap 122 bodies, GPU world + reduction, sleep on, 90 steps", flush=True)
    _run(oracle, _heap(), params(substeps=8, **world.GPU_WORLD), 90, "heap", every=10)


def test_chunks_on_fixed_chunks_bit_exact(oracle):
    print("case: 40 chunks dropped on fixed chunks, 60 steps", flush=True)
    rng = np.random.default_rng(9)
    H = hulls.count()
    items = []
    for k in range(40):
        x, z = (k % 10) * 3 - 15, (k // 10) * 3 - 6
        ia, ib = int(rng.integers(0, H)), int(rng.integers(0, H))
        ha = hulls.cell(ia)["box_h"] * 0.4 + 0.1
        hb = hulls.cell(ib)["box_h"] * 0.25 + 0.05
        items.append(scenes.body([scenes.shape((x, 1.5, z), ha, q=_qa(rng.random(3) - 0.5, rng.random() * 6), hull=ia)], fixed=True))
        items.append(scenes.body([scenes.shape((x + (rng.random() - 0.5) * 0.2, 2.4, z), hb, q=_qa(rng.random(3) - 0.5, rng.random() * 6), hull=ib)]))
    _run(oracle, scenes.bodies(items), params(substeps=12, sleep=False, **world.GPU_WORLD), 60, "hullpairs", every=5)


def test_chunks_thrown_at_the_ground_bit_exact(oracle):
    print("case: 60 spinning chunks thrown at the ground (two ground patches), 40 steps", flush=True)
    rng = np.random.default_rng(21)
    H = hulls.count()
    items = []
    for k in range(60):
        h0 = 1 + int(rng.integers(0, H - 1))
        hh = hulls.cell(h0)["box_h"] * 0.25 + 0.04
        items.append(scenes.body([scenes.shape(((k % 10) * 2 - 10, 1.5 + rng.random(), (k // 10) * 2 - 6), hh,
                                               q=_qa(rng.random(3) - 0.5, rng.random() * 6), hull=h0)],
                                 v=((rng.random() - 0.5) * 6, -4 - rng.random() * 10, (rng.random() - 0.5) * 6), w=(rng.random(3) - 0.5) * 30))
    w, o = _run(oracle, scenes.bodies(items), params(substeps=12, sleep=False, **world.GPU_WORLD), 40, "hullthrow", every=2)
    keys, _ = w.manifolds()
    assert (keys[:, 1] == -2).any()   # some chunk touched the ground at more than 4 corners


def test_redpairs_reduced_bit_exact(oracle):
    print("case: redpairs, 80 slabs of 36 boxes on a static floor, 3 steps", flush=True)
    floor = [scenes.shape((-15 + i * 0.5 + 0.25, -0.25, -12 + k * 0.5 + 0.25), (0.25, 0.25, 0.25)) for i in range(60) for k in range(48)]
    items = [scenes.body(floor, fixed=True)]
    for i in range(10):
        for k in range(8):
            items.append(scenes.body([scenes.shape((-15 + i * 3 + a * 0.5 + 0.25, 0.125, -12 + k * 3 + c * 0.5 + 0.25), (0.24, 0.125, 0.24))
                                      for a in range(6) for c in range(6)]))
    w, o = _run(oracle, scenes.bodies(items), params(substeps=8, sleep=False, **world.GPU_WORLD), 3, "redpairs")
    assert w.stats()["reduced"] == o.stats()["reduced"] > 0
    assert (w.colours() == -2).sum() == w.stats()["reduced"]


def test_compound_landing_reduced_bit_exact(oracle):
    print("case: 800-brick body landing on a floor of bricks, reduction on, 30 steps", flush=True)
    n = 20
    sh = [scenes.shape((-n / 2 * 0.5 + i * 0.5 + 0.25, 0.75 + j * 0.5, -n / 2 * 0.5 + k * 0.5 + 0.25), (0.25, 0.25, 0.25))
          for i in range(n) for k in range(n) for j in range(2)]
    m = n // 2 + 2
    fl = [scenes.shape((-m / 2 + i + 0.5, 0.25, -m / 2 + k + 0.5), (0.5, 0.25, 0.5)) for i in range(m) for k in range(m)]
    scene = scenes.bodies([scenes.body(fl, fixed=True), scenes.body(sh, v=(0, -5.9, 0))])
    w, o = _run(oracle, scene, params(substeps=12, sleep=False, **world.GPU_WORLD), 30, "compound")
    assert w.stats()["overflow"] == 0


def test_reduction_with_a_small_threshold_on_rubble_bit_exact(oracle):
    print("case: rubble(400) reduction threshold 8 (many reduced pairs), 20 steps", flush=True)
    b, sb, s = scenes.rubble(400, box=(8.0, 6.0, 8.0), seed=23)
    prm = params(substeps=4, **{**world.GPU_WORLD, "red_k": 8})
    w, o = _run(oracle, (b, sb, s), prm, 20, "rubble red 8", every=5)

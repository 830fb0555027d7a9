"""Particle debris on the GPU against the oracle (oracle/box_world.c's particle stage) through the C ABI: particle
records (position, velocity, drawing rotation, state, radius) and the bodies equal bit for bit after every step."""
import numpy as np
import pytest

import shatterforge_amd as sf
from shatterforge_amd import scenes, world
from shatterforge_amd.world import params

pytestmark = pytest.mark.gpu
DT = np.float32(1.0 / 60.0)


def _pair(oracle, items, prm, cap):
    b, sb, s = scenes.boxes(items)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    w.particles_capacity(cap)
    o.particles_capacity(cap)
    return w, o


def _same(w, o, tag):
    gp, op = w.particles(), o.particles()
    assert gp.shape == op.shape, tag
    assert np.array_equal(gp, op), f"{tag}: particles max diff {np.abs(gp - op).max()}"
    gb, ob = w.bodies(), o.bodies()
    assert np.array_equal(gb, ob), f"{tag}: bodies max diff {np.abs(gb - ob).max()}"
    assert np.array_equal(w.particle_stats(), o.particle_stats()), tag


def _pour(n, seed, lo=-0.6, hi=0.6, y0=1.5, dy=0.004, r=0.08):
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0] = rng.uniform(lo, hi, n)
    rec[:, 1] = y0 + dy * np.arange(n)
    rec[:, 2] = rng.uniform(lo, hi, n)
    rec[:, 6] = r
    rec[:, 7] = rng.choice([0.5, 0.6], n)
    return rec


def test_heap_and_a_dropped_box_bit_exact(oracle):
    print("case: 1,500 particles poured on a table, a box dropped on the heap, 150 steps", flush=True)
    table = scenes.box((0, 0.5, 0), (1.0, 0.5, 1.0), density=0.0)
    box = scenes.box((0.3, 5.0, 0.2), (0.3, 0.3, 0.3), density=800.0)
    w, o = _pair(oracle, [table, box], params(sleep=True), 2048)
    rec = _pour(1500, 3)
    assert np.array_equal(w.add_particles(rec), o.add_particles(rec))
    for k in range(150):
        w.step(DT)
        o.step(DT)
        _same(w, o, f"step {k}")
    st = w.particle_stats()
    assert st[0] == 1500 and st[1] > 0


def test_blast_on_a_sleeping_heap_and_slot_reuse_bit_exact(oracle):
    print("case: 1,000-particle heap asleep, a blast, kills and new particles in the freed slots", flush=True)
    w, o = _pair(oracle, [scenes.box((0, -20, 0), (0.5, 0.5, 0.5), density=0.0)], params(sleep=True), 1200)
    rec = _pour(1000, 5, y0=0.3, dy=0.002)
    w.add_particles(rec)
    o.add_particles(rec)
    for k in range(240):
        w.step(DT)
        o.step(DT)
        if k % 20 == 0:
            _same(w, o, f"settle {k}")
    _same(w, o, "settled")
    for i in range(0, 1000, 7):
        w.kill_particle(i)
        o.kill_particle(i)
    more = _pour(200, 6, y0=2.0, dy=0.01)
    assert np.array_equal(w.add_particles(more), o.add_particles(more))
    w.blast_particles((0.2, 0.0, -0.1), 1.5, 8.0)
    o.blast_particles((0.2, 0.0, -0.1), 1.5, 8.0)
    for k in range(60):
        w.step(DT)
        o.step(DT)
        _same(w, o, f"after blast {k}")
    w.wake_particles()
    o.wake_particles()
    w.step(DT)
    o.step(DT)
    _same(w, o, "wake all")


def test_debris_on_mixed_rubble_bit_exact(oracle):
    print("case: rubble(300) with the GPU world's solver, 4,000 particles raining in, 60 steps", flush=True)
    b, sb, s = scenes.rubble(300, box=(7.0, 5.0, 7.0), seed=13)
    prm = params(substeps=6, **world.GPU_WORLD)
    w = sf.World.from_scene(b, sb, s, prm)
    o = oracle.BoxOracle(b, sb, s, prm, oracle.ORDER_COLOURED)
    w.particles_capacity(4096)
    o.particles_capacity(4096)
    rec = _pour(4000, 9, lo=-3.0, hi=3.0, y0=6.0, dy=0.002, r=0.1)
    rec[:, 3] = np.random.default_rng(10).normal(0, 1.0, 4000)
    w.add_particles(rec)
    o.add_particles(rec)
    for k in range(60):
        w.step(DT)
        o.step(DT)
        _same(w, o, f"step {k}")

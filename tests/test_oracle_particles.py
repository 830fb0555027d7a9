"""Particle debris in the box oracle (gpu/gpu_world.js:1671-2038 restated), pinned by closed forms and the reference's
own heap check (docs/GPU.md:168: 2,000 particles poured on a 2 x 2 m table, none under the ground or inside the table,
all asleep by 5 s).

Each check is one a plausible mistake fails: gravity applied twice or not per sub-step, the speed cap missing, the
friction pair mean or the Coulomb limit wrong, the rest timer's threshold, the blast's fall-off or its upward share,
the pair push weights (a moving pair shares the push), and a particle pushing a body (the coupling is one way).
"""
import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd.world import params

DT = np.float32(1.0 / 60.0)
G = 9.81


def _world(oracle, items=(), prm=None, cap=4096):
    items = list(items) or [scenes.box((0, -20, 0), (0.5, 0.5, 0.5), density=0.0)]  # (a static box far below)
    b, sb, s = scenes.boxes(items)
    o = oracle.BoxOracle(b, sb, s, params(sleep=True) if prm is None else prm, oracle.ORDER_COLOURED)
    o.particles_capacity(cap)
    return o


def test_free_fall_four_sub_steps(oracle):
    o = _world(oracle)
    o.add_particles([[0, 10, 0, 0, 0, 0, 0.1, 0.6]])
    o.step(DT)
    p = o.particles()[0]
    h = float(DT) / 4
    # v after k sub-steps: k g h; the position drops g h^2 (1 + 2 + 3 + 4)
    assert abs(p[1] - (10.0 - 10.0 * G * h * h)) < 2e-5
    assert abs(p[4] + G * float(DT)) < 0.01 * G * float(DT)
    assert p[10] == 1.0


def test_speed_capped_at_40(oracle):
    o = _world(oracle, prm=params(sleep=True, gy=0.0))
    o.add_particles([[0, 10, 0, 100.0, 0, 0, 0.1, 0.6]])
    o.step(DT)
    v = o.particles()[0, 3:6]
    assert abs(np.linalg.norm(v) - 40.0) < 1e-3


@pytest.mark.parametrize("mu_p", [0.6, 0.2])
def test_slide_stops_at_the_coulomb_distance(oracle, mu_p):
    # on the ground the push each sub-step is g h^2 and the slide taken back is at most mu times it: deceleration
    # mu g with mu the mean of the particle's and the ground's (0.6), so the slide is v^2 / (2 mu g)
    o = _world(oracle)
    v0 = 3.0
    o.add_particles([[0, 0.1, 0, v0, 0, 0, 0.1, mu_p]])
    for _ in range(150):
        o.step(DT)
    p = o.particles()[0]
    mu = 0.5 * (mu_p + 0.6)
    want = v0 * v0 / (2 * mu * G)
    assert abs(p[0] - want) < 0.03 * want, (p[0], want)
    assert abs(p[3]) < 1e-6 and abs(p[1] - 0.1) < 1e-4


def test_rest_on_the_ground_sleeps_after_half_a_second(oracle):
    o = _world(oracle)
    o.add_particles([[0, 0.1, 0, 0, 0, 0, 0.1, 0.6]])
    states = []
    for _ in range(40):
        o.step(DT)
        states.append(o.particles()[0, 10])
    first = states.index(2.0)
    # the timer adds dt per step while the particle touches and stays within 1 cm: past 0.5 s after 30 or 31 steps
    # (30 sums of f32 1/60 land a hair above 0.5)
    assert first in (29, 30), states
    assert o.particle_stats()[2] == 1


def test_blast_gives_the_fall_off_impulse(oracle):
    o = _world(oracle, prm=params(sleep=True, gy=0.0))
    o.add_particles([[1.0, 5.0, 0, 0, 0, 0, 0.1, 0.6]])
    o.blast_particles((0.0, 5.0, 0.0), 4.0, 10.0)
    o.step(DT)
    v = o.particles()[0, 3:6]
    # (d / |d| + (0, 0.3, 0)) * speed * (1 - |d| / reach) with d = (1, 0, 0)
    want = np.array([1.0, 0.3, 0.0]) * 10.0 * (1 - 1.0 / 4.0)
    assert np.abs(v - want).max() < 1e-3, (v, want)
    # out of reach: nothing
    o2 = _world(oracle, prm=params(sleep=True, gy=0.0))
    o2.add_particles([[5.0, 5.0, 0, 0, 0, 0, 0.1, 0.6]])
    o2.blast_particles((0.0, 5.0, 0.0), 4.0, 10.0)
    o2.step(DT)
    assert np.abs(o2.particles()[0, 3:6]).max() == 0.0


def test_head_on_pair_separates_symmetrically(oracle):
    o = _world(oracle, prm=params(sleep=True, gy=0.0))
    o.add_particles([[-0.3, 5.0, 0, 2.0, 0, 0, 0.1, 0.0], [0.3, 5.0, 0, -2.0, 0, 0, 0.1, 0.0]])
    for _ in range(20):
        o.step(DT)
    p = o.particles()
    assert abs(p[0, 3] + p[1, 3]) < 1e-5 and abs(p[0, 0] + p[1, 0]) < 1e-5
    assert p[1, 0] - p[0, 0] >= 0.2 - 1e-3   # apart by at least the two radii
    assert p[0, 3] <= 0.0                     # and no longer approaching


def _heap(oracle, n=2000, seed=1, with_box=False):
    table = scenes.box((0, 0.5, 0), (1.0, 0.5, 1.0), density=0.0)
    items = [table]
    if with_box:
        items.append(scenes.box((0.3, 4.0, 0.2), (0.3, 0.3, 0.3), density=800.0))
    o = _world(oracle, items)
    rng = np.random.default_rng(seed)
    rec = np.zeros((n, 8), np.float32)
    rec[:, 0] = rng.uniform(-0.6, 0.6, n)
    rec[:, 1] = 1.5 + 0.004 * np.arange(n)
    rec[:, 2] = rng.uniform(-0.6, 0.6, n)
    rec[:, 6] = 0.08
    rec[:, 7] = 0.6
    o.add_particles(rec)
    return o


def test_heap_on_a_table_stays_out_of_the_table_and_sleeps(oracle):
    o = _heap(oracle)
    for _ in range(360):
        o.step(DT)
    p = o.particles()
    r = p[:, 11]
    assert (p[:, 1] >= r - 1e-3).all()                       # none under the ground
    inside = (np.abs(p[:, 0]) < 1.0 - 1e-3) & (np.abs(p[:, 2]) < 1.0 - 1e-3) & (p[:, 1] < 1.0 + r - 2e-3)
    assert not inside.any()                                   # none inside the table
    # a heap on it (the reference: 0.6 m on the table, the rest at the angle of repose around it)
    on = (np.abs(p[:, 0]) < 1.0) & (np.abs(p[:, 2]) < 1.0)
    assert on.mean() > 0.1 and 0.3 < p[on, 1].max() - 1.0 < 0.8
    st = o.particle_stats()
    assert st[2] == len(p) and (p[:, 10] == 2.0).all()       # all asleep by 6 s


def test_particles_never_push_a_body(oracle):
    # a box dropped onto a sleeping heap: the box's motion is the same as with no particles at all (one-way coupling),
    # and the heap wakes under it
    o = _heap(oracle, n=600, with_box=True)
    table = scenes.box((0, 0.5, 0), (1.0, 0.5, 1.0), density=0.0)
    box = scenes.box((0.3, 4.0, 0.2), (0.3, 0.3, 0.3), density=800.0)
    b, sb, s = scenes.boxes([table, box])
    ref = oracle.BoxOracle(b, sb, s, params(sleep=True), oracle.ORDER_COLOURED)
    woke = 0
    for f in range(90):
        o.step(DT)
        ref.step(DT)
        assert np.array_equal(o.bodies(), ref.bodies()), f"step {f}"
        if f > 40:
            woke = max(woke, int(o.particle_stats()[1]))
    p = o.particles()
    # no particle ends inside the box (its final pose)
    bx = o.bodies()[1]
    q = bx[3:7]
    R = np.array([[1 - 2 * (q[1] ** 2 + q[2] ** 2), 2 * (q[0] * q[1] - q[2] * q[3]), 2 * (q[0] * q[2] + q[1] * q[3])],
                  [2 * (q[0] * q[1] + q[2] * q[3]), 1 - 2 * (q[0] ** 2 + q[2] ** 2), 2 * (q[1] * q[2] - q[0] * q[3])],
                  [2 * (q[0] * q[2] - q[1] * q[3]), 2 * (q[1] * q[2] + q[0] * q[3]), 1 - 2 * (q[0] ** 2 + q[1] ** 2)]])
    lc = (p[:, 0:3] - bx[0:3]) @ R
    assert not (np.abs(lc) < 0.3 - 0.02).all(axis=1).any()
    assert woke > 0

"""Pins of the lattice-world oracle (oracle/lattice.c, the legacy brick engine's frame) against what the physics and
the reference's own checks fix, in both of its orders:

  free fall with drag ........ v' = (v - g h)(1 - 0.01 h) per substep (src/lib.rs:1769-1783, :1711-1714)
  column load field .......... phi_0 = n M / (4 s), phi_{i+1} = phi_i + M (n-1-i) / s after the warm-up sweeps
  blast ...................... bricks within r BLAST[m] / 2 of the charge removed, the rest of the structure kept
  connectivity ............... a column cut at its footing becomes one loose body of the bricks above the cut
  stability .................. a standing tower with its load field live for 5 s loses nothing (tests/stab.mjs)
  rest ....................... a brick on the ground settles and sleeps (integrate, :1727-1738)
  g - a at rest .............. a piece lying still has ext ~ 0 on its bricks (structural :1284-1308)
"""
import numpy as np
import pytest

import lattice_cases as lc

C = lc.C


def _world(oracle, order, nb=4096, nd=1024):
    return oracle.LatticeOracle(order=order, brick_cap=nb, body_cap=nd)


@pytest.mark.parametrize("order", [0, 1])
def test_free_fall_with_drag(oracle, order):
    O = _world(oracle, order)
    O.add_body(np.array([2], np.uint8), np.array([[C, 100, C]], np.int32))
    O.finalize()
    y0 = float(O.bodies()[0, 5])
    v, y, h, g = 0.0, y0, 1.0 / 120.0, 9.81
    for f in range(30):
        O.step(1.0 / 60.0)
        for _ in range(2):
            v -= g * h
            y += v * h
            v *= 1.0 - 0.01 * h
    b = O.bodies()[0]
    assert abs(b[12] - v) < 1e-4 * abs(v) and abs(b[5] - y) < 1e-4 * abs(y0 - y) + 1e-5
    # a dropped drag term or gravity applied after the move would miss by ~g h^2 per substep, far outside this


@pytest.mark.parametrize("order", [0, 1])
@pytest.mark.parametrize("mat", [1, 2, 3])
def test_column_load_field_closed_form(oracle, order, mat):
    n = 12
    O = _world(oracle, order)
    O.add_structure(np.full(n, mat, np.uint8), np.stack([np.full(n, C), np.arange(n), np.full(n, C)], 1).astype(np.int32))
    O.finalize()
    O.warm_all()
    O.step()
    M = [0, 300.0, 230.0, 250.0][mat]
    s = [0, 1.0, 0.8, 2.0][mat]
    phi = np.zeros(n)
    phi[0] = n * M / (4 * s)
    for i in range(n - 1):
        phi[i + 1] = phi[i] + M * (n - 1 - i) / s
    got = O.bricks()[:, 5].astype(np.float64)
    assert np.abs(got - phi).max() < 1e-4 * phi.max()


@pytest.mark.parametrize("order", [0, 1])
def test_blast_removes_the_bricks_in_range(oracle, order):
    mat, xyz, infill = lc.small_tower(8, 6)
    O = _world(oracle, order)
    O.add_structure(mat, xyz, infill)
    O.finalize()
    c = np.array([0.0, 0.6, 0.0])
    r = 2.0
    O.explode(c[0], c[1], c[2], r, 1.4)
    pos = np.stack([(xyz[:, 0] - C) * 0.5 + 0.25, xyz[:, 1] * 0.5 + 0.25, (xyz[:, 2] - C) * 0.5 + 0.25], 1)
    d = np.linalg.norm(pos - c, axis=1)
    blast = np.array([0.0, 0.75, 0.85, 0.35, 1.4, 0.95])[mat]
    gone = d < r * blast * 0.5
    got = O.bricks()[:, 0] == 0
    assert gone.sum() > 0 and np.array_equal(got, gone)


@pytest.mark.parametrize("order", [0, 1])
def test_cut_column_detaches_as_one_body(oracle, order):
    n = 8
    O = _world(oracle, order)
    xyz = np.stack([np.full(n, C), np.arange(n), np.full(n, C)], 1).astype(np.int32)
    O.add_structure(np.full(n, 2, np.uint8), xyz)
    O.finalize()
    # a charge at the footing: masonry within 0.85 r / 2 of (0.25, 0.25, 0.25) goes, r = 0.8 takes the bottom brick
    O.explode(0.25, 0.25, 0.25, 0.8, 1.0)
    assert (O.bricks()[:, 0] == 0).sum() == 1
    O.step()
    b = O.bodies()
    live = b[b[:, 0] > 0]
    assert len(live) == 1 and live[0, 2] == n - 1
    assert O.stats()[0] == 0


@pytest.mark.parametrize("order", [0, 1])
def test_tower_stands_with_live_load_field(oracle, order):
    # tests/stab.mjs: every structure's load field live (dbg_wake_all every second), nothing lost
    mat, xyz, infill = lc.small_tower(8, 16)
    O = _world(oracle, order)
    O.add_structure(mat, xyz, infill)
    O.finalize()
    O.warm_all()
    O.step()
    n0 = O.structs()[0, 5]
    for f in range(300):
        if f % 60 == 0:
            O.wake_all()
        O.step()
    st = O.structs()[0]
    assert st[5] == n0 == st[4] and O.stats()[2] == 0
    assert O.stats()[6] < 1.0


@pytest.mark.parametrize("order", [0, 1])
def test_brick_settles_and_sleeps(oracle, order):
    O = _world(oracle, order)
    O.add_body(np.array([2], np.uint8), np.array([[C, 2, C]], np.int32))
    O.finalize()
    for f in range(120):
        O.step()
    b = O.bodies()[0]
    assert b[0] == 1 and b[1] == 1, "asleep"
    assert abs(b[5] - 0.25) < 0.05


@pytest.mark.parametrize("order", [0, 1])
def test_g_minus_a_vanishes_at_rest(oracle, order):
    O = _world(oracle, order)
    cells = np.array([[x, y, z] for x in range(2) for y in range(2) for z in range(2)], np.int32) + [C, 0, C]
    O.add_body(np.full(8, 1, np.uint8), cells)
    O.finalize()
    ext = None
    for f in range(24):
        O.step()
        b = O.bodies()[0]
        if b[1] == 0 and b[3] > 0:
            ext = O.bricks()[:, 6].copy()
    assert ext is not None
    assert np.abs(ext).max() < 0.05 * 300.0


def test_orders_agree_statistically(oracle):
    # the device order (coloured Gauss-Seidel + Jacobi overflow, red-black sweeps) settles the drop scene like the
    # reference's sequential order: same pieces, resting heights within a few cm
    out = []
    for order in (0, 1):
        O = _world(oracle, order)
        lc.build(O, None, dict(kind="drop", seed=5, nbody=12))
        for f in range(240):
            O.step()
        b = O.bodies()
        out.append(b[b[:, 0] > 0])
    a, c = out
    assert len(a) == len(c)
    assert abs(np.sort(a[:, 5]).mean() - np.sort(c[:, 5]).mean()) < 0.1

"""Pins of the lattice oracle's substep pieces (oracle/lattice.c), each run on its own through the oracle's test hooks
and checked against what geometry and mechanics fix, with answers worked out by hand below rather than by the
oracle's formulas:

  gen_contacts (src/lib.rs:1405-1471)  a brick's sphere (r 0.27) against a static cube's face, edge and corner, a
                                       centre inside the cube (axis rule), sphere against sphere, an asleep partner
  prep_contacts (:1551-1577)           the bias: 0.2/h (depth - 0.01) capped at 1, raised to -0.12 vn when vn < -2
  solve (:1578-1601)                   one resting contact: jn = m (-vn + bias), the body leaves at exactly the bias;
                                       friction clamped to +-0.65 jn per tangent axis (a box, not a disk)
  momentum                             two bodies colliding off the ground keep their total linear momentum
  pinch_check (:1746-1768)             velocity change limited to 8 m/s (small pieces killed), rim spin change to 6 m/s
  structural g - a (:1284-1308)        free fall loads nothing (ext = -MASS, field 0); a stop at n g loads
                                       MASS (n - 1); n capped at 150
  crush_at (:1643-1702)                energy E crushes floor(E / EB) bricks nearest first; under 0.15 EB nothing
"""
import math

import numpy as np
import pytest

C = 320          # lattice index of x = 0 and z = 0
RB, H = 0.27, 0.25
MASS = [0.0, 300.0, 230.0, 250.0, 60.0, 70.0]
EB = [0.0, 15000.0, 12000.0, 1.5e6, 300.0, 6000.0]
HSUB = 1.0 / 120.0


def _col(n=3, mat=1):
    return np.full(n, mat, np.uint8), np.stack([np.full(n, C), np.arange(n), np.full(n, C)], 1).astype(np.int32)


def _world(oracle, n_col=3, col_mat=1, bodies=((2, (C + 1, 0, C)),)):
    """a static column at x = z = 0.25 (cube k spans y in [0.5 k, 0.5 k + 0.5]) and single-brick bodies"""
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    if n_col:
        O.add_structure(*_col(n_col, col_mat))
    ids = [O.add_body(np.array([m], np.uint8), np.array([cell], np.int32)) for m, cell in bodies]
    O.finalize()
    return O, ids


def _one(cs):
    assert len(cs) == 1, cs
    c = cs[0]
    return dict(a=int(c[0]), b=int(c[1]), bo=int(c[2]), n=c[7:10].astype(np.float64), p=c[10:13].astype(np.float64),
                depth=float(c[13]), jn=float(c[14]), jt1=float(c[15]), jt2=float(c[16]), vn0=float(c[17]))


# ---------------------------------------------------------------- contact geometry
def test_sphere_on_cube_face(oracle):
    O, (b,) = _world(oracle)
    # 0.2 m off the +x face of the top cube (face at x = 0.5, y in [1, 1.5]); the cubes below are out of reach
    O.set_body(b, pos=(0.7, 1.3, 0.35), v=(0, 0, 0))
    c = _one(O.probe_contacts())
    assert c["b"] == -1 and c["bo"] == -1
    assert np.allclose(c["n"], [1, 0, 0], atol=1e-7)
    assert abs(c["depth"] - (RB - 0.2)) < 1e-6
    assert np.allclose(c["p"], [0.5, 1.3, 0.35], atol=1e-6)


def test_sphere_on_cube_edge(oracle):
    O, (b,) = _world(oracle)
    # over the top cube's +x top edge (x = 0.5, y = 1.5): 0.1 m out along x and y
    O.set_body(b, pos=(0.6, 1.6, 0.35))
    c = _one(O.probe_contacts())
    r = math.hypot(0.1, 0.1)
    assert np.allclose(c["n"], [1 / math.sqrt(2), 1 / math.sqrt(2), 0], atol=2e-6)
    assert abs(c["depth"] - (RB - r)) < 2e-6
    assert np.allclose(c["p"], [0.5, 1.5, 0.35], atol=1e-6)


def test_sphere_on_cube_corner(oracle):
    O, (b,) = _world(oracle)
    O.set_body(b, pos=(0.6, 1.6, 0.6))
    c = _one(O.probe_contacts())
    r = math.sqrt(3) * 0.1
    assert np.allclose(c["n"], np.ones(3) / math.sqrt(3), atol=2e-6)
    assert abs(c["depth"] - (RB - r)) < 2e-6
    assert np.allclose(c["p"], [0.5, 1.5, 0.5], atol=1e-6)


def test_centre_inside_cube_takes_the_largest_axis(oracle):
    O, (b,) = _world(oracle)
    # inside the top cube (centre (0.25, 1.25, 0.25)), offset (+0.05, -0.1, +0.02): the closest point is the centre
    # itself, so the normal is the axis of the largest offset (-y) and the depth H + RB - 0.1; the cube below is a
    # face contact 0.15 m away
    O.set_body(b, pos=(0.30, 1.15, 0.27))
    cs = O.probe_contacts()
    assert len(cs) == 2
    below, inside = cs[0], cs[1]     # cells visited dy = -1 first
    assert np.allclose(below[7:10], [0, 1, 0]) and abs(below[13] - (RB - 0.15)) < 1e-6
    assert np.allclose(inside[7:10], [0, -1, 0]) and abs(inside[13] - (H + RB - 0.1)) < 1e-6
    assert np.allclose(inside[10:13], [0.30, 1.15, 0.27], atol=1e-7)


@pytest.mark.parametrize("d", [0.3, 0.5])
def test_sphere_against_sphere(oracle, d):
    O, (a, b) = _world(oracle, n_col=0, bodies=((2, (C, 10, C)), (1, (C + 1, 10, C))))
    u = np.array([2.0, 1.0, -2.0]) / 3.0
    pa = np.array([3.0, 5.0, 3.0])
    pb = pa - d * u
    O.set_body(a, pos=pa)
    O.set_body(b, pos=pb)
    c = _one(O.probe_contacts())
    # reported once, from the lower brick index (body a's brick), with the normal from b to a
    assert (c["a"], c["b"], c["bo"]) == (a, b, b)
    assert np.allclose(c["n"], u, atol=2e-6)
    assert abs(c["depth"] - (2 * RB - d)) < 2e-6
    assert np.allclose(c["p"], (pa + pb) / 2, atol=2e-6)


def test_sphere_out_of_reach_and_asleep_partner(oracle):
    O, (a, b) = _world(oracle, n_col=0, bodies=((2, (C, 10, C)), (1, (C + 1, 10, C))))
    O.set_body(a, pos=(3.0, 5.0, 3.0))
    O.set_body(b, pos=(3.0 + 2 * RB + 1e-3, 5.0, 3.0))
    assert len(O.probe_contacts()) == 0
    # partner asleep: the awake body reports it as static-like (b = -1) with bo naming the sleeper; the normal still
    # points from the sleeper to the awake brick
    O.set_body(b, pos=(3.4, 5.0, 3.0))
    O.set_asleep(b)
    c = _one(O.probe_contacts())
    assert (c["a"], c["b"], c["bo"]) == (a, -1, b)
    assert np.allclose(c["n"], [-1, 0, 0], atol=5e-6) and abs(c["depth"] - (2 * RB - 0.4)) < 2e-6


# ---------------------------------------------------------------- prep and the impulse solve
def _resting(oracle, vy, vx=0.0, mat=2):
    # the brick's centre 0.25 m above the top face (y 1.5): depth RB - 0.25 = 0.02, normal +y, arm (0, -0.25, 0)
    O, (b,) = _world(oracle, bodies=((mat, (C, 4, C)),))
    O.set_body(b, pos=(0.25, 1.75, 0.25), v=(vx, vy, 0.0), w=(0, 0, 0))
    c = _one(O.probe_contacts())
    assert np.allclose(c["n"], [0, 1, 0]) and abs(c["depth"] - 0.02) < 1e-6
    O.probe_prep(HSUB)
    return O, b


@pytest.mark.parametrize("vy,bias", [(-1.0, 0.24), (-3.0, 0.36), (-20.0, 2.4)])
def test_single_contact_impulse_closed_form(oracle, vy, bias):
    # bias = min(0.2/h * (0.02 - 0.01), 1) = 0.24 at h = 1/120; vn < -2 raises it to -0.12 vn (0.36 at -3 m/s; 2.4
    # at -20 m/s, above the cap of 1 because the cap applies before the raise)
    m = MASS[2]
    O, b = _resting(oracle, vy)
    c = _one(O.contacts())
    assert abs(c["vn0"] - (-vy)) < 1e-6
    O.probe_solve(1)
    c = _one(O.contacts())
    # the arm is parallel to the normal, so the effective mass is the body mass
    assert abs(c["jn"] - m * (-vy + bias)) < 1e-5 * m * (-vy + bias)
    v = O.bodies()[b, 11:14]
    assert abs(v[1] - bias) < 1e-5 and abs(v[0]) < 1e-7 and abs(v[2]) < 1e-7


def test_friction_is_clamped_per_axis(oracle):
    m = MASS[2]
    # prep with slip along +x: t1 = (1, 0, 0) (the slip direction), t2 = n x t1 = (0, 0, -1)
    O, b = _resting(oracle, -3.0, vx=8.0)
    # the body then also slides along -z (t2 = +8 m/s), so both tangent rows want more than the limit
    O.set_body(b, v=(8.0, -3.0, -8.0))
    O.probe_solve(1)
    c = _one(O.contacts())
    jn = m * (3.0 + 0.36)
    assert abs(c["jn"] - jn) < 1e-5 * jn
    # single brick: I = m S^2 / 6, tangential effective mass m / (1 + 0.25^2 * 6 / S^2) = 0.4 m; the unclamped
    # impulse 0.4 m * 8 = 3.2 m exceeds 0.65 jn = 2.184 m on both axes, so each is clamped on its own (a disk clamp
    # would give 0.65 jn / sqrt(2) per axis)
    assert abs(c["jt1"] + 0.65 * jn) < 1e-5 * jn
    assert abs(c["jt2"] + 0.65 * jn) < 1e-5 * jn


def test_friction_below_the_limit_stops_the_slip_at_the_contact(oracle):
    m = MASS[2]
    O, b = _resting(oracle, -3.0, vx=1.0)
    O.probe_solve(1)
    c = _one(O.contacts())
    # unclamped: jt1 = -0.4 m * 1; the contact point's tangential velocity is then zero
    assert abs(c["jt1"] + 0.4 * m) < 1e-5 * m and c["jt2"] == 0.0
    bod = O.bodies()[b]
    v, w = bod[11:14].astype(np.float64), bod[14:17].astype(np.float64)
    vp = v + np.cross(w, [0.0, -0.25, 0.0])
    assert abs(vp[0]) < 1e-5


def test_momentum_conserved_between_two_bodies(oracle):
    O, (a, b) = _world(oracle, n_col=0, bodies=((1, (C, 10, C)), (4, (C + 1, 10, C))))
    u = np.array([1.0, 2.0, 2.0]) / 3.0
    pa = np.array([2.0, 6.0, -1.0])
    O.set_body(a, pos=pa, v=(-1.5, 0.3, -2.0), w=(0.5, -1.0, 2.0))
    O.set_body(b, pos=pa - 0.45 * u, v=(2.0, 1.0, 1.0), w=(-3.0, 0.2, 0.0))
    assert len(O.probe_contacts()) == 1
    before = O.bodies()
    O.probe_prep(HSUB)
    O.probe_solve(10)
    after = O.bodies()
    p0 = before[a, 17] * before[a, 11:14].astype(np.float64) + before[b, 17] * before[b, 11:14].astype(np.float64)
    p1 = after[a, 17] * after[a, 11:14].astype(np.float64) + after[b, 17] * after[b, 11:14].astype(np.float64)
    assert not np.allclose(after[a, 11:14], before[a, 11:14])   # the impulse did something
    assert np.abs(p1 - p0).max() < 1e-5 * MASS[1] * 3


# ---------------------------------------------------------------- pinch
def test_pinch_kills_a_small_piece_and_limits_a_large_one(oracle):
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    small = O.add_body(np.full(2, 2, np.uint8), np.array([[C, 10, C], [C + 1, 10, C]], np.int32))
    big = O.add_body(np.full(8, 2, np.uint8), np.array([[C + 4 + x, 10 + y, C + z] for x in range(2) for y in range(2)
                                                       for z in range(2)], np.int32))
    O.finalize()
    O.set_body(small, v=(9.0, 0, 0), w=(0, 0, 0))
    O.set_body(big, v=(0, 0, 12.0), w=(0, 0, 0))
    O.probe_pinch(small, (0, 0, 0), (0, 0, 0))
    O.probe_pinch(big, (0, 0, 0), (0, 0, 0))
    bod = O.bodies()
    assert bod[small, 0] == 0                  # 9 m/s change on a 2-brick piece: killed
    assert bod[big, 0] == 1
    assert np.allclose(bod[big, 11:14], [0, 0, 8.0], atol=1e-5)   # 12 m/s change limited to 8


def test_pinch_limits_rim_spin_and_leaves_small_changes(oracle):
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    b = O.add_body(np.array([2], np.uint8), np.array([[C, 10, C]], np.int32))
    O.finalize()
    # single brick: rad = 0 + 1.8 H = 0.45; a spin change of 20 rad/s is a 9 m/s rim speed, limited to 6 m/s
    O.set_body(b, v=(7.9, 0, 0), w=(0, 0, 21.0))
    O.probe_pinch(b, (0, 0, 0), (0, 0, 1.0))
    bod = O.bodies()[b]
    assert bod[0] == 1
    assert np.allclose(bod[11:14], [7.9, 0, 0])
    assert np.allclose(bod[14:17], [0, 0, 1.0 + 20.0 * 6.0 / 9.0], atol=1e-5)


# ---------------------------------------------------------------- g - a loading
def _cube8(oracle, mat=1):
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    cells = np.array([[C + x, 20 + y, C + z] for x in range(2) for y in range(2) for z in range(2)], np.int32)
    b = O.add_body(np.full(8, mat, np.uint8), cells)
    O.finalize()
    O.set_nc(b, 1)                             # in contact with something: its bricks carry load
    return O, b


def test_free_fall_loads_nothing(oracle):
    dt = 1.0 / 60.0
    O, b = _cube8(oracle)
    O.set_body(b, v=(0.5, -4.0, 0.0), w=(0, 0, 0))
    O.set_last(b, (0.5, -4.0 + 9.81 * dt, 0.0), (0, 0, 0))
    O.set_nc(b, 1)
    O.probe_structural(dt)
    br = O.bricks()
    assert np.allclose(br[:, 6], -MASS[1], rtol=1e-3)     # ext = MASS (0 - 1)
    assert np.abs(br[:, 5]).max() < 1e-3                   # source MASS + ext = 0: no field
    assert abs(O.bodies()[b, 21]) < 1e-3                   # nmax = 0


@pytest.mark.parametrize("n_g,expect", [(50.0, 50.0), (400.0, 150.0)])
def test_impact_loads_n_g_capped_at_150(oracle, n_g, expect):
    dt = 1.0 / 60.0
    O, b = _cube8(oracle)
    # stopped from falling: the velocity rose by (n - 1) g dt in one frame, so g - a = n g upward... downward
    O.set_body(b, v=(0, 0, 0), w=(0, 0, 0))
    O.set_last(b, (0, -(n_g - 1.0) * 9.81 * dt, 0), (0, 0, 0))
    O.set_nc(b, 1)
    O.probe_structural(dt)
    br = O.bricks()
    assert np.allclose(br[:, 6], MASS[1] * (expect - 1.0), rtol=1e-4)
    assert abs(O.bodies()[b, 21] - expect) < 1e-3 * expect


# ---------------------------------------------------------------- crush
@pytest.mark.parametrize("k", [1, 3, 7])
def test_crush_spends_the_energy_brick_by_brick(oracle, k):
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    n = 20
    O.add_structure(*_col(n, 2))
    O.finalize()
    start = 10
    p = (0.25, start * 0.5 + 0.25, 0.25)
    # (k + 1/2) EB crushes k bricks; the half left over cracks the next one and stops
    got = O.probe_crush(start, (k + 0.5) * EB[2], p)
    assert got == k
    br = O.bricks()
    bond_y = br[:, 8]
    # a crushed brick is removed (its pulverised share) or severed: its own +y bond and the one from below are cut
    crushed = {i for i in range(1, n - 1) if br[i, 0] == 0 or (bond_y[i] == 0 and bond_y[i - 1] == 0)}
    # nearest first from the contact point in the middle of brick 10: the k bricks closest to it along the column
    assert crushed == set(range(start - (k - 1) // 2, start + (k - 1) // 2 + 1))


@pytest.mark.parametrize("k", [3, 4])
def test_crush_grows_nearest_first_around_an_off_centre_hit(oracle, k):
    # a 9 x 9 masonry wall in the x-y plane, hit 0.2 m right of and above the middle brick's centre: the crushed
    # bricks are the k nearest to the hit point (the breadth-first queue alone would take the -x neighbour second)
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    cells = np.array([[C + x, y, C] for y in range(9) for x in range(9)], np.int32)
    O.add_structure(np.full(len(cells), 2, np.uint8), cells)
    O.finalize()
    centre = np.array([0.25 + 4 * 0.5, 4 * 0.5 + 0.25, 0.25])
    hit = centre + [0.2, 0.2, 0.0]
    assert O.probe_crush(4 * 9 + 4, (k + 0.5) * EB[2], hit) == k
    br = O.bricks()
    bx, by = br[:, 7].reshape(9, 9), br[:, 8].reshape(9, 9)      # +x and +y bonds, [y, x]
    gone = br[:, 0].reshape(9, 9) == 0
    crushed = {(x, y) for y in range(1, 8) for x in range(1, 8)
               if gone[y, x] or (bx[y, x] == 0 and by[y, x] == 0 and bx[y, x - 1] == 0 and by[y - 1, x] == 0)}
    pos = {(x, y): np.array([0.25 + x * 0.5, y * 0.5 + 0.25, 0.25]) for y in range(9) for x in range(9)}
    nearest = sorted(pos, key=lambda c: float(np.linalg.norm(pos[c] - hit)))[:k]
    assert crushed == set(nearest)


def test_crush_under_the_threshold_does_nothing(oracle):
    O = oracle.LatticeOracle(order=0, brick_cap=256, body_cap=64)
    O.add_structure(*_col(6, 2))
    O.finalize()
    before = O.bricks().copy()
    assert O.probe_crush(3, 0.1 * EB[2], (0.25, 1.75, 0.25)) == 0
    assert np.array_equal(before, O.bricks())

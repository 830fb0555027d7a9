"""The spring-network structural analysis oracle (oracle/structure.c: engine/src/structure.rs:191-783 over the brick
store of engine/src/brick.rs) pinned by the reference's own validation against beam theory and axial stress
(engine/src/validation.rs:1-102, docs/VALIDATION.md:3-14) and its analysis tests (structure.rs:790-879), plus the
mathematics the network fixes: the solve equals a dense solve of the same stiffness, every cell is in force and moment
equilibrium, the joints are every face two bricks share.

A plausible mistake fails one of these: a section modulus off by a factor (the cantilevers' 3 rho g L^2 / h), a lever
arm from the wrong cell (moment equilibrium), a wrong sign in the rigid map (the dense solve), a support cut left out
(the spans' bracket), a joint missed or its area wrong (the brute-force joints), compression checked against tension
(the columns).
"""
import itertools

import numpy as np
import pytest

from oracle import pyoracle as po

C, M, S, G, T = 0, 1, 2, 3, 4
DENS = {C: 2400.0, M: 1900.0, S: 7850.0, T: 500.0, G: 2500.0}
FLEX = {C: 12e6, M: 0.5e6, T: 20e6}
STRENGTH_C = {C: 30e6, M: 10e6, T: 21e6}
GRAV = 9.81


class Store:
    """A brick store and the min corners it was given (the supports are the bricks resting on y = 0)."""

    def __init__(self):
        self.b = po.BricksOracle()
        self.mn = []
        self.size = []

    def add(self, mn, size, m, paint=0, structure=0):
        self.mn.append(list(mn))
        self.size.append(list(size))
        return self.b.add(mn, size, m, paint, structure)

    def run(self, extra=(), order=po.ORDER_REFERENCE, point=(), warm=None, dir3=(0.0, -1.0, 0.0)):
        """validation.rs run(): weights plus extra loads, every brick on the ground a support"""
        n = len(self.mn)
        load = np.array([self.b.mass(i) * np.float32(GRAV) for i in range(n)], np.float32)
        for i, f in extra:
            load[i] += np.float32(f)
        sink = np.array([mn[1] <= 0 for mn in self.mn], np.uint8)
        return po.analyse_oracle(self.b, np.arange(n, dtype=np.uint32), load, sink, dir3=dir3, point=point,
                                 order=order, warm=warm)


def _cantilever(m, length, h):
    s = Store()
    s.add([0, 0, 0], [8, 8, 8], m)
    s.add([0, 8, 0], [8 + length, h, 8], m)
    s.b.build_joints()
    a = s.run()
    l, d = length * 0.125, h * 0.125
    return a, 3.0 * DENS[m] * GRAV * l * l / d / FLEX[m]


@pytest.mark.parametrize("m,length,h", [(C, 24, 2), (C, 40, 2), (C, 56, 2), (C, 40, 4), (T, 40, 2)])
def test_cantilever_matches_beam_theory(m, length, h):
    # validation.rs:34-45: sigma = 3 rho g L^2 / h against the flexural strength; the reference reports 0.0 % error
    a, theory = _cantilever(m, length, h)
    assert abs(a["max_ratio"] - theory) / theory < 0.005, (a["max_ratio"], theory)


def _span(m, span, h, bear):
    s = Store()
    s.add([0, 0, 0], [bear, 8, 8], m)
    s.add([bear + span, 0, 0], [bear, 8, 8], m)
    s.add([0, 8, 0], [2 * bear + span, h, 8], m)
    s.b.build_joints()
    a = s.run()
    c, bw, d = span * 0.0625, bear * 0.125, h * 0.125
    l = 2.0 * c + bw
    w = DENS[m] * GRAV * d
    return a, w * l * l / 12.0 * 6.0 / (d * d) / FLEX[m]


@pytest.mark.parametrize("m,span,h", [(C, 48, 2), (C, 96, 2), (T, 96, 4)])
def test_span_between_clamped_and_simply_supported(m, span, h):
    # validation.rs:46-62, :95-101: between 0.6 of the clamped end moment and 1.6 of it (the simply supported midspan
    # moment with the bearings' width)
    a, clamped = _span(m, span, h, 2)
    assert 0.6 * clamped < a["max_ratio"] < 1.6 * clamped, (a["max_ratio"], clamped)


@pytest.mark.parametrize("m,p", [(C, 5.0e6), (M, 2.0e6), (T, 1.0e6)])
def test_column_axial_stress(m, p):
    # validation.rs:63-76: sigma = (P + W) / A against the compressive strength; the reference reports -0.4 to -0.7 %
    s = Store()
    s.add([0, 0, 0], [4, 24, 4], m)
    cap = s.add([0, 24, 0], [4, 2, 4], m)
    s.b.build_joints()
    a = s.run(extra=[(cap, p)])
    w = (s.b.mass(0) + s.b.mass(cap)) * GRAV
    theory = (p + w) / (0.5 * 0.5) / STRENGTH_C[m]
    err = (a["max_ratio"] - theory) / theory
    assert -0.01 < err < 0.0, (a["max_ratio"], theory)


@pytest.mark.parametrize("length,fails", [(60, True), (16, False)])
def test_cantilever_as_one_brick(length, fails):
    # structure.rs:818-830: a 7.5 m concrete strip cracks or fails, a 2 m one holds
    s = Store()
    s.add([0, 0, 0], [8, 8, 8], C)
    s.add([0, 8, 0], [8 + length, 2, 8], C)
    s.b.build_joints()
    a = s.run()
    over = (a["joint_ratios"] > 1.0).any() or (a["inner_ratios"] > 1.0).any()
    assert over == fails, a["max_ratio"]


@pytest.mark.parametrize("span,fails", [(60, False), (180, True)])
def test_simply_supported_span(span, fails):
    # structure.rs:831-846: a slab strip bonded to two blocks holds over 7.5 m and not over 22.5 m
    s = Store()
    s.add([0, 0, 0], [8, 8, 8], C)
    s.add([8 + span, 0, 0], [8, 8, 8], C)
    s.add([0, 8, 0], [16 + span, 2, 8], C)
    s.b.build_joints()
    assert (s.run()["max_ratio"] > 1.0) == fails


@pytest.mark.parametrize("off,tips", [(0, True), (20, False)])
def test_slab_on_one_column_tips_off_centre(off, tips):
    # structure.rs:847-865: a 6 x 6 m slab on one 0.5 m column tips with the column at its corner, stands with the
    # column near its centre
    s = Store()
    s.add([0, 0, 0], [80, 2, 80], C)
    s.add([off, 2, off], [4, 24, 4], C, paint=1)
    s.add([0, 26, 0], [48, 2, 48], C)
    s.b.build_joints()
    a = s.run()
    assert bool((a["joint_ratios"] > 1.0).any()) == tips, a["max_ratio"]


def _rigid(d):
    # a point at offset d of a body moving by (u, theta) moves by u + theta x d
    t = np.eye(6)
    for i in range(3):
        e = np.zeros(3)
        e[i] = 1.0
        t[:3, 3 + i] = np.cross(e, d)
    return t


def _stiffness(a):
    n = a["nodes"]
    K = np.zeros((6 * n, 6 * n))
    for (ia, ib), sp in zip(a["ends"], a["springs"]):
        D = np.diag(np.concatenate([sp[0:3], sp[3:6]]))
        Ba, Bb = _rigid(sp[6:9]), _rigid(sp[9:12])
        sa, sb = slice(6 * ia, 6 * ia + 6), slice(6 * ib, 6 * ib + 6)
        K[sa, sa] += Ba.T @ D @ Ba
        K[sb, sb] += Bb.T @ D @ Bb
        K[sa, sb] -= Ba.T @ D @ Bb
        K[sb, sa] -= Bb.T @ D @ Ba
    return K


def _frame():
    s = Store()
    # two columns, a beam across them and a short cantilever beyond, a block on top
    s.add([0, 0, 0], [4, 24, 4], C, paint=1)
    s.add([36, 0, 0], [4, 24, 4], C, paint=1)
    s.add([0, 24, 0], [56, 3, 4], C)
    s.add([16, 27, 0], [8, 8, 4], M)
    s.b.build_joints()
    return s


def test_solve_equals_a_dense_solve_of_the_same_network():
    s = _frame()
    a = s.run()
    K = _stiffness(a)
    free = ~a["fixed"]
    idx = np.concatenate([np.arange(6 * k, 6 * k + 6) for k in np.nonzero(free)[0]])
    f = np.zeros(6 * a["nodes"])
    f[1::6] = -a["node_load"].astype(np.float64)
    Kff = K[np.ix_(idx, idx)]
    # the solve's own small diagonal addition (1e-7 of a cell's largest diagonal entry) is far below the tolerance
    x = np.linalg.solve(Kff + np.diag(1e-7 * np.maximum(np.abs(np.diag(Kff)), 1.0)), f[idx])
    got = a["x"].reshape(-1)[idx]
    assert np.abs(got - x).max() < 1e-3 * np.abs(x).max()
    assert a["iterations"] > 0 and np.abs(a["x"][a["fixed"]]).max() == 0.0


def test_every_cell_in_force_and_moment_equilibrium():
    s = _frame()
    a = s.run()
    n = a["nodes"]
    K = _stiffness(a)
    y = (K @ a["x"].reshape(-1)).reshape(n, 6)
    load = a["node_load"].astype(np.float64)
    free = ~a["fixed"]
    carried = load[free].sum()
    # free cells: the springs balance the weight (to the solve's tolerance)
    res = y[free].copy()
    res[:, 1] += load[free]
    assert np.abs(res).max() < 2e-3 * load.max()
    # the internal forces cancel, so the supports take the whole weight of what they carry; and the moments about
    # the origin cancel too, which holds only if both arms of every spring end at the same face centre
    centre = 0.5 * (a["node"][:, 1:4] + a["node"][:, 4:7]) * 0.125
    assert np.abs(y[:, :3].sum(0)).max() < 1e-9 * carried
    assert abs(y[a["fixed"], 1].sum() - carried) < 2e-3 * carried
    mom = y[:, 3:] + np.cross(centre, y[:, :3])
    assert np.abs(mom.sum(0)).max() < 1e-9 * carried * 10.0


def test_device_order_agrees_with_the_reference_order():
    s = _frame()
    a = s.run(order=po.ORDER_REFERENCE)
    b = s.run(order=po.ORDER_DEVICE)
    assert np.array_equal(a["joint_ids"], b["joint_ids"])
    assert np.abs(a["x"] - b["x"]).max() < 1e-9 * np.abs(a["x"]).max()
    assert abs(a["max_ratio"] - b["max_ratio"]) < 1e-6 * a["max_ratio"]


def test_warm_start_keeps_the_answer_in_no_iterations():
    s = _frame()
    w = po.WarmOracle()
    a = s.run(warm=w)
    assert len(w) == int((~a["fixed"]).sum())
    b = s.run(warm=w)
    assert b["iterations"] == 0
    assert np.array_equal(a["joint_ratios"], b["joint_ratios"])


def test_point_load_goes_to_the_cell_under_it():
    # structure.rs:181-186: a point load on one bay of a slab loads that bay's cell, not the whole slab
    s = Store()
    s.add([0, 0, 0], [8, 8, 8], C)
    s.add([0, 8, 0], [48, 2, 8], C)
    s.b.build_joints()
    a = s.run(point=[(1, 44.0, 4.0, 20000.0)])
    b = s.run()
    cells = a["node"][a["node"][:, 0] == 1]
    hit = (cells[:, 1] <= 44) & (cells[:, 4] > 44)
    diff = a["node_load"][a["node"][:, 0] == 1] - b["node_load"][b["node"][:, 0] == 1]
    assert hit.sum() == 1 and abs(diff[hit][0] - 20000.0) < 1e-2 and np.abs(diff[~hit]).max() == 0.0
    assert a["max_ratio"] > b["max_ratio"]


def test_cells_cut_at_two_metres_and_at_the_support_edges():
    # structure.rs:201-229: a 12 m slab on a 1 m block is cut every 2 m (and at most 16 ticks a cell) and at the
    # block's edge
    s = Store()
    s.add([20, 0, 0], [8, 8, 8], C)
    s.add([0, 8, 0], [96, 2, 8], C)
    s.b.build_joints()
    a = s.run()
    cells = a["node"][a["node"][:, 0] == 1]
    xs = sorted(set(cells[:, 1].tolist()) | set(cells[:, 4].tolist()))
    assert 20 in xs and 28 in xs and xs[0] == 0 and xs[-1] == 96
    assert max(np.diff(xs)) <= 16


def test_joints_are_every_shared_face():
    # brute force: every pair of boxes that share a face of positive area, on the face's axis, with its area in
    # tick^2, the min-side brick first
    rng = np.random.default_rng(5)
    s = Store()
    occ = set()
    for _ in range(220):
        size = [int(v) for v in rng.integers(1, 5, 3)]
        mn = [int(v) for v in rng.integers(0, 12, 3)]
        cells = set(itertools.product(*[range(mn[k], mn[k] + size[k]) for k in range(3)]))
        if cells & occ:
            continue
        occ |= cells
        s.add(mn, size, C)
    n = s.b.build_joints()
    want = set()
    boxes = [(np.array(mn), np.array(mn) + np.array(sz)) for mn, sz in zip(s.mn, s.size)]
    for i, j in itertools.combinations(range(len(boxes)), 2):
        (pa, pb), (qa, qb) = boxes[i], boxes[j]
        for ax in range(3):
            lo = pb[ax] == qa[ax]
            hi = qb[ax] == pa[ax]
            if not (lo or hi):
                continue
            o = [k for k in range(3) if k != ax]
            ext = [min(pb[k], qb[k]) - max(pa[k], qa[k]) for k in o]
            if ext[0] > 0 and ext[1] > 0:
                want.add((i, j, ax, ext[0] * ext[1]) if lo else (j, i, ax, ext[0] * ext[1]))
            break
    got = {tuple(int(v) for v in r[:4]) for r in s.b.joints() if r[0] >= 0}
    assert got == want and n == len(want) and len(want) > 50


def test_removed_brick_takes_its_joints_with_it():
    s = Store()
    for k in range(4):
        s.add([0, 8 * k, 0], [8, 8, 8], C)
    s.b.build_joints()
    s.b.remove(1)
    live = [tuple(r[:2]) for r in s.b.joints() if r[0] >= 0]
    assert live == [(2, 3)]

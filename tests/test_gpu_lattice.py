"""Lattice world (the legacy brick engine's frame) on the GPU against the oracle (oracle/lattice.c) in the device's
orders, through the C ABI.

The bar: every body (alive, asleep, brick count, contact count, position, rotation, velocities, mass, rest timer,
joint ratio, load-field residual, g - a factor), every brick (material, body, position, load potential, external
load, bonds, stress, support age), the frame statistics and the random stream equal bit for bit, frame after frame:
the kernels perform the oracle's f32 operations in the same order and the host bookkeeping follows it.
"""
import numpy as np
import pytest

import lattice_cases as lc
from shatterforge_amd.lattice import Lattice

pytestmark = pytest.mark.gpu
DT = np.float32(1.0 / 60.0)


def _first_diff(a, b):
    d = np.argwhere(~((a == b) | (np.isnan(a) & np.isnan(b))))
    return d[:5].tolist()


def _assert_same(G, s, O, tag):
    gb, ob = G.bodies(s), O.bodies()
    assert gb.shape == ob.shape, f"{tag}: body slots {gb.shape} vs {ob.shape}"
    assert np.array_equal(gb, ob, equal_nan=True), f"{tag}: bodies differ at {_first_diff(gb, ob)}"
    gk, ok = G.bricks(s), O.bricks()
    assert np.array_equal(gk, ok, equal_nan=True), f"{tag}: bricks differ at {_first_diff(gk, ok)}"
    gs, os_ = G.stats(s), O.stats()
    assert np.array_equal(gs, os_), f"{tag}: stats {gs} vs {os_}"
    assert G.rng(s) == O.rng(), f"{tag}: random stream"


def _run(oracle, case, frames, check_every=1, fused=True, sor_lds=-1, conn_device=True):
    nb = lc.brick_count(case) + 8
    nd = 4096
    G = Lattice(nb, nd, scenes=1)
    if not fused:
        G.set_fused(0)
    G.set_conn_device(conn_device)
    G.set_sor_lds(sor_lds)
    s = G.add_scene(nb, nd, structs=4)
    O = oracle.LatticeOracle(order=oracle.ORDER_DEV, brick_cap=nb, body_cap=nd)
    lc.build(G, s, case)
    lc.build(O, None, case)
    G.step(DT)
    O.step(DT)
    _assert_same(G, s, O, "warm-up frame")
    for (x, y, z, r, p) in lc.charges(case):
        G.explode(s, x, y, z, r, p)
        O.explode(x, y, z, r, p)
    _assert_same(G, s, O, "after charges")
    for f in range(frames):
        G.step(DT)
        O.step(DT)
        if f % check_every == 0 or f == frames - 1:
            _assert_same(G, s, O, f"frame {f}")
    return G, s, O


def test_charges_among_bodies_bit_exact(oracle):
    """Charges set off after the tower has come apart: the blast reaches bricks of moving bodies (removed, their
    bonds damaged, the bodies split at the next rebuild) and pushes every body near it, beside the standing
    structure's bricks; the dropped bodies of the drop case take a charge of their own."""
    print("case: 8x8x16 tower, 4 charges, 60 frames, then charges at its two largest bodies, 60 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="tower", w=8, h=16), 60, check_every=10)
    ob = O.bodies()
    live = np.flatnonzero(ob[:, 0] > 0)
    assert live.size >= 2
    big = live[np.argsort(-ob[live, 2], kind="stable")[:2]]
    for k in big:
        x, y, z = (float(v) for v in ob[k, 4:7])
        G.explode(s, x, y, z, 2.0, 1.2)
        O.explode(x, y, z, 2.0, 1.2)
        _assert_same(G, s, O, f"after the charge at body {k}")
    for f in range(60):
        G.step(DT)
        O.step(DT)
        if f % 5 == 0 or f == 59:
            _assert_same(G, s, O, f"frame {f} after the body charges")
    print("case: 24 loose bodies dropped around a column, 30 frames, charges at the two largest, 30 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="drop", seed=3, nbody=24), 30, check_every=10)
    ob, bk0 = O.bodies(), O.bricks()
    live = np.flatnonzero(ob[:, 0] > 0)
    for k in live[np.argsort(-ob[live, 2], kind="stable")[:2]]:
        x, y, z = (float(v) for v in ob[k, 4:7])
        G.explode(s, x, y, z, 1.5, 1.0)
        O.explode(x, y, z, 1.5, 1.0)
        _assert_same(G, s, O, f"after the charge at dropped body {k}")
    bk1 = O.bricks()
    in_body = bk0[:, 1] >= 0
    assert ((bk0[:, 0] != 0) & (bk1[:, 0] == 0) & in_body).sum() > 0      # bricks of bodies blown away
    assert ((bk0 != bk1).any(1) & in_body & (bk1[:, 0] != 0)).sum() > 0  # and bonds inside bodies damaged
    most = 0
    for f in range(30):
        G.step(DT)
        O.step(DT)
        _assert_same(G, s, O, f"frame {f} after the charges at the dropped bodies")
        most = max(most, int((O.bodies()[:, 0] > 0).sum()))
    assert most > live.size                                                # the damaged bodies split


def test_rams_plough_a_wall_bit_exact(oracle):
    """Two concrete rams thrown broadside at 20 m/s into a masonry wall: the small one ploughs bricks out of the wall
    (each a body of its own, thrown with a random push and spin, the ram slowed), the large one (over 40 bricks) only
    damages what it strikes, a brick already gone voids its contact; crushes on impact besides. Every frame equal to
    the oracle bit for bit, the random stream included."""
    print("case: masonry wall, a 2x2x2 and a 4x4x3 concrete ram at 20 m/s, 90 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="ram", speed=20.0), 90)
    assert (O.bodies()[:, 0] > 0).sum() > 2


def test_heavy_ram_long_crush_walks_bit_exact(oracle):
    """A steel ram at 40 m/s through a masonry wall: crush walks of over 300 bricks (the nearest of each 24-brick window,
    queues near 300 long), pulverised and severed bricks from the scene's random stream, the wall in pieces. Every
    frame equal to the oracle bit for bit."""
    print("case: 16x10x3 masonry wall, a 6x6x4 steel ram at 40 m/s, 60 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="heavy_ram"), 60)
    assert (O.bodies()[:, 0] > 0).sum() > 50


def test_drop_bodies_bit_exact(oracle):
    print("case: 24 loose bodies dropped around a column, 90 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="drop", seed=3, nbody=24), 90)
    st = O.stats()
    assert st[2] > 10 and st[4] > 0


@pytest.mark.parametrize("route", ["workgroup per scene", "colour launches"])
def test_tower_blast_bit_exact(oracle, route):
    # the contact solve as one workgroup per scene and as colour-by-colour launches: both equal the oracle
    print(f"case: 8x8x16 tower, 4 charges at the column bases, 120 frames, solve as {route}", flush=True)
    G, s, O = _run(oracle, dict(kind="tower", w=8, h=16, nch=4, r=2.2, p=1.4), 120, fused=route.startswith("work"))
    assert O.stats()[2] > 0


@pytest.mark.parametrize("search", ["device", "host"])
def test_connectivity_search_on_device_and_host_bit_exact(oracle, search):
    # connectivity's grounded search as the device's breadth-first search per structure (default) and as the host's
    # search: the detached parts, their bodies and everything after equal the oracle either way
    print(f"case: 2500-brick frame building, 4 charges r=3, 150 frames, grounded search on the {search}", flush=True)
    G, s, O = _run(oracle, dict(kind="building", n=2500, seed=7, nch=4, r=3.0, p=1.4), 150, check_every=5,
                   conn_device=search == "device")
    assert O.stats()[2] > 0


@pytest.mark.parametrize("sweeps", ["potential in LDS", "potential in global memory"])
def test_building_collapse_bit_exact(oracle, sweeps):
    # the structure sweeps from packed rows with the potential in LDS (structures up to the LDS limit) and from the
    # full rows with the potential in global memory (larger structures; forced here): both equal the oracle
    print(f"case: 2500-brick frame building, 4 charges r=3, 150 frames, sweeps with the {sweeps}", flush=True)
    G, s, O = _run(oracle, dict(kind="building", n=2500, seed=4, nch=4, r=3.0, p=1.4), 150, check_every=5,
                   sor_lds=-1 if sweeps.endswith("LDS") else 0)
    assert O.stats()[2] > 0


def test_office_collapse_bit_exact(oracle):
    # every fourth ground column cell of a six-storey concrete frame blown: the floors come down over the columns,
    # so large pieces run long static chains (hundreds of contacts against the standing columns) every substep
    print("case: six-storey office, 16 charges at the ground columns, 420 frames", flush=True)
    G, s, O = _run(oracle, dict(kind="office"), 420, check_every=20)
    assert O.stats()[2] > 100


def test_overlapping_structures_are_rejected():
    # two structures sharing a lattice cell: the neighbour table would be ambiguous, so finalize refuses the scene
    from shatterforge_amd._native import SfError
    G = Lattice(64, 64, scenes=1)
    s = G.add_scene(64, 64, structs=2)
    col = np.stack([np.full(4, lc.C), np.arange(4), np.full(4, lc.C)], 1).astype(np.int32)
    G.add_structure(s, np.full(4, 2, np.uint8), col)
    G.add_structure(s, np.full(2, 1, np.uint8), col[2:4] + [0, 0, 0])
    with pytest.raises(SfError, match="two bricks in lattice cell"):
        G.finalize(s)
    G.close()


MIX = [dict(kind="drop", seed=3, nbody=24), dict(kind="tower", w=8, h=16, nch=4, r=2.2, p=1.4),
       dict(kind="building", n=2500, seed=4, nch=4, r=3.0, p=1.4), dict(kind="drop", seed=11, nbody=30),
       dict(kind="ram", speed=20.0), dict(kind="heavy_ram")]


def test_batched_scenes_same_alone_and_in_mix(oracle):
    # six scenes of uneven size stepped together in one world: each equals the oracle run of that scene alone and
    # the GPU run of that scene alone, bit for bit (the batch solves one workgroup per scene, the lone worlds by
    # colour launches; two of the scenes plough and crush at once, each with its own random stream and slot cursor)
    print("case: 6 scenes batched, 100 frames, vs each alone (GPU and oracle)", flush=True)
    caps = [lc.brick_count(c) + 8 for c in MIX]
    G = Lattice(sum(caps), 4096 * len(MIX), scenes=len(MIX))
    sc = [G.add_scene(caps[k], 4096, structs=4, seed=0x9E3779B9) for k in range(len(MIX))]
    alone = []
    for k, c in enumerate(MIX):
        A = Lattice(caps[k], 4096, scenes=1)
        A.set_fused(0)
        a = A.add_scene(caps[k], 4096, structs=4)
        alone.append((A, a))
    Os = [oracle.LatticeOracle(order=oracle.ORDER_DEV, brick_cap=caps[k], body_cap=4096) for k in range(len(MIX))]
    for k, c in enumerate(MIX):
        lc.build(G, sc[k], c)
        lc.build(alone[k][0], alone[k][1], c)
        lc.build(Os[k], None, c)
    def step_all():
        G.step(DT)
        for A, _ in alone:
            A.step(DT)
        for O in Os:
            O.step(DT)
    step_all()
    for k, c in enumerate(MIX):
        for ch in lc.charges(c):
            G.explode(sc[k], *ch)
            alone[k][0].explode(alone[k][1], *ch)
            Os[k].explode(*ch)
    for f in range(100):
        step_all()
        if f % 10 == 9:
            for k in range(len(MIX)):
                _assert_same(G, sc[k], Os[k], f"frame {f} scene {k} vs oracle")
                A, a = alone[k]
                assert np.array_equal(G.bodies(sc[k]), A.bodies(a), equal_nan=True), f"frame {f} scene {k}: mix vs alone bodies"
                assert np.array_equal(G.bricks(sc[k]), A.bricks(a), equal_nan=True), f"frame {f} scene {k}: mix vs alone bricks"


def test_lattice_tuner_changes_no_result(oracle):
    # the batched frame's tuner steps one world through every candidate (solve route, sweep route), one frame each
    # per round; the world it leaves equals, bit for bit, the oracle and a world stepped the same frames with the
    # defaults
    from shatterforge_amd import autotune
    print("case: 4 scenes batched, the tuner's 6 candidates x 4 rounds, vs the defaults and the oracle", flush=True)
    caps = [lc.brick_count(c) + 8 for c in MIX]
    worlds = []
    for _ in range(2):
        W = Lattice(sum(caps), 4096 * len(MIX), scenes=len(MIX))
        sc = [W.add_scene(caps[k], 4096, structs=4, seed=0x9E3779B9) for k in range(len(MIX))]
        for k, c in enumerate(MIX):
            lc.build(W, sc[k], c)
        worlds.append((W, sc))
    Os = [oracle.LatticeOracle(order=oracle.ORDER_DEV, brick_cap=caps[k], body_cap=4096) for k in range(len(MIX))]
    for k, c in enumerate(MIX):
        lc.build(Os[k], None, c)
    for W, sc in worlds:
        W.step(DT)
        for k, c in enumerate(MIX):
            for ch in lc.charges(c):
                W.explode(sc[k], *ch)
    for k, c in enumerate(MIX):
        Os[k].step(DT)
        for ch in lc.charges(c):
            Os[k].explode(*ch)
    (T, st), (D, sd) = worlds
    table = autotune.tune_lattice(T, rounds=4)
    n = table["frames"]
    assert n == 24
    for _ in range(n):
        D.step(DT)
        for O in Os:
            O.step(DT)
    for k in range(len(MIX)):
        assert np.array_equal(T.bodies(st[k]), D.bodies(sd[k]), equal_nan=True), f"scene {k}: tuned vs default bodies"
        assert np.array_equal(T.bricks(st[k]), D.bricks(sd[k]), equal_nan=True), f"scene {k}: tuned vs default bricks"
        _assert_same(T, st[k], Os[k], f"scene {k} vs oracle after the tuner")
    assert table["best"]["ms"] > 0.0

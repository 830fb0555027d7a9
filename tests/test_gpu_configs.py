"""The brick engine's workloads at full size on the GPU.

  tower (5,760 bricks, 600 frames) ... every body, brick and statistic equal to the oracle in the device's orders,
                                       checked every 25 frames through the whole run
  site (37 structures, 382,857 bricks, 1,200 frames)
      standing ...................... with converged load fields and every structure's joints live (woken once a
                                       second, as the reference's stability check does: tests/stab.mjs:6-15),
                                       nothing is lost in 20 s: every structure keeps every brick, no piece forms
                                       (the check counts bricks and pieces, not joints)
      collapse ...................... eight charge sets: every position finite, every brick accounted for (standing,
                                       in a piece, or removed by a blast or a crush), the pieces asleep or moving
                                       below the engine's speed clamp
"""
import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd.lattice import Lattice
from shatterforge_amd.load import LoadBatch

pytestmark = pytest.mark.gpu
DT = np.float32(1.0 / 60.0)
C = 320


def _same(G, s, O, tag):
    gb, ob = G.bodies(s), O.bodies()
    assert gb.shape == ob.shape and np.array_equal(gb, ob, equal_nan=True), f"{tag}: bodies differ"
    gk, ok = G.bricks(s), O.bricks()
    assert np.array_equal(gk, ok, equal_nan=True), f"{tag}: bricks differ"
    assert np.array_equal(G.stats(s), O.stats()), f"{tag}: stats {G.stats(s)} vs {O.stats()}"
    assert G.rng(s) == O.rng(), f"{tag}: random stream"


def test_tower_full_size_600_frames(oracle):
    print("case: 12x12x40 tower (5,760 bricks), warm-up, 4 charges at the column bases, 600 frames", flush=True)
    mat, xyz, infill = scenes.tower()
    xyz = np.array(xyz, np.int32)
    xyz[:, 0] += C - 6
    xyz[:, 2] += C - 6
    nb, nd = len(mat) + 64, 8192
    G = Lattice(nb, nd)
    s = G.add_scene(nb, nd, structs=2)
    O = oracle.LatticeOracle(order=oracle.ORDER_DEV, brick_cap=nb, body_cap=nd)
    G.add_structure(s, mat, xyz, infill)
    O.add_structure(mat, xyz, infill)
    G.finalize(s)
    O.finalize()
    G.warm_all(s)
    O.warm_all()
    n = 0
    while O.stats()[14] > 0 or n < 30:
        G.step(DT)
        O.step(DT)
        n += 1
    _same(G, s, O, f"after {n} warm-up frames")
    for cx, cz in [(1, 1), (1, 11), (11, 1), (11, 11)]:
        G.explode(s, (cx - 6) * 0.5, 0.6, (cz - 6) * 0.5, 2.2, 1.4)
        O.explode((cx - 6) * 0.5, 0.6, (cz - 6) * 0.5, 2.2, 1.4)
    _same(G, s, O, "after the charges")
    for f in range(600):
        G.step(DT)
        O.step(DT)
        if f % 25 == 24:
            _same(G, s, O, f"frame {f}")
    st = O.stats()
    print(f"tower after 600 frames: {int(st[0])} standing, {int(st[2])} pieces", flush=True)
    assert st[2] > 0
    G.close()


def _site_world(seed=1):
    st = scenes.site(seed)
    nbr = sum(len(m) for m, _, _ in st)
    L = Lattice(nbr + 1024, 32768)
    s = L.add_scene(nbr + 1024, 32768, structs=len(st))
    for mat, xyz, infill in st:
        L.add_structure(s, mat, xyz, infill)
    L.finalize(s)
    lb = LoadBatch()
    for mat, xyz, infill in st:
        lb.add(mat, xyz, infill)
    assert lb.solve(rel_tol=1e-6) == len(st)
    for k in range(len(st)):
        L.set_field(s, k, lb.field(k).astype(np.float32))
    lb.close()
    return L, s, st, nbr


def test_site_stands_for_20_seconds():
    print("case: site of 37 structures, converged fields, joints live, no charges, 1,200 frames", flush=True)
    L, s, st, nbr = _site_world()
    for f in range(1200):
        if f % 60 == 0:
            L.wake_all(s)
        L.step(DT)
        if f % 300 == 299:
            print(f"frame {f}: {L.stats(s)[:5].tolist()}", flush=True)
    ss = L.structs(s)
    stt = L.stats(s)
    assert len(ss) == 37 and stt[15] == 37
    assert np.array_equal(ss[:, 5], ss[:, 4]), np.flatnonzero(ss[:, 5] != ss[:, 4]).tolist()
    assert stt[0] == nbr and stt[2] == 0 and stt[1] == 0
    # joints may fatigue where a structure's own weight overloads them (the check is that nothing falls)
    print(f"site after 20 s: every brick standing, joints failed {L.failures(s).tolist()}", flush=True)
    L.close()


def test_site_collapse_accounts_for_every_brick():
    print("case: site of 37 structures, 8 charge sets of 4 charges, 1,200 frames", flush=True)
    L, s, st, nbr = _site_world()
    for ch in scenes.site_charges(st, 1):
        for c in ch:
            L.explode(s, *c)
    for f in range(1200):
        L.step(DT)
        if f % 300 == 299:
            print(f"frame {f}: {L.stats(s)[:5].tolist()}", flush=True)
    stt = L.stats(s)
    br = L.bricks(s)
    bd = L.bodies(s)
    assert np.isfinite(br[:, 2:7]).all() and np.isfinite(bd[:, 4:17]).all()
    live = bd[bd[:, 0] > 0]
    removed = int((br[:, 0] == 0).sum())
    assert int(stt[0]) + int(stt[1]) + removed == nbr, (stt[:3].tolist(), removed, nbr)
    assert int(stt[1]) == int(live[:, 2].sum())
    assert int(stt[2]) > 100 and removed > 0
    # the speed clamp (60 m/s) and the out-of-bounds kill keep every piece in the site
    assert np.linalg.norm(live[:, 11:14], axis=1).max() <= 60.0 + 1e-3
    assert np.abs(live[:, [4, 6]]).max() < 190.0 and live[:, 5].min() > -15.0
    print(f"site after 1,200 frames: {int(stt[0])} standing, {int(stt[2])} pieces ({int(stt[1])} bricks), "
          f"{removed} removed, joints failed {L.failures(s).tolist()}", flush=True)
    L.close()

"""The legacy engine's own exports (sf_legacy_*: src/lib.rs init/step/explode/throw_charge/detonate/n_structs/
ptr_structs/ptr_stats/events, tools/debug_exports.rs dbg_*) on the default world, called through ctypes the way the
reference's page and node tests call the wasm exports, and checked against the handle API on the same world."""
import ctypes

import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd._native import lib

pytestmark = pytest.mark.gpu
DT = 1.0 / 60.0


def _f(p, n):
    assert p
    return np.ctypeslib.as_array(ctypes.cast(p, ctypes.POINTER(ctypes.c_float)), shape=(n,)).copy()


def test_legacy_exports_on_the_default_world():
    print("case: sf_legacy_init (site scene), 30 frames, 8 thrown charges at a tower's columns, 120 frames", flush=True)
    L = lib()
    site = scenes.site(1)
    sizes = np.array([len(m) for m, _, _ in site])
    nb = L.sf_legacy_init()
    assert nb == sizes.sum()
    n = L.sf_legacy_n_structs()
    assert n == len(site) == 37
    S = _f(L.sf_legacy_ptr_structs(), 12 * n).reshape(n, 12)
    assert np.array_equal(S[:, 4], sizes) and np.array_equal(S[:, 5], sizes)
    # (the statistics are written by a frame, as the reference's step writes them)
    for _ in range(30):
        L.sf_legacy_step(ctypes.c_float(DT))
    st = _f(L.sf_legacy_ptr_stats(), 16)
    assert st[0] == nb and st[2] == 0 and st[14] == 0 and st[15] == 37
    # eight charges at the ground-floor concrete of structure 0 (a frame tower); a ninth does not fit
    mat, xyz, _ = site[0]
    g = xyz[(xyz[:, 1] == 0) & (mat == 1)]
    pick = g[np.linspace(0, len(g) - 1, 8).astype(int)]
    for x, _, z in pick:
        assert L.sf_legacy_throw_charge(ctypes.c_float((x - 320) * 0.5 + 0.25), ctypes.c_float(0.6),
                                        ctypes.c_float((z - 320) * 0.5 + 0.25), 0.0, 0.0, 0.0) == 1
    assert L.sf_legacy_throw_charge(0.0, 0.6, 0.0, 0.0, 0.0, 0.0) == 0
    assert L.sf_legacy_detonate() == 8
    assert L.sf_legacy_detonate() == 0
    for _ in range(120):
        L.sf_legacy_step(ctypes.c_float(DT))
    st = _f(L.sf_legacy_ptr_stats(), 16)
    S = _f(L.sf_legacy_ptr_structs(), 12 * n).reshape(n, 12)
    assert st[2] > 0 and S[0, 5] < S[0, 4]
    # the same world through the handle API
    scene = ctypes.c_int(0)
    h = L.sf_legacy_world(ctypes.byref(scene))
    assert h
    counts = np.zeros(3, np.int32)
    L.sf_lat_counts(h, scene.value, counts.ctypes.data)
    bodies = np.zeros((counts[1], 22), np.float32)
    L.sf_lat_bodies(h, scene.value, bodies.ctypes.data)
    bricks = np.zeros((counts[0], 13), np.float32)
    L.sf_lat_bricks(h, scene.value, bricks.ctypes.data)
    alive = bodies[bodies[:, 0] > 0]
    assert st[2] == len(alive) and st[3] == (alive[:, 1] == 0).sum() and st[1] == alive[:, 2].sum()
    # dbg_hist: every brick of structure 0 standing, in a piece or removed, per height; standing equals ptr_structs
    hist = _f(L.sf_legacy_dbg_hist(0), 304)
    h3 = hist[:300].reshape(100, 3)
    assert h3.sum() == sizes[0] and h3[:, 0].sum() == S[0, 5]
    if h3[:, 1].sum() > 0:
        assert hist[300] <= hist[301] and hist[302] <= hist[303]
    # dbg_top: the 12 most stressed standing bricks of structure 0, in order, and the structure's worst ratio
    top = _f(L.sf_legacy_dbg_top(0), 61)
    r = top[:60].reshape(12, 5)
    assert (np.diff(r[:, 0]) <= 0).all() and top[60] >= r[0, 0]
    lo = int(sizes[:0].sum())
    own = bricks[lo:lo + sizes[0]]
    standing = own[(own[:, 0] > 0) & (own[:, 1] < 0)]
    assert np.isclose(r[0, 0], standing[:, 11].max())
    # dbg_matstats equals the handle's fragment table
    ms = np.zeros(50, np.float32)
    L.sf_lat_matstats(h, scene.value, ms.ctypes.data)
    assert np.array_equal(_f(L.sf_legacy_dbg_matstats(), 50), ms)
    # dbg_bodies: the 8 largest pieces in order; dbg_bigbody: the largest
    bb = _f(L.sf_legacy_dbg_bodies(), 24).reshape(8, 3)
    assert bb[0, 0] == alive[:, 2].max() and (np.diff(bb[:, 0]) <= 0).all() and (bb[bb[:, 0] > 0, 1:] >= 0.5).all()
    big = _f(L.sf_legacy_dbg_bigbody(), 12)
    assert big[0] == alive[:, 2].max() and 0.0 <= big[4] <= 180.0
    # dbg_speed: awake pieces counted
    sp = _f(L.sf_legacy_dbg_speed(), 10)
    assert sp[6] == st[3] and sp[0] >= 0.0
    # dbg_wake_all: every structure's joints live after the next frame
    L.sf_legacy_dbg_wake_all()
    L.sf_legacy_step(ctypes.c_float(DT))
    assert (_f(L.sf_legacy_ptr_structs(), 12 * n).reshape(n, 12)[:, 10] == 1.0).all()
    L.sf_legacy_clear_events()
    assert L.sf_legacy_n_events() == 0 and L.sf_legacy_ptr_events()
    # init again: a fresh world
    assert L.sf_legacy_init() == nb
    L.sf_legacy_step(ctypes.c_float(DT))
    assert _f(L.sf_legacy_ptr_stats(), 16)[2] == 0

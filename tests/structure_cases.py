"""Structures for the structural-analysis tests: the same bricks in the product's store and in the oracle's."""
import numpy as np

from shatterforge_amd import scenes, structure
from oracle import pyoracle as po

C, M, S, G, T = 0, 1, 2, 3, 4


def both(bricks):
    """(product Bricks, oracle store, min y per brick) holding `bricks` [(min, size, mat, paint, flags)], joints built"""
    b = structure.Bricks()
    o = po.BricksOracle()
    ys = []
    for mn, size, m, paint, fl in bricks:
        i = b.add(mn, size, m, paint)
        k = o.add(mn, size, m, paint)
        assert i == k
        if fl != 1:
            b.set_flags(i, fl)
            o.set_flags(k, fl)
        ys.append(mn[1])
    b.build_joints()
    o.build_joints()
    return b, o, ys


def gravity(o, ys, extra=()):
    n = len(ys)
    load = np.array([o.mass(i) * np.float32(9.81) for i in range(n)], np.float32)
    for i, f in extra:
        load[i] += np.float32(f)
    return np.arange(n, dtype=np.uint32), load, np.array([y <= 0 for y in ys], np.uint8)


def cantilever(length, h, m=C):
    return [([0, 0, 0], [8, 8, 8], m, 0, 1), ([0, 8, 0], [8 + length, h, 8], m, 0, 1)]


def span(span_, h, bear=2, m=C):
    return [([0, 0, 0], [bear, 8, 8], m, 0, 1), ([bear + span_, 0, 0], [bear, 8, 8], m, 0, 1),
            ([0, 8, 0], [2 * bear + span_, h, 8], m, 0, 1)]


def slab_on_column(off):
    return [([0, 0, 0], [80, 2, 80], C, 0, 1), ([off, 2, off], [4, 24, 4], C, 1, 1), ([0, 26, 0], [48, 2, 48], C, 0, 1)]


def mixed_frame():
    """steel beams on masonry piers, a timber floor on them, a cracked joint, a cast stair flight"""
    return [([0, 0, 0], [8, 24, 8], M, 6, 1), ([40, 0, 0], [8, 24, 8], M, 6, 1), ([0, 24, 0], [48, 2, 8], S, 7, 1),
            ([0, 26, 0], [48, 2, 8], T, 8, 1), ([8, 0, 0], [32, 8, 2], C, 4, 1 | 8), ([8, 8, 0], [4, 16, 1], G, 0, 3)]

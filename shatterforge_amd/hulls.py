"""The convex-chunk library (engine/src/hull.rs, gpu/gpu_world.js hullLibrary :2113-2208) as the library builds it:
16 fracture patterns of the cube [-1,1]^3, each a Voronoi partition of 2-6 seeds; hull 0 is the whole cube.

    ids = pattern(3)                       # the hull ids of pattern 3's cells
    c = cell(ids[0])                       # counts, box, volume shares, centroid, covariance
    pieces = shapes(p, q, h, 3, mat=1)     # a brick broken by pattern 3: SF_SHAPE_IN records for scenes.body
"""
import numpy as np

from ._native import lib, check, ptr

MAXV, MAXF, MAXFV, MAXS, MAXD = 16, 12, 8, 24, 12


def count():
    return int(lib().sf_hull_count())


def pattern_count():
    return int(lib().sf_hull_pattern_count())


def pattern(p):
    ids = np.zeros(8, np.int32)
    n = check(lib().sf_hull_pattern(int(p), ptr(ids)))
    return ids[:n].tolist()


def cell(h):
    o = np.zeros(26, np.float64)
    check(lib().sf_hull_cell(int(h), ptr(o)))
    return dict(nv=int(o[0]), nf=int(o[1]), nseg=int(o[2]), ndir=int(o[3]), box_c=o[4:7].copy(), box_h=o[7:10].copy(),
                box_vol=float(o[10]), cube_vol=float(o[11]), cen=o[12:15].copy(), cov=o[15:24].reshape(3, 3).copy(),
                pattern=int(o[24]), slot=int(o[25]))


def vertices(h, cube=False):
    n = check(lib().sf_hull_vertices(int(h), 1 if cube else 0, None))
    o = np.zeros((max(n, 1), 3), np.float64)
    check(lib().sf_hull_vertices(int(h), 1 if cube else 0, ptr(o)))
    return o[:n]


def faces(h, cube=False):
    """[(polygon [k, 3], plane (n, d) in box coordinates, what lies across: cube face 0..5 or 6 + slot)]"""
    out = []
    for f in range(cell(h)["nf"]):
        xyz = np.zeros((MAXFV, 3), np.float64)
        pl = np.zeros(4, np.float64)
        other = np.zeros(1, np.int32)
        n = check(lib().sf_hull_face(int(h), f, 1 if cube else 0, ptr(xyz), ptr(pl), ptr(other)))
        out.append((xyz[:n].copy(), pl, int(other[0])))
    return out


def edges(h):
    c = cell(h)
    pairs = np.zeros((MAXS, 2), np.int32)
    dirs = np.zeros((MAXD, 3), np.float32)
    check(lib().sf_hull_edges(int(h), ptr(pairs), ptr(dirs)))
    return pairs[:c["nseg"]], dirs[:c["ndir"]]


def tables():
    """The narrow phase's f32 tables (v4, pl4, dir4, fidx, fcnt, seg, head), field by field"""
    H = count()
    t = (np.zeros(H * MAXV * 4, np.float32), np.zeros(H * MAXF * 4, np.float32), np.zeros(H * MAXD * 4, np.float32),
         np.zeros(H * MAXF * MAXFV, np.uint8), np.zeros(H * MAXF, np.uint8), np.zeros(H * MAXS * 2, np.uint8),
         np.zeros(H * 4, np.int32))
    check(lib().sf_hull_tables(*[ptr(x) for x in t]))
    return t


def shapes(p, q, h, pat, mat=1):
    """A brick (centre p, rotation q, half size h) broken by pattern `pat`: one SF_SHAPE_IN record per cell"""
    p = np.asarray(p, np.float32)
    q = np.asarray(q, np.float32)
    h = np.asarray(h, np.float32)
    n = check(lib().sf_hull_shapes(ptr(p), ptr(q), ptr(h), int(pat), int(mat), None))
    out = np.zeros((n, 12), np.float32)
    check(lib().sf_hull_shapes(ptr(p), ptr(q), ptr(h), int(pat), int(mat), ptr(out)))
    return out

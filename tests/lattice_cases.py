"""Scenes of the lattice-world tests, built identically for the GPU library and the oracle."""
import numpy as np

from shatterforge_amd import scenes

C = 320  # lattice index of x = 0 and z = 0


def centred(mat, xyz, infill):
    xyz = np.array(xyz, np.int32)
    wx = int(xyz[:, 0].max()) + 1
    wz = int(xyz[:, 2].max()) + 1
    xyz[:, 0] += C - wx // 2
    xyz[:, 2] += C - wz // 2
    return mat, xyz, infill


def polycube(rng, n):
    """a connected random polycube of n lattice cells, grown from the origin"""
    cells = [(0, 0, 0)]
    seen = {cells[0]}
    dirs = [(1, 0, 0), (-1, 0, 0), (0, 1, 0), (0, -1, 0), (0, 0, 1), (0, 0, -1)]
    while len(cells) < n:
        c = cells[int(rng.integers(len(cells)))]
        d = dirs[int(rng.integers(6))]
        q = (c[0] + d[0], c[1] + d[1], c[2] + d[2])
        if q not in seen:
            seen.add(q)
            cells.append(q)
    return np.array(cells, np.int32)


def drop_scene(seed=3, nbody=24):
    """a small column (static) and loose bodies of 1-12 bricks dropped around it with random velocities"""
    rng = np.random.default_rng(seed)
    col = (np.full(10, 2, np.uint8), np.stack([np.full(10, C), np.arange(10), np.full(10, C)], 1).astype(np.int32),
           np.zeros(10, np.uint8))
    bodies = []
    for k in range(nbody):
        n = int(rng.integers(1, 13))
        cells = polycube(rng, n)
        off = np.array([C + int(rng.integers(-6, 7)), 4 + int(rng.integers(0, 14)), C + int(rng.integers(-6, 7))])
        xyz = cells + off
        mat = rng.integers(1, 6, size=n).astype(np.uint8)
        v = rng.normal(0, 2.0, 3).astype(np.float32)
        w = rng.normal(0, 1.5, 3).astype(np.float32)
        bodies.append((mat, xyz.astype(np.int32), v, w))
    # drop bodies whose cells overlap the column or each other
    used = set(map(tuple, col[1].tolist()))
    keep = []
    for b in bodies:
        cs = set(map(tuple, b[1].tolist()))
        if cs & used:
            continue
        used |= cs
        keep.append(b)
    return col, keep


def small_tower(w=8, h=16):
    """masonry tower with 2x2 concrete corner columns and a concrete slab every 6th layer"""
    mats, xyz = [], []
    for y in range(h):
        for z in range(w):
            for x in range(w):
                col = (x < 2 or x >= w - 2) and (z < 2 or z >= w - 2)
                mats.append(1 if (col or y % 6 == 5) else 2)
                xyz.append((x, y, z))
    return centred(np.array(mats, np.uint8), np.array(xyz, np.int32), np.zeros(len(mats), np.uint8))


def frame_building(n_target=2500, seed=4):
    return centred(*scenes.building_for(n_target, seed=seed))


def office_all_columns(storeys=6, w=16, seed=3):
    """a concrete-frame office of `storeys` storeys on a w x w footprint, with a charge of radius 3 m at every fourth
    ground-floor column cell (the reference's office collapse fixture brings down all ground columns,
    tests/office_collapse.json)"""
    mat, xyz, inf = centred(*scenes.building(w, w, storeys, 6, 0, seed))
    g = xyz[(xyz[:, 1] == 0) & (mat == 1)]
    cols = sorted(set((int(a), int(b)) for a, b in g[:, [0, 2]]))
    charges = [((x - C) * 0.5 + 0.5, 0.6, (z - C) * 0.5 + 0.5, 3.0, 1.4) for x, z in cols[::4]]
    return (mat, xyz, inf), charges


def masonry_shaft():
    """a hollow 10 x 10 masonry shaft 40 bricks high with timber floors every 8 layers and two charges of radius
    4 m at opposite corners of its base (the reference's masonry collapse fixture: a church tower base,
    tests/masonry_collapse.json)"""
    cells, m = [], []
    for y in range(40):
        for z in range(10):
            for x in range(10):
                if x in (0, 9) or z in (0, 9):
                    cells.append((x, y, z)); m.append(2)
                elif y % 8 == 7:
                    cells.append((x, y, z)); m.append(5)
    st = centred(np.array(m, np.uint8), np.array(cells, np.int32), np.zeros(len(m), np.uint8))
    return st, [(-2.25, 0.6, -2.5, 4.0, 1.4), (2.25, 0.6, 2.5, 4.0, 1.4)]


def ram_scene(speed=14.0):
    """a masonry wall 10 wide, 6 high and 2 deep, struck broadside by two concrete rams: a small one (2 x 2 x 2, which
    ploughs bricks out of the wall) and a large one (4 x 4 x 3, which only damages the bricks it strikes)"""
    wall = [(x, y, z) for y in range(6) for z in range(2) for x in range(10)]
    mat = np.full(len(wall), 2, np.uint8)
    xyz = np.array(wall, np.int32) + np.array([C - 5, 0, C - 1], np.int32)
    small = np.array([(x, y, z) for x in range(2) for y in range(2) for z in range(2)], np.int32) + np.array([C - 3, 2, C - 6], np.int32)
    large = np.array([(x, y, z) for x in range(4) for y in range(4) for z in range(3)], np.int32) + np.array([C, 1, C + 4], np.int32)
    bodies = [(np.full(len(small), 1, np.uint8), small, (0.0, 0.0, speed), (0.0, 0.0, 0.0)),
              (np.full(len(large), 1, np.uint8), large, (0.0, 0.0, -speed), (0.0, 0.0, 0.0))]
    return (mat, xyz, np.zeros(len(mat), np.uint8)), bodies


def heavy_ram_scene():
    """a masonry wall 16 wide, 10 high and 3 deep struck broadside at 40 m/s by a 6 x 6 x 4 steel ram: long crush walks
    (over 300 bricks crushed in one walk, its queue near 300 bricks) and a wall torn into pieces"""
    wall = [(x, y, z) for y in range(10) for z in range(3) for x in range(16)]
    mat = np.full(len(wall), 2, np.uint8)
    xyz = np.array(wall, np.int32) + np.array([C - 8, 0, C - 1], np.int32)
    ram = np.array([(x, y, z) for x in range(6) for y in range(6) for z in range(4)], np.int32) + np.array([C - 3, 1, C - 6], np.int32)
    bodies = [(np.full(len(ram), 3, np.uint8), ram, (0.0, 0.0, 40.0), (0.0, 0.0, 0.0))]
    return (mat, xyz, np.zeros(len(mat), np.uint8)), bodies


def column_bases(mat, xyz, rng, k):
    g = xyz[(xyz[:, 1] == 0) & (mat == 1)]
    cols = sorted(set((int(a), int(b)) for a, b in g[:, [0, 2]]))
    pick = [cols[i] for i in rng.choice(len(cols), size=min(k, len(cols)), replace=False)]
    return [((x - C) * 0.5 + 0.5, 0.6, (z - C) * 0.5 + 0.5) for x, z in pick]


def build(world, scene_or_none, case):
    """put a case into a GPU Lattice scene (world, scene) or an oracle (world, None)"""
    kind = case["kind"]
    def add_struct(mat, xyz, infill):
        if scene_or_none is None:
            world.add_structure(mat, xyz, infill)
        else:
            world.add_structure(scene_or_none, mat, xyz, infill)
    def add_body(mat, xyz, v, w):
        if scene_or_none is None:
            world.add_body(mat, xyz, v, w)
        else:
            world.add_body(scene_or_none, mat, xyz, v, w)
    if kind == "drop":
        col, bodies = drop_scene(case.get("seed", 3), case.get("nbody", 24))
        add_struct(*col)
        for mat, xyz, v, w in bodies:
            add_body(mat, xyz, v, w)
    elif kind == "heavy_ram":
        wall, bodies = heavy_ram_scene()
        add_struct(*wall)
        for mat, xyz, v, w in bodies:
            add_body(mat, xyz, v, w)
    elif kind == "ram":
        wall, bodies = ram_scene(case.get("speed", 14.0))
        add_struct(*wall)
        for mat, xyz, v, w in bodies:
            add_body(mat, xyz, v, w)
    elif kind == "tower":
        add_struct(*small_tower(case.get("w", 8), case.get("h", 16)))
    elif kind == "building":
        add_struct(*frame_building(case.get("n", 2500), case.get("seed", 4)))
    elif kind == "office":
        add_struct(*office_all_columns()[0])
    elif kind == "shaft":
        add_struct(*masonry_shaft()[0])
    if scene_or_none is None:
        world.finalize()
        world.warm_all()
    else:
        world.finalize(scene_or_none)
        world.warm_all(scene_or_none)


def charges(case):
    kind = case["kind"]
    if kind == "tower":
        mat, xyz, _ = small_tower(case.get("w", 8), case.get("h", 16))
    elif kind == "building":
        mat, xyz, _ = frame_building(case.get("n", 2500), case.get("seed", 4))
    elif kind == "office":
        return office_all_columns()[1]
    elif kind == "shaft":
        return masonry_shaft()[1]
    else:
        return []
    rng = np.random.default_rng(case.get("cseed", 1))
    return [(x, y, z, case.get("r", 2.5), case.get("p", 1.4)) for x, y, z in column_bases(mat, xyz, rng, case.get("nch", 4))]


def brick_count(case):
    kind = case["kind"]
    if kind == "drop":
        col, bodies = drop_scene(case.get("seed", 3), case.get("nbody", 24))
        return len(col[0]) + sum(len(b[0]) for b in bodies)
    if kind == "ram":
        wall, bodies = ram_scene(case.get("speed", 14.0))
        return len(wall[0]) + sum(len(b[0]) for b in bodies)
    if kind == "heavy_ram":
        wall, bodies = heavy_ram_scene()
        return len(wall[0]) + sum(len(b[0]) for b in bodies)
    if kind == "tower":
        return len(small_tower(case.get("w", 8), case.get("h", 16))[0])
    if kind == "office":
        return len(office_all_columns()[0][0])
    if kind == "shaft":
        return len(masonry_shaft()[0][0])
    return len(frame_building(case.get("n", 2500), case.get("seed", 4))[0])

"""Seeded scenes as numpy arrays (thin wrappers over the library's host-side generators).

A box scene is ``(bodies, shape_body, shapes)``: ``bodies`` float32 [nb, 20] (p, q, v, w, inv_m, inv_i, awake,
sleep_t, reserved), ``shape_body`` int32 [ns], ``shapes`` float32 [ns, 13] (local p, local q, half extents, mu, restitution,
hull: 0 the whole box, else a convex cell of the chunk library).
A structure is ``(mat, xyz, infill)`` on the 0.5 m lattice.
"""
import numpy as np

from ._native import lib, check, ptr

BODY_F = 20
SHAPE_F = 13


def _box_scene(fn, *args):
    L = lib()
    ns = np.zeros(1, np.int32)
    nb = check(fn(*args, None, None, None, ptr(ns)))
    bodies = np.zeros((nb, BODY_F), np.float32)
    shape_body = np.zeros(int(ns[0]), np.int32)
    shapes = np.zeros((int(ns[0]), SHAPE_F), np.float32)
    check(fn(*args, ptr(bodies), ptr(shape_body), ptr(shapes), ptr(ns)))
    return bodies, shape_body, shapes


def pile(n):
    """The reference's throughput scene (engine/src/phys.rs:632-646): n brick-sized boxes in a loose grid."""
    return _box_scene(lib().sf_scene_pile, int(n))


def stack(n):
    """n 0.5 m cubes stacked exactly (engine/src/wasm.rs:33-38)."""
    return _box_scene(lib().sf_scene_stack, int(n))


def rubble(nbody=50000, p=0.3, maxsize=64, box=(40.0, 20.0, 40.0), vsig=3.0, wsig=2.0, seed=1):
    """Rubble pile of polycubes (sizes ~ Geometric(p) truncated at maxsize), materials uniform over five,
    packed into `box` (m) on the ground, velocities N(0, vsig) per axis, angular N(0, wsig)."""
    return _box_scene(lib().sf_scene_rubble, int(nbody), float(p), int(maxsize), float(box[0]), float(box[1]),
                      float(box[2]), float(vsig), float(wsig), int(seed) & 0xFFFFFFFF)


def compound(xyz, mat):
    """One compound body from bricks at lattice cells: returns (body[20], shapes[n, 12])."""
    xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
    mat = np.ascontiguousarray(mat, np.uint8).reshape(-1)
    body = np.zeros(BODY_F, np.float32)
    shapes = np.zeros((len(mat), SHAPE_F), np.float32)
    check(lib().sf_build_compound(len(mat), ptr(xyz), ptr(mat), ptr(body), ptr(shapes)))
    return body, shapes


SHAPE_IN = 12  # a shape in the world for the body builder: p[3] q[4] h[3] material (1..5) hull


def shape(p, h, q=(0.0, 0.0, 0.0, 1.0), mat=1, hull=0):
    """A shape given in the world: centre, half size, rotation, material code (1 concrete .. 5 timber), hull id."""
    r = np.zeros(SHAPE_IN, np.float32)
    r[0:3] = p
    r[3:7] = q
    r[7:10] = h
    r[10] = mat
    r[11] = hull
    return r


def body(shapes, fixed=False, v=(0, 0, 0), w=(0, 0, 0)):
    """One body from shapes given in the world, as gpu/gpu_world.js buildBodies (:2212-2249) builds it: returns
    (body[20], shapes[n, 13])."""
    shp = np.ascontiguousarray(np.stack([np.asarray(x, np.float32) for x in shapes]), np.float32)
    b = np.zeros(BODY_F, np.float32)
    out = np.zeros((len(shp), SHAPE_F), np.float32)
    vv = np.asarray(v, np.float32)
    ww = np.asarray(w, np.float32)
    check(lib().sf_build_body(len(shp), ptr(shp), 1 if fixed else 0, ptr(vv), ptr(ww), ptr(b), ptr(out)))
    return b, out


def bodies(items):
    """A scene from (body, shapes[n, 13]) pairs (body() results): (bodies, shape_body, shapes)."""
    bs = np.stack([b for b, _ in items]).astype(np.float32)
    sb = np.concatenate([np.full(len(s), k, np.int32) for k, (_, s) in enumerate(items)])
    ss = np.concatenate([s for _, s in items]).astype(np.float32)
    return bs, sb, ss


def box(p, h, q=(0.0, 0.0, 0.0, 1.0), density=1900.0, mu=0.6, v=(0, 0, 0), w=(0, 0, 0), e=0.0):
    """A single box body as World::add_box builds it (engine/src/phys.rs:139-148), in f32 arithmetic."""
    f = np.float32
    hx, hy, hz = f(h[0]), f(h[1]), f(h[2])
    m = f(f(f(f(density) * f(8.0)) * hx) * hy) * hz
    body = np.zeros(BODY_F, np.float32)
    body[0:3] = p
    body[3:7] = q
    body[7:10] = v
    body[10:13] = w
    if m > 0:
        x2, y2, z2 = f(4.0) * hx * hx, f(4.0) * hy * hy, f(4.0) * hz * hz
        body[13] = f(1.0) / m
        body[14:17] = [f(12.0) / (m * (y2 + z2)), f(12.0) / (m * (x2 + z2)), f(12.0) / (m * (x2 + y2))]
        body[17] = 1.0
    shape = np.zeros(SHAPE_F, np.float32)
    shape[6] = 1.0
    shape[7:10] = h
    shape[10] = mu
    shape[11] = e
    return body, shape


def boxes(items):
    """Scene from a list of (body, shape) pairs."""
    bodies = np.stack([b for b, _ in items]).astype(np.float32)
    shapes = np.stack([s for _, s in items]).astype(np.float32)
    return bodies, np.arange(len(items), dtype=np.int32), shapes


def building(wx, wz, storeys, bay=6, remove=0, seed=1):
    """Frame/wall building: returns (mat, xyz, infill)."""
    L = lib()
    n = check(L.sf_scene_building(int(wx), int(wz), int(storeys), int(bay), int(remove), int(seed), None, None, None))
    mat = np.zeros(n, np.uint8)
    xyz = np.zeros((n, 3), np.int32)
    infill = np.zeros(n, np.uint8)
    check(L.sf_scene_building(int(wx), int(wz), int(storeys), int(bay), int(remove), int(seed), ptr(mat), ptr(xyz), ptr(infill)))
    return mat, xyz, infill


def building_for(n_target, seed=1):
    """A seeded building with about n_target bricks: square footprint and storey count chosen to fit, 1-8 ground
    columns removed."""
    rng = np.random.default_rng(seed)
    best = None
    for storeys in range(1, 40):
        for w in range(6, 120):
            n = lib().sf_scene_building(w, w, storeys, 6, 0, 1, None, None, None)
            if n >= n_target:
                err = abs(n - n_target) / n_target
                aspect = abs(storeys * 7 - w) / max(w, storeys * 7)
                score = err + 0.2 * aspect
                if best is None or score < best[0]:
                    best = (score, w, storeys)
                break
    _, w, storeys = best
    remove = int(rng.integers(1, 9))
    return building(w, w, storeys, 6, remove, seed)


def tower():
    """The 12 x 12 x 40 brick tower of the plain-method workload."""
    L = lib()
    n = 5760
    mat = np.zeros(n, np.uint8)
    xyz = np.zeros((n, 3), np.int32)
    infill = np.zeros(n, np.uint8)
    check(L.sf_scene_tower(ptr(mat), ptr(xyz), ptr(infill)))
    return mat, xyz, infill


def column(height, mat=1):
    """A vertical column of `height` bricks of one material."""
    n = int(height)
    return (np.full(n, mat, np.uint8), np.stack([np.zeros(n), np.arange(n), np.zeros(n)], 1).astype(np.int32),
            np.zeros(n, np.uint8))


def _box_shell(x0, z0, wx, wz, h, mat, y0=0):
    """a one-brick wall ring of height h around the footprint (x0, z0, wx, wz)"""
    out = []
    for y in range(y0, y0 + h):
        for x in range(wx):
            for z in range(wz):
                if x in (0, wx - 1) or z in (0, wz - 1):
                    out.append((x0 + x, y, z0 + z))
    return np.array(out, np.int32), np.full(len(out), mat, np.uint8)


def site(seed=1):
    """Site-scale scene (sf_scene_site): a 320 x 320 m site (lattice indices 0..639 around 320) of 37 structures of
    about 380k bricks: 16 frame towers of 8-16 storeys (3.5 m each, concrete frame, glass infill band, masonry walls),
    12 three-storey masonry row blocks, one 56 m single-brick-ring chimney, one arched masonry viaduct and 7
    steel-frame halls with glass walls, one structure per 40 m lot.
    Returns a list of (mat, xyz, infill) per structure, lattice coordinates."""
    L = lib()
    out = []
    for k in range(check(L.sf_scene_site(int(seed), -1, None, None, None))):
        n = check(L.sf_scene_site(int(seed), k, None, None, None))
        mat = np.zeros(n, np.uint8)
        xyz = np.zeros((n, 3), np.int32)
        infill = np.zeros(n, np.uint8)
        check(L.sf_scene_site(int(seed), k, ptr(mat), ptr(xyz), ptr(infill)))
        out.append((mat, xyz, infill))
    return out


def site_charges(structs, seed=1, sets=8, per_set=4, r=2.2, power=1.4):
    """seeded charge sets at ground-floor column bases (concrete or steel bricks on the ground) of random structures,
    world coordinates"""
    rng = np.random.default_rng(seed + 99)
    out = []
    for s in range(sets):
        ch = []
        for c in range(per_set):
            mat, xyz, _ = structs[int(rng.integers(len(structs)))]
            g = xyz[(xyz[:, 1] == 0) & ((mat == 1) | (mat == 3))]
            if not len(g):
                g = xyz[xyz[:, 1] == 0]
            x, _, z = g[int(rng.integers(len(g)))]
            ch.append(((x - 320) * 0.5 + 0.25, 0.6, (z - 320) * 0.5 + 0.25, r, power))
        out.append(ch)
    return out


def frame_tower(storeys=12, bays=3, kind=0, remove=(), bay=60, storey=30, slab=2, col=5):
    """A frame building for the structural analysis (the new engine's brick store, ticks of 0.125 m), laid out after
    engine/src/tower.rs: `bays` x `bays` bays of `bay` ticks, storeys of `storey` ticks (a `slab`-tick slab, then
    columns `col` ticks square on the bay grid), a concrete core of 0.25 m walls round the centre bay with a door each
    storey, glass infill on the facade between the columns, a roof slab. kind 0: concrete frame; 1: masonry columns;
    2: timber frame. `remove`: ground-storey columns (i, k) on the grid left out. Slabs wider than 255 ticks are laid
    as one brick per bay. Returns [(min, size, material, paint, flags)]; the ground slab rests on y = 0."""
    CONCRETE, MASONRY, GLASS, TIMBER = 0, 1, 3, 4
    INFILL = 2 | 1
    w = bays * bay
    colm = {0: CONCRETE, 1: MASONRY, 2: TIMBER}[kind]
    slabm = TIMBER if kind == 2 else CONCRETE
    out = []

    def put(mn, size, m, paint=0, flags=1):
        out.append((list(mn), list(size), m, paint, flags))

    def slab_at(y0):
        if w <= 255:
            put([0, y0, 0], [w, slab, w], slabm, 0)
        else:
            for i in range(bays):
                for k in range(bays):
                    put([i * bay, y0, k * bay], [bay, slab, bay], slabm, 0)

    grid = [min(i * bay, w - col) for i in range(bays + 1)]
    c0 = (bays // 2) * bay
    c1 = c0 + bay
    for s in range(storeys):
        y0 = s * storey
        slab_at(y0)
        ya, h = y0 + slab, storey - slab
        for i, gx in enumerate(grid):
            for k, gz in enumerate(grid):
                if s == 0 and (i, k) in remove:
                    continue
                put([gx, ya, gz], [col, h, col], colm, 1)
        # the core: south wall with a door, north wall, west and east walls between them
        door = 8
        mid = (c0 + col + c1) // 2
        put([c0 + col, ya, c0], [mid - door // 2 - (c0 + col), h, 2], CONCRETE, 3)
        put([mid + door // 2, ya, c0], [c1 - (mid + door // 2), h, 2], CONCRETE, 3)
        put([c0 + col, ya, c1 - 2], [c1 - (c0 + col), h, 2], CONCRETE, 3)
        put([c0, ya, c0 + col], [2, h, c1 - 2 - (c0 + col)], CONCRETE, 3)
        put([c1 - 2, ya, c0 + col], [2, h, c1 - 2 - (c0 + col)], CONCRETE, 3)
        # glass between the perimeter columns
        for i in range(bays):
            x0, x1 = grid[i] + col, grid[i + 1]
            put([x0, ya, 0], [x1 - x0, h, 1], GLASS, 0, INFILL)
            put([x0, ya, w - 1], [x1 - x0, h, 1], GLASS, 0, INFILL)
            put([0, ya, x0], [1, h, x1 - x0], GLASS, 0, INFILL)
            put([w - 1, ya, x0], [1, h, x1 - x0], GLASS, 0, INFILL)
    slab_at(storeys * storey)
    return out


def frame_towers(count, seed=1, storeys=(4, 16), bays=(2, 3, 4)):
    """`count` frame buildings of seeded sizes (storeys uniform in the range, bays from the list, kind from the three),
    some with ground-storey columns left out; a list of (bricks, description)."""
    rng = np.random.default_rng(seed)
    out = []
    for _ in range(count):
        st = int(rng.integers(storeys[0], storeys[1] + 1))
        nb = int(rng.choice(bays))
        kind = int(rng.integers(0, 3))
        nrem = int(rng.integers(0, 4))
        rem = {(int(rng.integers(0, nb + 1)), int(rng.integers(0, nb + 1))) for _ in range(nrem)}
        out.append((frame_tower(st, nb, kind, remove=rem), dict(storeys=st, bays=nb, kind=kind, removed=sorted(rem))))
    return out

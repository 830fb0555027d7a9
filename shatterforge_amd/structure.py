"""Structural analysis on the device: the host-side mirror of the new engine's structural solver
(engine/src/structure.rs:191-783: analyse, analyse_at, analyse_warm, Analysis, Crack, Warm, strength, flexural,
mass) over its brick store (engine/src/brick.rs Bricks), through the C ABI in libsf_hip (sf_bricks_*, sf_net_*).

    b = Bricks()
    b.add([0, 0, 0], [8, 8, 8], mat.CONCRETE)
    b.add([0, 8, 0], [48, 2, 8], mat.CONCRETE)
    b.build_joints()
    a = analyse(b, *gravity_case(b))          # a cantilever strip: a.max_ratio, a.failed, a.cracks
    many = analyse_many([(b, *gravity_case(b)), ...])   # one launch, a workgroup per case

Whole boxes only: the store has no fracture cells (convex hulls), so no brick is a hull cell.
"""
import ctypes
from dataclasses import dataclass, field

import numpy as np

from ._native import lib, check, ptr

TICK = 0.125
NONE = 0xFFFFFFFF
F_ALIVE, F_INFILL, F_DYNAMIC, F_MONO = 1, 2, 4, 8


class mat:
    """Materials of the new engine (engine/src/brick.rs:9-19)."""
    CONCRETE, MASONRY, STEEL, GLASS, TIMBER = 0, 1, 2, 3, 4


_STRENGTH = {mat.CONCRETE: (30e6, 4e6, 4e6), mat.MASONRY: (10e6, 0.3e6, 0.5e6),
             mat.STEEL: (250e6 * 0.16, 250e6 * 0.16, 150e6 * 0.10), mat.GLASS: (50e6, 7e6, 7e6)}
_FLEX = {mat.CONCRETE: 12e6, mat.MASONRY: 0.5e6, mat.STEEL: 250e6 * 0.12, mat.GLASS: 40e6}
_CRUSH = {mat.CONCRETE: 1.0e6, mat.MASONRY: 0.4e6, mat.STEEL: 20e6, mat.GLASS: 0.05e6}


def strength(m):
    """(compression, tension, shear) in Pa (structure.rs:24-37)."""
    return np.array(_STRENGTH.get(m, (21e6, 14e6, 4e6)), np.float32)


def flexural(m):
    """Bending strength in Pa (structure.rs:102-104)."""
    return np.float32(_FLEX.get(m, 20e6))


def crush_energy(m):
    """Energy per volume to crush the material to rubble, J/m^3 (structure.rs:64-66)."""
    return np.float32(_CRUSH.get(m, 0.2e6))


class Bricks:
    """The brick store (engine/src/brick.rs Bricks): boxes on the 0.125 m tick grid, joints as shared faces."""

    def __init__(self):
        self._h = lib().sf_bricks_create()
        self.min = []
        self.size = []

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sf_bricks_destroy(self._h)
            self._h = None

    def __len__(self):
        return len(self.min)

    def add(self, mn, size, material, paint=0, structure=0):
        """Bricks::add (brick.rs:157-170); returns the brick's index"""
        a = np.ascontiguousarray(mn, np.int32)
        s = np.ascontiguousarray(size, np.int32)
        i = check(lib().sf_bricks_add(self._h, ptr(a), ptr(s), int(material), int(paint), int(structure)))
        self.min.append(a.tolist())
        self.size.append(s.tolist())
        return i

    def y(self, i):
        return self.min[i][1]

    def set_flags(self, i, flags):
        check(lib().sf_bricks_set_flags(self._h, int(i), int(flags)))

    def flags(self, i):
        return check(lib().sf_bricks_flags(self._h, int(i)))

    def alive(self, i):
        return bool(self.flags(i) & F_ALIVE)

    def remove(self, i):
        check(lib().sf_bricks_remove(self._h, int(i)))

    def build_joints(self):
        """Bricks::build_joints (brick.rs:447-458); returns the live joints"""
        return check(lib().sf_bricks_build_joints(self._h))

    def joints(self):
        """[slots, 5]: ja (min side), jb (-1 once removed), axis, area in tick^2, damage"""
        n = check(lib().sf_bricks_joints(self._h, None, 0))
        out = np.zeros((max(n, 1), 5), np.int32)
        check(lib().sf_bricks_joints(self._h, ptr(out), n))
        return out[:n]

    def set_joint_damage(self, j, dmg):
        check(lib().sf_bricks_set_joint_damage(self._h, int(j), int(dmg)))

    def mass(self, i):
        """structure.rs mass (:66-71), kg"""
        return np.float32(lib().sf_bricks_mass(self._h, int(i)))

    def set_cell(self, i, body=-1, damage=0, hull=0):
        """the per-brick fields the store carries for its format: owning body (-1 none), damage byte, convex cell"""
        check(lib().sf_bricks_set_cell(self._h, int(i), int(body), int(damage), int(hull)))

    def cell(self, i):
        out = np.zeros(3, np.int32)
        check(lib().sf_bricks_cell(self._h, int(i), ptr(out)))
        return tuple(int(v) for v in out)

    def save(self):
        """Bricks::save (brick.rs:116-127): the store as bytes"""
        n = check(lib().sf_bricks_save(self._h, None, 0))
        out = np.zeros(n, np.uint8)
        check(lib().sf_bricks_save(self._h, ptr(out), n))
        return out.tobytes()

    @classmethod
    def load(cls, data):
        """Bricks::load (brick.rs:128-145): a store from bytes (SfError for malformed bytes)"""
        buf = np.frombuffer(bytes(data), np.uint8).copy()
        h = lib().sf_bricks_load(ptr(buf), len(buf))
        if not h:
            check(-1)
        b = cls.__new__(cls)
        b._h = h
        b.min, b.size = [], []
        box = np.zeros(6, np.int32)
        for i in range(check(lib().sf_bricks_len(h))):
            check(lib().sf_bricks_box(h, i, ptr(box)))
            b.min.append(box[0:3].tolist())
            b.size.append(box[3:6].tolist())
        return b


class Warm:
    """The last solution of a load case, per cell (structure.rs Warm, :178-189)."""

    def __init__(self):
        self._h = lib().sf_warm_create()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sf_warm_destroy(self._h)
            self._h = None

    def __len__(self):
        return check(lib().sf_warm_size(self._h))


@dataclass
class Crack:
    """A crack through a brick from bending (structure.rs:110-112)."""
    brick: int
    axis: int
    at: int
    ratio: float


@dataclass
class Analysis:
    """structure.rs Analysis (:119-133), plus the cells and their displacements."""
    failed: list
    cracks: list
    max_ratio: float
    iterations: int
    nodes: int
    worst_in: tuple
    ratios: list
    inner: list
    links: int = 0
    levels: list = field(default_factory=list)
    joint_ids: np.ndarray = None
    joint_ratios: np.ndarray = None
    inner_faces: np.ndarray = None
    inner_ratios: np.ndarray = None
    cells: np.ndarray = None
    x: np.ndarray = None


def gravity_case(b, extra=()):
    """The reference's gravity case (structure.rs:793-799, validation.rs:17-23): every live brick under its weight plus
    extra loads (index, N), the bricks resting on y = 0 the supports. Returns (set, load, sink)."""
    live = [i for i in range(len(b)) if b.alive(i)]
    set_ = np.array(live, np.uint32)
    load = np.array([b.mass(i) * np.float32(9.81) for i in live], np.float32)
    for i, f in extra:
        load[live.index(i)] += np.float32(f)
    sink = np.array([b.y(i) <= 0 for i in live], np.uint8)
    return set_, load, sink


class Net:
    """Analyses queued and solved together on the device (sf_net_*)."""

    def __init__(self):
        self._h = lib().sf_net_create()
        self._keep = []

    def __del__(self):
        if getattr(self, "_h", None):
            lib().sf_net_destroy(self._h)
            self._h = None

    def add(self, b, set_, load, sink, dir3=(0.0, -1.0, 0.0), point=(), warm=None):
        set_ = np.ascontiguousarray(set_, np.uint32)
        load = np.ascontiguousarray(load, np.float32)
        sink = np.ascontiguousarray(sink, np.uint8)
        d = np.ascontiguousarray(dir3, np.float32)
        pt = np.ascontiguousarray(point, np.float32).reshape(-1, 4)
        self._keep.append(b)
        return check(lib().sf_net_add(self._h, b._h, len(set_), ptr(set_), ptr(load), ptr(sink), ptr(d), len(pt),
                                      ptr(pt) if len(pt) else None, warm._h if warm is not None else None))

    def solve(self):
        check(lib().sf_net_solve(self._h))

    def clear(self):
        check(lib().sf_net_clear(self._h))
        self._keep = []

    def set_host_threads(self, n):
        check(lib().sf_net_set_host_threads(self._h, int(n)))

    def set_route(self, min_cells=-1, max_cases=16, block_its=6):
        """a case of at least min_cells cells in a batch of at most max_cases is solved a kernel per step over the
        whole chip (min_cells -1, the default: every case in one workgroup). The kernel-per-step route is switched
        off: it made a GPU memory fault whose cause is not found yet."""
        check(lib().sf_net_set_route(self._h, int(min_cells), int(max_cases), int(block_its)))

    def ms(self):
        """[host build, upload and read-back, device solve, device checks] of the last solve, ms"""
        out = np.zeros(4, np.float64)
        check(lib().sf_net_ms(self._h, ptr(out)))
        return out

    def host_ms(self):
        """the host build of the last solve summed over its cases: cells and links, levels, loads and checks; packing"""
        out = np.zeros(4, np.float64)
        check(lib().sf_net_host_ms(self._h, ptr(out)))
        return out

    def result(self, c):
        L = lib()
        sc = np.zeros(12, np.float64)
        check(L.sf_net_result(self._h, int(c), ptr(sc)))
        lv = np.zeros(8, np.int32)
        nlv = check(L.sf_net_levels(self._h, int(c), ptr(lv)))
        nj = check(L.sf_net_joints(self._h, int(c), None, None, 0))
        jid = np.zeros(max(nj, 1), np.int32)
        jr = np.zeros(max(nj, 1), np.float32)
        check(L.sf_net_joints(self._h, int(c), ptr(jid), ptr(jr), nj))
        ni = check(L.sf_net_inner(self._h, int(c), None, None, 0))
        bat = np.zeros((max(ni, 1), 3), np.int32)
        ir = np.zeros(max(ni, 1), np.float32)
        check(L.sf_net_inner(self._h, int(c), ptr(bat), ptr(ir), ni))
        n = check(L.sf_net_cells(self._h, int(c), None, None, 0))
        cells = np.zeros((max(n, 1), 7), np.int32)
        x = np.zeros((max(n, 1), 6), np.float64)
        check(L.sf_net_cells(self._h, int(c), ptr(cells), ptr(x), n))
        jid, jr, bat, ir = jid[:nj], jr[:nj], bat[:ni], ir[:ni]
        ratios = [(int(j), float(r)) for j, r in zip(jid, jr)]
        cracks = [Crack(int(q[0]), int(q[1]), int(q[2]), float(r)) for q, r in zip(bat, ir) if r > 1.0]
        inner = [(int(q[0]), int(q[1]), int(q[2]), float(r)) for q, r in zip(bat, ir) if 0.7 < r <= 1.0]
        return Analysis(failed=[(j, r) for j, r in ratios if r > 1.0], cracks=cracks, max_ratio=np.float32(sc[0]),
                        iterations=int(sc[1]), nodes=int(sc[2]),
                        worst_in=(np.float32(sc[5]), int(sc[6]), int(sc[7]), np.float32(sc[8]), np.float32(sc[9]), 0.0),
                        ratios=ratios, inner=inner, links=int(sc[3]), levels=lv[:nlv].tolist(), joint_ids=jid,
                        joint_ratios=jr, inner_faces=bat, inner_ratios=ir, cells=cells[:n], x=x[:n])


def analyse_warm(b, set_, load, point, sink, dir3, warm):
    """structure.rs analyse_warm (:191): analyse `set_` under per-brick loads (N) along dir3, supported at the sink
    bricks, plus point loads (index into set_, x, z in ticks, N), starting from and keeping `warm`"""
    n = Net()
    n.add(b, set_, load, sink, dir3, point, warm)
    n.solve()
    return n.result(0)


def analyse_at(b, set_, load, point, sink, dir3):
    return analyse_warm(b, set_, load, point, sink, dir3, None)


def analyse(b, set_, load, sink, dir3=(0.0, -1.0, 0.0)):
    return analyse_warm(b, set_, load, (), sink, dir3, None)


def analyse_many(cases):
    """Each case (b, set, load, sink[, dir3[, point[, warm]]]) analysed, all in one launch; the Analyses in order."""
    n = Net()
    for c in cases:
        n.add(*c)
    n.solve()
    return [n.result(k) for k in range(len(cases))]

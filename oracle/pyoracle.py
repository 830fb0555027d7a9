"""ctypes binding of the CPU oracle (oracle/liboracle.so).

TEST INFRASTRUCTURE ONLY: imported by tests/, by the smoke check and by the benchmark's CPU-baseline leg, always as
the checker. The product package never imports it.
"""
import ctypes
import os

import numpy as np

_HERE = os.path.dirname(os.path.abspath(__file__))
LIB_PATH = os.path.join(_HERE, "liboracle.so")
ORDER_REFERENCE, ORDER_COLOURED = 0, 1
BODY_F, SHAPE_F, MAN_F = 20, 13, 40

_lib = None
vp, ci, cf, cd = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_double


class _OlStruct(ctypes.Structure):
    _fields_ = [("n", ci), ("mat", vp), ("xyz", vp), ("nb6", vp), ("bond", vp), ("ext", vp)]


def lib():
    global _lib
    if _lib is None:
        if not os.path.exists(LIB_PATH):
            raise RuntimeError(f"{LIB_PATH} missing: run make -C oracle")
        L = ctypes.CDLL(LIB_PATH)
        protos = {
            "ob_create": (vp, [ci]), "ob_destroy": (None, [vp]), "ob_set_params": (None, [vp, vp, ci]),
            "ob_load": (ci, [vp, ci, vp, ci, vp, vp]), "ob_step": (None, [vp, cf]), "ob_nbody": (ci, [vp]),
            "ob_read_bodies": (None, [vp, vp]), "ob_stats": (None, [vp, vp]), "ob_reduced": (ci, [vp]), "ob_set_hulls": (ci, [vp, ci, vp, vp, vp, vp, vp, vp, vp]), "ob_pair_contact": (ci, [vp, ci, vp, vp, cf, vp]), "ob_candidates": (ci, [vp, vp, ci]),
            "ob_manifolds": (ci, [vp, vp, vp, ci]), "ob_colours": (ci, [vp, vp, ci]), "ob_islands": (None, [vp, vp]), "ob_levels": (None, [vp, vp]),
            "ob_particles_capacity": (ci, [vp, ci]), "ob_add_particles": (ci, [vp, ci, vp, vp]),
            "ob_kill_particle": (None, [vp, ci]), "ob_blast_particles": (None, [vp, cf, cf, cf, cf, cf]),
            "ob_wake_particles": (None, [vp]), "ob_read_particles": (ci, [vp, vp]), "ob_particle_stats": (None, [vp, vp]),
            "ob_bruteforce_pairs": (ci, [vp, cf, vp, ci]),
            "ol_sor_sweep": (cf, [vp, vp, cf]), "ol_solve64": (ci, [vp, vp, cd, ci]), "ol_residual64": (cd, [vp, vp]),
            "ol_verdicts64": (None, [vp, vp, vp, vp, vp]), "ol_verdicts32": (None, [vp, vp, vp, vp]),
            "ol_grounded": (ci, [vp, vp]), "ol_finalize": (None, [ci, vp, vp, vp, vp, vp]),
            "os_create": (vp, []), "os_destroy": (None, [vp]), "os_len": (ci, [vp]),
            "os_add": (ci, [vp, vp, vp, ci, ci, ci]), "os_set_flags": (None, [vp, ci, ci]), "os_flags": (ci, [vp, ci]),
            "os_remove": (None, [vp, ci]), "os_build_joints": (ci, [vp]), "os_njoint_slots": (ci, [vp]),
            "os_live_joints": (ci, [vp]), "os_joints": (ci, [vp, vp, ci]), "os_set_joint_damage": (None, [vp, ci, ci]),
            "os_mass": (cf, [vp, ci]), "os_warm_create": (vp, []), "os_warm_destroy": (None, [vp]),
            "os_warm_size": (ci, [vp]),
            "os_analyse": (vp, [vp, ci, vp, vp, vp, vp, ci, vp, ci, vp]), "os_result_free": (None, [vp]),
            "os_result_scalars": (None, [vp, vp]), "os_result_levels": (None, [vp, vp]),
            "os_result_joints": (ci, [vp, vp, vp, ci]), "os_result_inner": (ci, [vp, vp, vp, ci]),
            "os_result_nodes": (ci, [vp, vp, vp, vp, vp, ci]), "os_result_springs": (ci, [vp, vp, vp, ci]),
        }
        for k, (r, a) in protos.items():
            f = getattr(L, k)
            f.restype = r
            f.argtypes = a
        _lib = L
    return _lib


def _p(a):
    return None if a is None else a.ctypes.data_as(vp)


class BoxOracle:
    """The reference box step on the CPU (engine/src/phys.rs restated, compound shapes)."""

    def __init__(self, bodies, shape_body, shapes, prm=None, order=ORDER_COLOURED, hulls=None):
        """hulls: the convex-chunk tables (v4, pl4, dir4, fidx, fcnt, seg, head), needed when a shape has a hull"""
        L = lib()
        self._h = L.ob_create(order)
        if prm is not None:
            prm = np.ascontiguousarray(prm, np.float32)
            L.ob_set_params(self._h, _p(prm), len(prm))
        if hulls is not None:
            t = [np.ascontiguousarray(x) for x in hulls]
            self._hulls = t
            L.ob_set_hulls(self._h, len(t[6]) // 4, *[_p(x) for x in t])
        b = np.ascontiguousarray(bodies, np.float32)
        sb = np.ascontiguousarray(shape_body, np.int32)
        s = np.ascontiguousarray(shapes, np.float32)
        rc = L.ob_load(self._h, len(b), _p(b), len(s), _p(sb), _p(s))
        if rc != 0:
            raise ValueError(f"ob_load failed {rc}")
        self.nb = len(b)

    def __del__(self):
        if getattr(self, "_h", None):
            lib().ob_destroy(self._h)
            self._h = None

    def step(self, dt=1.0 / 60.0):
        lib().ob_step(self._h, ctypes.c_float(dt))

    def pair_contact(self, a, b, margin, chunk_path=True):
        """(normal, points [k, 4]: p, separation) of one pair before reduce4; a, b: p[3] q[4] h[3] hull (b None: ground)"""
        fa = np.ascontiguousarray(a, np.float32)
        fb = None if b is None else np.ascontiguousarray(b, np.float32)
        out = np.zeros(4 + 32, np.float32)
        n = lib().ob_pair_contact(self._h, 1 if chunk_path else 0, _p(fa), _p(fb), ctypes.c_float(margin), _p(out))
        return out[0:3].copy(), out[4:4 + 4 * n].reshape(n, 4).copy()

    def bodies(self):
        out = np.zeros((self.nb, BODY_F), np.float32)
        lib().ob_read_bodies(self._h, _p(out))
        return out

    def stats(self):
        out = np.zeros(8, np.int32)
        lib().ob_stats(self._h, _p(out))
        d = dict(zip(["awake", "pairs", "manifolds", "points", "colours", "overflow", "ground", "wakes"], out.tolist()))
        d["reduced"] = int(lib().ob_reduced(self._h))
        return d

    def candidates(self):
        n = lib().ob_candidates(self._h, None, 0)
        out = np.zeros((max(n, 1), 2), np.int32)
        lib().ob_candidates(self._h, _p(out), n)
        return out[:n]

    def manifolds(self):
        n = lib().ob_manifolds(self._h, None, None, 0)
        keys = np.zeros((max(n, 1), 3), np.int32)
        data = np.zeros((max(n, 1), MAN_F), np.float32)
        lib().ob_manifolds(self._h, _p(keys), _p(data), n)
        return keys[:n], data[:n]

    def colours(self):
        n = lib().ob_colours(self._h, None, 0)
        out = np.zeros(max(n, 1), np.int32)
        lib().ob_colours(self._h, _p(out), n)
        return out[:n]

    def islands(self):
        out = np.zeros(self.nb, np.int32)
        lib().ob_islands(self._h, _p(out))
        return out

    def levels(self):
        out = np.zeros(self.nb, np.int32)
        lib().ob_levels(self._h, _p(out))
        return out

    # particle debris
    def particles_capacity(self, cap):
        assert lib().ob_particles_capacity(self._h, int(cap)) == 0

    def add_particles(self, rec):
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 8)
        slots = np.zeros(len(rec), np.int32)
        lib().ob_add_particles(self._h, len(rec), _p(rec), _p(slots))
        return slots

    def kill_particle(self, i):
        lib().ob_kill_particle(self._h, int(i))

    def blast_particles(self, c, reach, speed):
        lib().ob_blast_particles(self._h, ctypes.c_float(c[0]), ctypes.c_float(c[1]), ctypes.c_float(c[2]),
                                 ctypes.c_float(reach), ctypes.c_float(speed))

    def wake_particles(self):
        lib().ob_wake_particles(self._h)

    def particles(self):
        """12 floats per slot: x y z vx, vy vz qx qy, qz qw state radius"""
        st = self.particle_stats()
        out = np.zeros((max(st[0], 1), 12), np.float32)
        lib().ob_read_particles(self._h, _p(out))
        return out[:st[0]]

    def particle_stats(self):
        out = np.zeros(4, np.int32)
        lib().ob_particle_stats(self._h, _p(out))
        return out

    def bruteforce_pairs(self, dt=1.0 / 60.0, cap=4_000_000):
        out = np.zeros((cap, 2), np.int32)
        n = lib().ob_bruteforce_pairs(self._h, ctypes.c_float(dt), _p(out), cap)
        return out[:n]


class LoadOracle:
    """The legacy load-path solve of one standing structure on the CPU (src/lib.rs restated)."""

    def __init__(self, mat, xyz, infill=None, ext=None):
        self.mat = np.ascontiguousarray(mat, np.uint8)
        self.xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        self.n = len(self.mat)
        self.infill = np.zeros(self.n, np.uint8) if infill is None else np.ascontiguousarray(infill, np.uint8)
        self.ext = None if ext is None else np.ascontiguousarray(ext, np.float32)
        self.nb6 = np.zeros((self.n, 6), np.int32)
        self.bond = np.zeros((self.n, 4), np.float32)
        lib().ol_finalize(self.n, _p(self.mat), _p(self.xyz), _p(self.infill), _p(self.nb6), _p(self.bond))
        self._s = _OlStruct(self.n, _p(self.mat).value, _p(self.xyz).value, _p(self.nb6).value, _p(self.bond).value,
                            _p(self.ext).value if self.ext is not None else None)

    def _sp(self):
        return ctypes.byref(self._s)

    def sor_sweep(self, phi32, omega):
        return lib().ol_sor_sweep(self._sp(), _p(phi32), ctypes.c_float(omega))

    def solve64(self, rel_tol=1e-12, max_iter=200000):
        phi = np.zeros(self.n, np.float64)
        it = lib().ol_solve64(self._sp(), _p(phi), ctypes.c_double(rel_tol), max_iter)
        return phi, it

    def residual64(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        return lib().ol_residual64(self._sp(), _p(phi))

    def verdicts64(self, phi):
        phi = np.ascontiguousarray(phi, np.float64)
        r = np.zeros((self.n, 4), np.float64)
        m = np.zeros((self.n, 4), np.uint8)
        f = np.zeros((self.n, 4), np.uint8)
        lib().ol_verdicts64(self._sp(), _p(phi), _p(r), _p(m), _p(f))
        return r, m, f

    def verdicts32(self, phi32):
        phi32 = np.ascontiguousarray(phi32, np.float32)
        r = np.zeros((self.n, 4), np.float32)
        m = np.zeros((self.n, 4), np.uint8)
        lib().ol_verdicts32(self._sp(), _p(phi32), _p(r), _p(m))
        return r, m

    def grounded(self):
        out = np.zeros(self.n, np.uint8)
        lib().ol_grounded(self._sp(), _p(out))
        return out


ORDER_SEQ = 0
ORDER_DEV = 1
LL_BODY_F = 22
LL_BRICK_F = 13
LL_CONTACT_F = 18


def _lattice_protos():
    L = lib()
    if getattr(L, "_ll_done", False):
        return L
    vp, i32, f32, u32 = ctypes.c_void_p, ctypes.c_int, ctypes.c_float, ctypes.c_uint32
    L.ll_create.restype = vp
    L.ll_create.argtypes = [i32, i32, i32, i32, i32, u32]
    L.ll_destroy.argtypes = [vp]
    L.ll_add_structure.argtypes = [vp, i32, vp, vp, vp]
    L.ll_add_body.argtypes = [vp, i32, vp, vp, vp, vp]
    for f in ("ll_finalize", "ll_warm_all", "ll_wake_all"):
        getattr(L, f).argtypes = [vp]
    L.ll_step.argtypes = [vp, f32]
    L.ll_explode.argtypes = [vp, f32, f32, f32, f32, f32]
    for f in ("ll_stats", "ll_bodies", "ll_bricks", "ll_structs", "ll_matstats"):
        getattr(L, f).argtypes = [vp, vp]
    for f in ("ll_nbricks", "ll_nbodies", "ll_nstructs"):
        getattr(L, f).argtypes = [vp]
    L.ll_contacts.argtypes = [vp, vp, i32]
    L.ll_rng.argtypes = [vp]
    L.ll_rng.restype = u32
    L.ll_set_field.argtypes = [vp, i32, vp]
    L.ll_failures.argtypes = [vp, vp]
    L.ll_set_body.argtypes = [vp, i32, vp, vp, vp, vp]
    L.ll_probe_contacts.argtypes = [vp]
    L.ll_probe_contacts.restype = i32
    L.ll_probe_prep.argtypes = [vp, f32]
    L.ll_probe_solve.argtypes = [vp, i32]
    L.ll_probe_pinch.argtypes = [vp, i32, vp, vp]
    L.ll_probe_crush.argtypes = [vp, i32, f32, vp, u32]
    L.ll_probe_crush.restype = u32
    L.ll_set_last.argtypes = [vp, i32, vp, vp]
    L.ll_probe_structural.argtypes = [vp, f32]
    L.ll_set_nc.argtypes = [vp, i32, u32]
    L.ll_set_asleep.argtypes = [vp, i32]
    L._ll_done = True
    return L


class LatticeOracle:
    """The legacy brick engine's physics frame on the CPU (src/lib.rs restated; oracle/lattice.c)."""

    def __init__(self, order=ORDER_DEV, brick_cap=1 << 16, body_cap=1 << 14, colours=16, rounds=8, seed=0x9E3779B9):
        L = _lattice_protos()
        self._h = L.ll_create(int(order), int(brick_cap), int(body_cap), int(colours), int(rounds), int(seed))

    def close(self):
        if getattr(self, "_h", None):
            lib().ll_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_structure(self, mat, xyz, infill=None):
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32)
        infill = np.zeros(len(mat), np.uint8) if infill is None else np.ascontiguousarray(infill, np.uint8)
        return lib().ll_add_structure(self._h, len(mat), _p(xyz), _p(mat), _p(infill))

    def add_body(self, mat, xyz, v=(0, 0, 0), w=(0, 0, 0)):
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32)
        v = np.asarray(v, np.float32)
        w = np.asarray(w, np.float32)
        return lib().ll_add_body(self._h, len(mat), _p(xyz), _p(mat), _p(v), _p(w))

    def finalize(self):
        lib().ll_finalize(self._h)

    def warm_all(self):
        lib().ll_warm_all(self._h)

    def wake_all(self):
        lib().ll_wake_all(self._h)

    def step(self, dt=1.0 / 60.0):
        lib().ll_step(self._h, ctypes.c_float(dt))

    def explode(self, x, y, z, r, power):
        lib().ll_explode(self._h, *(ctypes.c_float(v) for v in (x, y, z, r, power)))

    def stats(self):
        out = np.zeros(16, np.float32)
        lib().ll_stats(self._h, _p(out))
        return out

    def bodies(self):
        n = lib().ll_nbodies(self._h)
        out = np.zeros((max(n, 1), LL_BODY_F), np.float32)
        lib().ll_bodies(self._h, _p(out))
        return out[:n]

    def bricks(self):
        n = lib().ll_nbricks(self._h)
        out = np.zeros((max(n, 1), LL_BRICK_F), np.float32)
        lib().ll_bricks(self._h, _p(out))
        return out[:n]

    def structs(self):
        n = lib().ll_nstructs(self._h)
        out = np.zeros((max(n, 1), 12), np.float32)
        lib().ll_structs(self._h, _p(out))
        return out[:n]

    def matstats(self):
        out = np.zeros(50, np.float32)
        lib().ll_matstats(self._h, _p(out))
        return out

    def contacts(self):
        n = lib().ll_contacts(self._h, None, 0)
        out = np.zeros((max(n, 1), LL_CONTACT_F), np.float32)
        lib().ll_contacts(self._h, _p(out), n)
        return out[:n]

    def rng(self):
        return lib().ll_rng(self._h)

    def set_field(self, s, phi):
        phi = np.ascontiguousarray(phi, np.float32)
        lib().ll_set_field(self._h, int(s), _p(phi))

    def failures(self):
        out = np.zeros(4, np.uint32)
        lib().ll_failures(self._h, _p(out))
        return out

    # ---- test hooks: one piece of a substep on its own
    @staticmethod
    def _f(v, n):
        return None if v is None else np.ascontiguousarray(np.asarray(v, np.float32).reshape(n))

    def set_body(self, b, pos=None, q=None, v=None, w=None):
        a = [self._f(pos, 3), self._f(q, 4), self._f(v, 3), self._f(w, 3)]
        lib().ll_set_body(self._h, int(b), *(_p(x) for x in a))

    def probe_contacts(self):
        lib().ll_probe_contacts(self._h)
        return self.contacts()

    def probe_prep(self, h):
        lib().ll_probe_prep(self._h, ctypes.c_float(h))

    def probe_solve(self, passes=1):
        lib().ll_probe_solve(self._h, int(passes))

    def probe_pinch(self, b, vpre, wpre):
        lib().ll_probe_pinch(self._h, int(b), _p(self._f(vpre, 3)), _p(self._f(wpre, 3)))

    def probe_crush(self, brick, e, p, cap=2500):
        return lib().ll_probe_crush(self._h, int(brick), ctypes.c_float(e), _p(self._f(p, 3)), int(cap))

    def set_last(self, b, v, w):
        lib().ll_set_last(self._h, int(b), _p(self._f(v, 3)), _p(self._f(w, 3)))

    def probe_structural(self, dt=1.0 / 60.0):
        lib().ll_probe_structural(self._h, ctypes.c_float(dt))

    def set_nc(self, b, n):
        lib().ll_set_nc(self._h, int(b), int(n))

    def set_asleep(self, b):
        lib().ll_set_asleep(self._h, int(b))


ORDER_DEVICE = 1


class BricksOracle:
    """The brick store of engine/src/brick.rs on the CPU (whole boxes on the 0.125 m tick grid, joints as shared
    faces), for the structural analysis oracle."""

    def __init__(self):
        self._h = lib().os_create()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().os_destroy(self._h)
            self._h = None

    def __len__(self):
        return lib().os_len(self._h)

    def add(self, mn, size, mat, paint=0, structure=0):
        mn = np.ascontiguousarray(mn, np.int32)
        size = np.ascontiguousarray(size, np.int32)
        i = lib().os_add(self._h, _p(mn), _p(size), int(mat), int(paint), int(structure))
        if i < 0:
            raise ValueError(f"brick size {size.tolist()} outside 1..255 ticks")
        return i

    def set_flags(self, i, flags):
        lib().os_set_flags(self._h, int(i), int(flags))

    def flags(self, i):
        return lib().os_flags(self._h, int(i))

    def remove(self, i):
        lib().os_remove(self._h, int(i))

    def build_joints(self):
        return lib().os_build_joints(self._h)

    def joints(self):
        n = lib().os_njoint_slots(self._h)
        out = np.zeros((max(n, 1), 5), np.int32)
        lib().os_joints(self._h, _p(out), n)
        return out[:n]

    def set_joint_damage(self, j, dmg):
        lib().os_set_joint_damage(self._h, int(j), int(dmg))

    def mass(self, i):
        return np.float32(lib().os_mass(self._h, int(i)))


class WarmOracle:
    def __init__(self):
        self._h = lib().os_warm_create()

    def __del__(self):
        if getattr(self, "_h", None):
            lib().os_warm_destroy(self._h)
            self._h = None

    def __len__(self):
        return lib().os_warm_size(self._h)


def analyse_oracle(b, set_, load, sink, dir3=(0.0, -1.0, 0.0), point=(), order=ORDER_REFERENCE, warm=None):
    """structure.rs analyse_warm on the CPU; a dict of the analysis, plus the cells, displacements and springs."""
    set_ = np.ascontiguousarray(set_, np.uint32)
    load = np.ascontiguousarray(load, np.float32)
    sink = np.ascontiguousarray(sink, np.uint8)
    d = np.ascontiguousarray(dir3, np.float32)
    pt = np.ascontiguousarray(point, np.float32).reshape(-1, 4)
    L = lib()
    R = L.os_analyse(b._h, len(set_), _p(set_), _p(load), _p(sink), _p(d), len(pt), _p(pt) if len(pt) else None,
                     int(order), warm._h if warm is not None else None)
    try:
        sc = np.zeros(12, np.float64)
        L.os_result_scalars(R, _p(sc))
        lv = np.zeros(8, np.int32)
        L.os_result_levels(R, _p(lv))
        nj, ni, n, nl = int(sc[10]), int(sc[11]), int(sc[2]), int(sc[3])
        jid = np.zeros(max(nj, 1), np.int32)
        jr = np.zeros(max(nj, 1), np.float32)
        L.os_result_joints(R, _p(jid), _p(jr), nj)
        bat = np.zeros((max(ni, 1), 3), np.int32)
        ir = np.zeros(max(ni, 1), np.float32)
        L.os_result_inner(R, _p(bat), _p(ir), ni)
        node = np.zeros((max(n, 1), 7), np.int32)
        nload = np.zeros(max(n, 1), np.float32)
        fixed = np.zeros(max(n, 1), np.uint8)
        x = np.zeros((max(n, 1), 6), np.float64)
        L.os_result_nodes(R, _p(node), _p(nload), _p(fixed), _p(x), n)
        ends = np.zeros((max(nl, 1), 2), np.uint32)
        sp = np.zeros((max(nl, 1), 12), np.float64)
        L.os_result_springs(R, _p(ends), _p(sp), nl)
    finally:
        L.os_result_free(R)
    return dict(max_ratio=np.float32(sc[0]), iterations=int(sc[1]), nodes=n, links=nl, levels=lv[:int(sc[4])].tolist(),
                worst_in=(np.float32(sc[5]), int(sc[6]), int(sc[7]), np.float32(sc[8]), np.float32(sc[9])),
                joint_ids=jid[:nj], joint_ratios=jr[:nj], inner=bat[:ni], inner_ratios=ir[:ni],
                node=node[:n], node_load=nload[:n], fixed=fixed[:n].astype(bool), x=x[:n], ends=ends[:nl], springs=sp[:nl])

"""Box world: the host-side mirror of the reference's box-core API (engine/src/phys.rs World, engine/src/wasm.rs
exports) over the GPU step in libsf_hip.

    w = World.from_scene(*scenes.pile(2000), substeps=4, sleep=False)
    for _ in range(120): w.step(1 / 60)
    pose = w.pose()          # [nb, 10]: position, rotation, half extents (wasm.rs pose)
"""
import ctypes

import numpy as np

from ._native import lib, check, ptr

# parameter indices (include/sf.h SF_P_*)
P = dict(GX=0, GY=1, GZ=2, SUBSTEPS=3, HERTZ=4, DAMPING=5, MAXPUSH=6, FRICTION=7, SPEC=8, SLEEP_LIN=9, SLEEP_ANG=10,
         SLEEP_TIME=11, SLEEP_DIST=12, SLEEP_ROT=13, ALLOW_SLEEP=14, CELL=15, LEAK=16, ANCHOR=17, ANCHOR_PUSH=18,
         ANCHOR_SLOP=19, COL_LIMIT=20, COL_ROUNDS=21, COL_SPREAD=22, TAIL_MIN=23, SOLVE_NT=24, RATIO=25, SHOCK=26,
         REST_THRESHOLD=27, REST_ITERS=28, COL_SUPPORT=29, PSUB=30, PCELL=31, RED_K=32)
NPARAM = 33
# the GPU world's solver extensions (gpu/gpu_world.js:1152-1203, :954-985, :834-920): mass-ratio limit,
# shock-propagation level passes, supports-first colouring rounds, contact reduction; params(**GPU_WORLD) turns them on
GPU_WORLD = dict(ratio=20.0, shock=6, col_support=2, red_k=32)
# statistics (SF_S_*)
S = dict(AWAKE=0, PAIRS=1, MANIFOLDS=2, POINTS=3, COLOURS=4, OVERFLOW=5, GROUND=6, WAKES=7, ENTRIES=8, CANDIDATES=9,
         DROPPED=10, ROUNDS=11, REDUCED=12)
NSTAT = 13
MAN_F = 40


def default_params():
    """Params::default (engine/src/phys.rs:101-105) plus the colouring limits; the GPU world's extensions off."""
    p = np.zeros(NPARAM, np.float32)
    p[[P["GX"], P["GY"], P["GZ"]]] = [0.0, -9.81, 0.0]
    p[P["SUBSTEPS"]] = 12
    p[P["DAMPING"]] = 10.0
    p[P["MAXPUSH"]] = 3.0
    p[P["FRICTION"]] = 0.6
    p[P["SPEC"]] = 0.02
    p[P["SLEEP_LIN"]] = 0.05
    p[P["SLEEP_ANG"]] = 0.08
    p[P["SLEEP_TIME"]] = 0.5
    p[P["SLEEP_DIST"]] = 0.01
    p[P["SLEEP_ROT"]] = 0.02
    p[P["ALLOW_SLEEP"]] = 1.0
    p[P["CELL"]] = 1.0
    p[P["LEAK"]] = 1.0
    p[P["ANCHOR"]] = 1.0
    p[P["ANCHOR_PUSH"]] = 3.0
    p[P["COL_LIMIT"]] = 16
    p[P["COL_ROUNDS"]] = 32
    p[P["COL_SPREAD"]] = 4
    p[P["TAIL_MIN"]] = 1024
    p[P["SOLVE_NT"]] = 64
    p[P["REST_THRESHOLD"]] = 1.0
    p[P["REST_ITERS"]] = 8
    p[P["PSUB"]] = 4
    p[P["PCELL"]] = 0.7
    return p


def _full(prm):
    """prm as NPARAM float32s, defaults after the ones given (parameters added later keep their defaults)"""
    prm = np.ascontiguousarray(prm, np.float32)
    if len(prm) < NPARAM:
        full = default_params()
        full[:len(prm)] = prm
        prm = full
    return prm


def params(substeps=12, sleep=True, **kw):
    p = default_params()
    p[P["SUBSTEPS"]] = substeps
    p[P["ALLOW_SLEEP"]] = 1.0 if sleep else 0.0
    for k, v in kw.items():
        p[P[k.upper()]] = v
    return p


class World:
    """One box-physics world on the GPU (a handle of the C ABI)."""

    def __init__(self, body_cap, shape_cap, prm=None):
        L = lib()
        prm = default_params() if prm is None else _full(prm)
        self._prm = prm
        self._h = L.sf_world_create(int(body_cap), int(shape_cap), ptr(prm))
        if not self._h:
            check(-1)
        self.nb = 0
        self.ns = 0

    @classmethod
    def from_scene(cls, bodies, shape_body, shapes, prm=None, substeps=None, sleep=None, headroom=1.0):
        if prm is None:
            prm = default_params()
        prm = np.array(prm, np.float32)
        if substeps is not None:
            prm[P["SUBSTEPS"]] = substeps
        if sleep is not None:
            prm[P["ALLOW_SLEEP"]] = 1.0 if sleep else 0.0
        w = cls(max(1, int(len(bodies) * headroom)), max(1, int(len(shapes) * headroom)), prm)
        w.load(bodies, shape_body, shapes)
        return w

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_world_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def params(self):
        """A copy of the world's parameters."""
        return self._prm.copy()

    def set_params(self, prm):
        prm = _full(prm)
        self._prm = prm
        check(lib().sf_world_set_params(self._h, ptr(prm), len(prm)))

    def load(self, bodies, shape_body, shapes):
        bodies = np.ascontiguousarray(bodies, np.float32)
        shape_body = np.ascontiguousarray(shape_body, np.int32)
        shapes = np.ascontiguousarray(shapes, np.float32)
        check(lib().sf_world_load(self._h, len(bodies), ptr(bodies), len(shapes), ptr(shape_body), ptr(shapes)))
        self.nb, self.ns = len(bodies), len(shapes)

    def step(self, dt=1.0 / 60.0):
        """One step; returns the bodies in the solver (phys.rs World::step, wasm.rs step)."""
        return check(lib().sf_world_step(self._h, ctypes.c_float(dt)))

    def step_many(self, n, dt=1.0 / 60.0):
        check(lib().sf_world_step_many(self._h, ctypes.c_float(dt), int(n)))

    def bodies(self):
        out = np.zeros((self.nb, 20), np.float32)
        check(lib().sf_world_read_bodies(self._h, ptr(out)))
        return out

    def pose(self):
        """[nb, 10]: position, rotation (x, y, z, w), half extents of the first shape (wasm.rs:47-53)."""
        p = lib().sf_world_pose(self._h)
        if not p:
            check(-1)
        return np.ctypeslib.as_array(p, shape=(self.nb * 10,)).reshape(self.nb, 10).copy()

    def stats(self):
        out = np.zeros(NSTAT, np.int32)
        check(lib().sf_world_stats(self._h, ptr(out)))
        return {k.lower(): int(out[v]) for k, v in S.items()}

    def candidates(self):
        n = check(lib().sf_world_candidates(self._h, None, 0))
        out = np.zeros((max(n, 1), 2), np.int32)
        check(lib().sf_world_candidates(self._h, ptr(out), n))
        return out[:n]

    def manifolds(self):
        n = check(lib().sf_world_manifolds(self._h, None, None, 0))
        keys = np.zeros((max(n, 1), 3), np.int32)
        data = np.zeros((max(n, 1), MAN_F), np.float32)
        check(lib().sf_world_manifolds(self._h, ptr(keys), ptr(data), n))
        return keys[:n], data[:n]

    def colours(self):
        n = check(lib().sf_world_colours(self._h, None, 0))
        out = np.zeros(max(n, 1), np.int32)
        check(lib().sf_world_colours(self._h, ptr(out), n))
        return out[:n]

    def islands(self):
        out = np.zeros(self.nb, np.int32)
        check(lib().sf_world_islands(self._h, ptr(out)))
        return out

    def set_timing(self, on=True):
        check(lib().sf_world_set_timing(self._h, 1 if on else 0))

    def set_kernel_timing(self, on=True):
        check(lib().sf_world_set_kernel_timing(self._h, 1 if on else 0))

    def kernel_stats(self):
        """(ms, launches, manifold solves) of the colour-bucket solve kernel since kernel timing was switched on."""
        ms = np.zeros(1, np.float64)
        n = np.zeros(1, np.int64)
        it = np.zeros(1, np.int64)
        check(lib().sf_world_kernel_stats(self._h, ptr(ms), ptr(n), ptr(it)))
        return float(ms[0]), int(n[0]), int(it[0])

    def set_graph(self, on=True):
        """Replay the sub-steps' launches as one captured HIP graph (on by default; off: launched one by one)."""
        check(lib().sf_world_set_graph(self._h, 1 if on else 0))

    def graph_stats(self):
        """(captures, replays) of the sub-step graph since the world was made."""
        c = np.zeros(1, np.int64)
        r = np.zeros(1, np.int64)
        check(lib().sf_world_graph_stats(self._h, ptr(c), ptr(r)))
        return int(c[0]), int(r[0])

    def stage_ms(self):
        out = np.zeros(5, np.float32)
        check(lib().sf_world_stage_ms(self._h, ptr(out)))
        return dict(zip(["broad", "narrow", "colour", "solve", "sleep"], out.tolist()))

    # ---- particle debris (gpu/gpu_world.js addParticles/killParticle/blastParticles/readParticles :2863-2896)
    def particles_capacity(self, cap):
        check(lib().sf_world_particles_capacity(self._h, int(cap)))

    def add_particles(self, rec):
        """rec: [n, 8] float32 (p, v, radius, friction); returns the slots (-1: full or not finite)"""
        rec = np.ascontiguousarray(rec, np.float32).reshape(-1, 8)
        slots = np.zeros(len(rec), np.int32)
        check(lib().sf_world_add_particles(self._h, len(rec), ptr(rec), ptr(slots)))
        return slots

    def kill_particle(self, i):
        check(lib().sf_world_kill_particle(self._h, int(i)))

    def blast_particles(self, c, reach, speed):
        check(lib().sf_world_blast_particles(self._h, float(c[0]), float(c[1]), float(c[2]), float(reach), float(speed)))

    def wake_particles(self):
        check(lib().sf_world_wake_particles(self._h))

    def particles(self):
        """[slots, 12]: x y z vx, vy vz qx qy, qz qw state radius"""
        n = check(lib().sf_world_read_particles(self._h, None))
        out = np.zeros((max(n, 1), 12), np.float32)
        if n:
            check(lib().sf_world_read_particles(self._h, ptr(out)))
        return out[:n]

    def particle_ms(self):
        """device ms of the last step's particle stage (timing on)"""
        ms = np.zeros(1, np.float32)
        check(lib().sf_world_particle_ms(self._h, ptr(ms)))
        return float(ms[0])

    def particle_stats(self):
        """[slots in use, moving at the last step's start, asleep at its end, free slots]"""
        out = np.zeros(4, np.int32)
        check(lib().sf_world_particle_stats(self._h, ptr(out)))
        return out


# ---- the reference's singleton exports (engine/src/wasm.rs:27-53)
def pile_setup(n, substeps):
    lib().sf_pile_setup(int(n), int(substeps))


def stack_setup(n, substeps, sleep):
    lib().sf_stack_setup(int(n), int(substeps), int(sleep))


def step(dt):
    return lib().sf_step(ctypes.c_float(dt))


def body_count():
    return lib().sf_body_count()


def manifolds():
    return lib().sf_manifolds()


def points():
    return lib().sf_points()


def max_speed():
    return lib().sf_max_speed()


def body_pos(i, axis):
    return lib().sf_body_pos(int(i), int(axis))


def pose():
    n = body_count()
    p = lib().sf_pose()
    return np.ctypeslib.as_array(p, shape=(n * 10,)).copy() if p else np.zeros(0, np.float32)

"""Lattice world: the legacy brick engine's physics frame on the GPU (src/lib.rs step :3639-3721), one or many
independent scenes per handle.

    L = Lattice(bricks=20000, bodies=4000)
    s = L.add_scene(bricks=20000, bodies=4000)
    L.add_structure(s, *scenes.tower())       # mat, xyz, infill
    L.finalize(s); L.warm_all(s)
    L.step()                                  # warm-up sweeps
    L.explode(s, x, 0.6, z, 2.2, 1.4)
    for _ in range(600): L.step()
"""
import ctypes

import numpy as np

from ._native import lib, check, ptr

BODY_F = 22
BRICK_F = 13
CONTACT_F = 18


class Lattice:
    """A handle of the lattice C ABI (sf_lat_*); scene indices are returned by add_scene."""

    def __init__(self, bricks, bodies, scenes=1, colours=16, rounds=8):
        self._h = lib().sf_lat_create(int(bricks), int(bodies), int(scenes), int(colours), int(rounds))
        if not self._h:
            check(-1)

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_lat_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add_scene(self, bricks, bodies, structs=8, seed=0x9E3779B9):
        return check(lib().sf_lat_scene_add(self._h, int(bricks), int(bodies), int(structs), int(seed) & 0xFFFFFFFF))

    def clear_scene(self, scene, seed=0x9E3779B9):
        check(lib().sf_lat_scene_clear(self._h, int(scene), int(seed) & 0xFFFFFFFF))

    def add_structure(self, scene, mat, xyz, infill=None):
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32)
        infill = np.zeros(len(mat), np.uint8) if infill is None else np.ascontiguousarray(infill, np.uint8)
        return check(lib().sf_lat_add_structure(self._h, int(scene), len(mat), ptr(xyz), ptr(mat), ptr(infill)))

    def add_body(self, scene, mat, xyz, v=(0, 0, 0), w=(0, 0, 0)):
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32)
        v = np.asarray(v, np.float32)
        w = np.asarray(w, np.float32)
        return check(lib().sf_lat_add_body(self._h, int(scene), len(mat), ptr(xyz), ptr(mat), ptr(v), ptr(w)))

    def finalize(self, scene):
        check(lib().sf_lat_finalize(self._h, int(scene)))

    def warm_all(self, scene):
        check(lib().sf_lat_warm_all(self._h, int(scene)))

    def wake_all(self, scene):
        check(lib().sf_lat_wake_all(self._h, int(scene)))

    def explode(self, scene, x, y, z, r, power):
        check(lib().sf_lat_explode(self._h, int(scene), *(ctypes.c_float(v) for v in (x, y, z, r, power))))

    def step(self, dt=1.0 / 60.0):
        check(lib().sf_lat_step(self._h, ctypes.c_float(dt)))

    def counts(self, scene):
        out = np.zeros(3, np.int32)
        check(lib().sf_lat_counts(self._h, int(scene), ptr(out)))
        return out

    def stats(self, scene):
        out = np.zeros(16, np.float32)
        check(lib().sf_lat_stats(self._h, int(scene), ptr(out)))
        return out

    def bodies(self, scene):
        n = int(self.counts(scene)[1])
        out = np.zeros((max(n, 1), BODY_F), np.float32)
        check(lib().sf_lat_bodies(self._h, int(scene), ptr(out)))
        return out[:n]

    def bricks(self, scene):
        n = int(self.counts(scene)[0])
        out = np.zeros((max(n, 1), BRICK_F), np.float32)
        check(lib().sf_lat_bricks(self._h, int(scene), ptr(out)))
        return out[:n]

    def structs(self, scene):
        n = int(self.counts(scene)[2])
        out = np.zeros((max(n, 1), 12), np.float32)
        check(lib().sf_lat_structs(self._h, int(scene), ptr(out)))
        return out[:n]

    def matstats(self, scene):
        out = np.zeros(50, np.float32)
        check(lib().sf_lat_matstats(self._h, int(scene), ptr(out)))
        return out

    def contacts(self, scene):
        n = check(lib().sf_lat_contacts(self._h, int(scene), None, 0))
        out = np.zeros((max(n, 1), CONTACT_F), np.float32)
        check(lib().sf_lat_contacts(self._h, int(scene), ptr(out), n))
        return out[:n]

    def rng(self, scene):
        out = np.zeros(1, np.uint32)
        check(lib().sf_lat_rng(self._h, int(scene), ptr(out)))
        return int(out[0])

    def set_field(self, scene, structure, phi):
        phi = np.ascontiguousarray(phi, np.float32)
        check(lib().sf_lat_set_field(self._h, int(scene), int(structure), ptr(phi)))

    def failures(self, scene):
        out = np.zeros(4, np.uint32)
        check(lib().sf_lat_failures(self._h, int(scene), ptr(out)))
        return out

    def phase_ms(self):
        """wall ms of the last frame: substeps, plough, impact, topology, structural, warm-up, statistics, all"""
        out = np.zeros(8, np.float64)
        check(lib().sf_lat_phase_ms(self._h, ptr(out)))
        return dict(zip(["substeps", "plough", "impact", "topology", "structural", "warmup", "stats", "frame"], out.tolist()))

    def subphase_ms(self):
        """finer split of the last frame's host work (ms; "scenes" is a count)"""
        out = np.zeros(16, np.float64)
        check(lib().sf_lat_subphase_ms(self._h, ptr(out)))
        keys = ["topo_pull", "topo_rebuild", "topo_conn", "topo_flush", "topo_csr", "topo_upload", "topo_scenes",
                "st_upload", "st_launch", "st_wait", "st_apply", "st_flush", "topo_reach"]
        return dict(zip(keys, out.tolist()))

    def set_conn_device(self, on):
        """connectivity's grounded search on the device (default) or on the host; results do not depend on it"""
        check(lib().sf_lat_set_conn_device(self._h, 1 if on else 0))

    def set_fused(self, max_contacts):
        """scenes of at most max_contacts contacts are solved by one workgroup each; 0: colour-by-colour launches"""
        check(lib().sf_lat_set_fused(self._h, int(max_contacts)))

    def set_host_threads(self, n):
        """worker threads of the per-scene host bookkeeping (0: the default); results do not depend on it"""
        check(lib().sf_lat_set_host_threads(self._h, int(n)))

    def set_sor_lds(self, max_bricks):
        """structures of at most max_bricks bricks are swept with their potential in LDS (packed rows); -1: the
        device's limit, 0: every structure from the full rows with the potential in global memory"""
        check(lib().sf_lat_set_sor_lds(self._h, int(max_bricks)))

    def timing(self):
        ms = np.zeros(1, np.float64)
        syncs = np.zeros(1, np.int32)
        check(lib().sf_lat_timing(self._h, ptr(ms), ptr(syncs)))
        return float(ms[0]), int(syncs[0])

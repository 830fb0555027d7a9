"""Served what-if stream: demolition queries answered by continuous batching on one GPU.

The server itself is the library's (sf_serve_*, shatterforge_amd/csrc/serve.cpp); this module generates the queries
(seeded frame/wall buildings of 2k-40k bricks with 1-8 charges at ground-floor column bases), drives a Poisson
arrival stream in real time and binds the server. A query is answered in three parts, all on the GPU:
  1. the load-path solve of the intact building (FP32 red-black sweeps, FP64 residual and refinement, sf_load);
     the converged field becomes the scene's load field, so the structure starts solved and settled;
  2. the charges (explode_at, src/lib.rs:2697-2716);
  3. `frames` frames of the legacy brick step (src/lib.rs:3639-3721), collapse and fracture included.
The answer: joints failed by mode (compression, tension, shear, footing), the fragment table per dominant material
(pieces, bricks, largest, pieces of 50+ bricks, crumbs of 1-2 bricks; tools/debug_exports.rs:41-56), bricks still
standing and pieces made.

Continuous batching: every query lives in a scene of one lattice world; all live scenes advance together in one
batched frame. Queries that arrive while others are on the GPU are admitted as soon as a slot is free and their
scene's resident bytes fit the HBM budget, take brick and body ranges of their own size from the world's pool, and
leave the batch the frame they finish. Each scene's result does not depend on the others in the batch (every order
and priority in a frame is scene local), which the tests check.
"""
import ctypes
import time
from collections import deque

import numpy as np

from . import scenes
from ._native import check, lib, ptr

C = 320  # lattice index of x = 0 and z = 0

_SIZE_TABLE = None


def _size_table():
    """(bricks, width, storeys) of every square frame building up to 40 storeys and 120 bricks wide"""
    global _SIZE_TABLE
    if _SIZE_TABLE is None:
        from ._native import lib
        rows = []
        for w in range(6, 121):
            one = lib().sf_scene_building(w, w, 1, 6, 0, 1, None, None, None)  # every storey is the same
            for storeys in range(1, 41):
                rows.append((one * storeys, w, storeys))
        _SIZE_TABLE = np.array(rows, np.int64)
    return _SIZE_TABLE


def building_near(n_target, seed):
    """a seeded frame/wall building of about n_target bricks (square footprint, storeys and width chosen for a
    sensible aspect) with 1-8 ground-floor columns removed, centred on x = z = 0"""
    t = _size_table()
    err = np.abs(t[:, 0] - n_target) / n_target
    aspect = np.abs(t[:, 2] * 7 - t[:, 1]) / np.maximum(t[:, 1], t[:, 2] * 7)
    k = int(np.argmin(err + 0.2 * aspect))
    _, w, storeys = t[k]
    rng = np.random.default_rng(seed)
    mat, xyz, infill = scenes.building(int(w), int(w), int(storeys), 6, int(rng.integers(1, 9)), seed)
    xyz = np.array(xyz, np.int32)
    xyz[:, 0] += C - int(w) // 2
    xyz[:, 2] += C - int(w) // 2
    return mat, xyz, infill


def make_query(qid, seed, n_min=2000, n_max=40000):
    """one what-if query: building size log-uniform in [n_min, n_max], 1-8 charges of radius U(1.5, 3) m at random
    ground-floor column bases, power 1.4"""
    rng = np.random.default_rng(seed)
    n = int(np.exp(rng.uniform(np.log(n_min), np.log(n_max))))
    bseed = int(rng.integers(1, 2 ** 31))
    mat, xyz, infill = building_near(n, bseed)
    g = xyz[(xyz[:, 1] == 0) & (mat == 1)]
    cols = sorted(set((int(a), int(b)) for a, b in g[:, [0, 2]]))
    k = int(rng.integers(1, 9))
    pick = [cols[i] for i in rng.choice(len(cols), size=min(k, len(cols)), replace=False)]
    charges = [((x - C) * 0.5 + 0.5, 0.6, (z - C) * 0.5 + 0.5, float(rng.uniform(1.5, 3.0)), 1.4) for x, z in pick]
    return dict(id=qid, seed=seed, bricks=len(mat), mat=mat, xyz=xyz, infill=infill, charges=charges,
                rng_seed=bseed & 0xFFFFFFFF or 1)


class _Answer(ctypes.Structure):
    """sf_answer (include/sf.h)"""
    _fields_ = [("id", ctypes.c_uint64), ("bricks", ctypes.c_int32), ("standing", ctypes.c_int32),
                ("pieces", ctypes.c_int32), ("load_sweeps", ctypes.c_int32), ("failed", ctypes.c_uint32 * 4),
                ("fragments", ctypes.c_float * 25), ("load_resid", ctypes.c_double), ("latency_ms", ctypes.c_double),
                ("queue_ms", ctypes.c_double)]


class WhatIfServer:
    """The library's continuous-batching server (sf_serve_*): `slots` scenes in flight at most, a brick pool of
    slots x slot_bricks shared by them, admission against `hbm_budget` bytes of resident scenes (default: the
    whole pool's)."""

    def __init__(self, slots=64, slot_bricks=40960, slot_bodies=8192, frames=300, hbm_budget=None, dt=1.0 / 60.0,
                 max_solving=3):
        self.frames = frames
        self.dt = dt
        self.slot_bricks = slot_bricks
        self._h = lib().sf_serve_create(int(slots), int(slots * slot_bricks), int(slot_bodies), int(frames),
                                        int(hbm_budget or 0), int(max_solving))
        if not self._h:
            check(-1)
        self.done = []
        self._by_id = {}

    def query_bytes(self, q):
        """resident bytes the query's scene would hold: what admission charges against the budget"""
        return check(lib().sf_serve_query_bytes(self._h, int(q["bricks"])))

    def status(self):
        out = np.zeros(8, np.int64)
        check(lib().sf_serve_status(self._h, ptr(out)))
        return dict(zip(["queued", "solving", "live", "answers", "bytes_in_use", "budget", "brick_steps", "frames"],
                        out.tolist()))

    def timing(self):
        """wall ms spent in admission, activation, frames and retirement since the server was made"""
        out = np.zeros(4, np.float64)
        check(lib().sf_serve_timing(self._h, ptr(out)))
        return dict(zip(["admit", "activate", "frames", "retire"], out.tolist()))

    @property
    def brick_steps(self):
        return self.status()["brick_steps"]

    @property
    def frames_run(self):
        return self.status()["frames"]

    def submit(self, q, t_arrival=None):
        """queue a query; one that could never be admitted (its scene's bytes beyond the budget) is refused here
        with ValueError"""
        xyz = np.ascontiguousarray(q["xyz"], np.int32)
        mat = np.ascontiguousarray(q["mat"], np.uint8)
        inf = np.ascontiguousarray(q["infill"], np.uint8)
        ch = np.ascontiguousarray(np.array(q["charges"], np.float32).reshape(-1, 5))
        rc = lib().sf_serve_submit(self._h, int(q["id"]), len(mat), ptr(xyz), ptr(mat), ptr(inf), len(ch), ptr(ch),
                                   int(q.get("rng_seed", 1)))
        if rc == -3:
            raise ValueError(lib().sf_last_error_message().decode())
        check(rc)
        self._by_id[int(q["id"])] = {k: v for k, v in q.items() if k not in ("mat", "xyz", "infill")}

    def _collect(self):
        buf = (_Answer * 64)()
        while True:
            n = check(lib().sf_serve_poll(self._h, ctypes.cast(buf, ctypes.c_void_p), 64))
            for a in buf[:n]:
                q = dict(self._by_id.pop(int(a.id), {"id": int(a.id)}))
                q["result"] = dict(failed=list(a.failed), fragments=np.array(a.fragments, np.float32).reshape(5, 5).tolist(),
                                   standing=int(a.standing), pieces=int(a.pieces))
                q["latency"] = a.latency_ms * 1e-3
                q["queue"] = a.queue_ms * 1e-3
                q["load_resid"] = a.load_resid
                q["load_sweeps"] = int(a.load_sweeps)
                self.done.append(q)
            if n < 64:
                return

    def busy(self):
        """queries waiting for a slot, solving their load path, or stepping"""
        st = self.status()
        return bool(st["queued"] or st["solving"] or st["live"])

    def tick(self):
        """admit what can be admitted, bring in queries whose solve finished, advance every live scene one frame,
        retire finished queries; returns the live scenes"""
        n = check(lib().sf_serve_tick(self._h, ctypes.c_float(self.dt)))
        self._collect()
        return n

    def run_stream(self, queries, arrivals):
        """serve queries arriving at the given times (s, from now) in real time; returns the finished queries"""
        t0 = time.perf_counter()
        pending = deque(sorted(zip(arrivals, range(len(queries)))))
        while pending or self.busy():
            now = time.perf_counter() - t0
            while pending and pending[0][0] <= now:
                _, i = pending.popleft()
                self.submit(queries[i])
            if not self.busy() and pending:
                time.sleep(max(0.0, min(pending[0][0] - now, 0.01)))
                continue
            self.tick()
        return self.done

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_serve_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass


def poisson_arrivals(rate, n, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))

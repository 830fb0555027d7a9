"""Served what-if stream: demolition queries answered by continuous batching on one GPU.

A query is a seeded frame/wall building of 2k-40k bricks plus 1-8 charges at its ground-floor column bases. It is
answered in three parts, all on the GPU:
  1. the load-path solve of the intact building (FP32 red-black sweeps, FP64 residual and refinement, sf_load);
     the converged field becomes the scene's load field, so the structure starts solved and settled;
  2. the charges (explode_at, src/lib.rs:2697-2716);
  3. `frames` frames of the legacy brick step (src/lib.rs:3639-3721), collapse and fracture included.
The answer: joints failed by mode (compression, tension, shear, footing), the fragment table per dominant material
(pieces, bricks, largest, pieces of 50+ bricks, crumbs of 1-2 bricks; tools/debug_exports.rs:41-56), bricks still
standing and pieces made.

Continuous batching: every query lives in a scene slot of one lattice world; all live slots advance together in one
batched frame. Queries that arrive while others are on the GPU are admitted as soon as a slot is free and their
resident bytes fit the HBM budget, and leave the batch the frame they finish, so scenes of uneven size and age share
every launch. Each scene's result does not depend on the others in the batch (every order and priority in a frame
is scene local), which the tests check.
"""
import time
from collections import deque

import numpy as np

from . import scenes
from .lattice import Lattice
from .load import LoadBatch

C = 320  # lattice index of x = 0 and z = 0

# device bytes a resident scene holds per brick and per body slot (lattice world arrays, contacts at 4 per brick)
BYTES_PER_BRICK = 176 + 4 * 196
BYTES_PER_BODY = 420

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


class WhatIfServer:
    """Continuous-batching server over `slots` resident scene slots of `slot_bricks` bricks each."""

    def __init__(self, slots=64, slot_bricks=40960, slot_bodies=8192, frames=300, hbm_budget=None, dt=1.0 / 60.0,
                 max_solving=3):
        self.frames = frames
        self.max_solving = max_solving
        self.dt = dt
        self.slot_bricks = slot_bricks
        self.lat = Lattice(slots * slot_bricks, slots * slot_bodies, scenes=slots)
        self.slot = [self.lat.add_scene(slot_bricks, slot_bodies, structs=2) for _ in range(slots)]
        self.free = deque(range(slots))
        self.budget = hbm_budget if hbm_budget is not None else slots * (slot_bricks * BYTES_PER_BRICK + slot_bodies * BYTES_PER_BODY)
        self.in_use = 0
        self.live = {}          # slot -> query state
        self.queue = deque()    # admitted to the server, waiting for a slot
        self.pending = []       # (slot and query list, load batch) whose load-path solve is running
        self.done = []
        self.brick_steps = 0
        self.frames_run = 0

    @staticmethod
    def query_bytes(q):
        return q["bricks"] * BYTES_PER_BRICK + 4096 * BYTES_PER_BODY

    def submit(self, q, t_arrival):
        """queue a query; one that can never be admitted (more bricks than a slot holds, or more resident bytes than
        the whole budget) is refused here with ValueError, before it reaches the queue"""
        if q["bricks"] > self.slot_bricks:
            raise ValueError(f"query {q.get('id')}: {q['bricks']} bricks, a slot holds {self.slot_bricks}")
        if self.query_bytes(q) > self.budget:
            raise ValueError(f"query {q.get('id')}: {self.query_bytes(q)} resident bytes, the budget is {self.budget}")
        q = dict(q)
        q["t_arrival"] = t_arrival
        self.queue.append(q)

    def _admit(self, now):
        """slots and HBM bytes permitting, queued queries take a slot and their load-path solves start on the load
        batch's own stream; the scenes already in flight keep stepping meanwhile. At most `max_solving` solve
        batches run at once (each on a stream of its own beside the frame's); later arrivals form the next batch"""
        if len(self.pending) >= self.max_solving:
            return
        batch = []
        while self.queue and self.free and self.in_use + self.query_bytes(self.queue[0]) <= self.budget:
            q = self.queue.popleft()
            s = self.free.popleft()
            self.in_use += self.query_bytes(q)
            batch.append((s, q))
        if not batch:
            return
        load = LoadBatch()
        for s, q in batch:
            load.add(q["mat"], q["xyz"], q["infill"])
        load.solve_begin(rel_tol=1e-6)
        self.pending.append((batch, load))

    def _activate(self, now, wait=False):
        """queries whose load-path solve has finished join the batch: scene, converged field, charges"""
        keep = []
        for batch, load in self.pending:
            if not (wait or load.ready()):
                keep.append((batch, load))
                continue
            load.solve_end()
            for k, (s, q) in enumerate(batch):
                self.lat.clear_scene(s, q["rng_seed"])
                q["structure"] = self.lat.add_structure(s, q["mat"], q["xyz"], q["infill"])
                self.lat.finalize(s)
                inf = load.info(k)
                q["load_resid"] = inf["resid"]
                q["load_sweeps"] = inf["sweeps"]
                self.lat.set_field(s, q["structure"], load.field(k).astype(np.float32))
                for ch in q["charges"]:
                    self.lat.explode(s, *ch)
                q["t_start"] = now
                q["frame"] = 0
                self.live[s] = q
            load.close()
        self.pending = keep

    def _retire(self, now):
        for s in [s for s, q in self.live.items() if q["frame"] >= self.frames]:
            q = self.live.pop(s)
            st = self.lat.stats(s)
            ms = self.lat.matstats(s).reshape(10, 5)
            q["result"] = dict(failed=self.lat.failures(s).tolist(), fragments=ms[1:6].tolist(),
                               standing=int(st[0]), pieces=int(st[2]))
            q["t_done"] = time.perf_counter()
            q["latency"] = q["t_done"] - q["t_arrival"]
            for k in ("mat", "xyz", "infill"):
                q.pop(k, None)
            self.done.append(q)
            self.in_use -= self.query_bytes(q)
            self.lat.clear_scene(s, 1)  # an empty slot costs the batched frame nothing
            self.free.append(s)

    def busy(self):
        """queries waiting for a slot, solving their load path, or stepping"""
        return bool(self.queue or self.pending or self.live)

    def tick(self):
        """admit what can be admitted, activate queries whose solve finished, advance every live scene one frame,
        retire finished queries"""
        now = time.perf_counter()
        self._admit(now)
        self._activate(now, wait=not self.live)
        if self.live:
            self.lat.step(self.dt)
            self.frames_run += 1
            for q in self.live.values():
                q["frame"] += 1
                self.brick_steps += q["bricks"]
        self._retire(time.perf_counter())
        return len(self.live)

    def run_stream(self, queries, arrivals):
        """serve queries arriving at the given times (s, from now) in real time; returns the finished queries"""
        t0 = time.perf_counter()
        pending = deque(sorted(zip(arrivals, range(len(queries)))))
        while pending or self.busy():
            now = time.perf_counter() - t0
            while pending and pending[0][0] <= now:
                ta, i = pending.popleft()
                self.submit(queries[i], t0 + ta)
            if not self.busy() and pending:
                time.sleep(max(0.0, min(pending[0][0] - now, 0.01)))
                continue
            self.tick()
        return self.done

    def close(self):
        for _, load in self.pending:
            load.close()
        self.pending = []
        self.lat.close()


def poisson_arrivals(rate, n, seed=0):
    rng = np.random.default_rng(seed)
    return np.cumsum(rng.exponential(1.0 / rate, size=n))

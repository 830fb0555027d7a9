"""Load path: per-structure scalar-potential solve and joint verdicts on the GPU (src/lib.rs structures / SOR).

    L = LoadBatch()
    sid = L.add(*scenes.tower())
    L.solve(rel_tol=1e-6)
    phi = L.field(sid); ratio, mode, fail = L.verdicts(sid)
"""
import numpy as np

from ._native import lib, check, ptr

MODES = {0: "none", 1: "compression", 2: "tension", 3: "shear", 4: "ground"}


class LoadBatch:
    def __init__(self):
        self._h = lib().sf_load_create()
        if not self._h:
            check(-1)
        self.sizes = []

    def close(self):
        if getattr(self, "_h", None):
            lib().sf_load_destroy(self._h)
            self._h = None

    def __del__(self):
        try:
            self.close()
        except Exception:
            pass

    def add(self, mat, xyz, infill=None, ext=None):
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32).reshape(-1, 3)
        infill = np.zeros(len(mat), np.uint8) if infill is None else np.ascontiguousarray(infill, np.uint8)
        ext = None if ext is None else np.ascontiguousarray(ext, np.float32)
        sid = check(lib().sf_load_add(self._h, len(mat), ptr(mat), ptr(xyz), ptr(infill), ptr(ext)))
        self.sizes.append(len(mat))
        return sid

    def set_packed(self, on=True):
        """FP32 sweeps from packed 16-byte rows (default) or from the full neighbour and conductance arrays; the
        field is the same bit for bit"""
        check(lib().sf_load_set_packed(self._h, 1 if on else 0))

    def solve(self, rel_tol=1e-6, omega=0.0, max_sweeps=200000):
        """FP32 sweeps with FP64 residual and refinement to max|b - A phi| / max|b| <= rel_tol; returns the number
        of structures that converged."""
        return check(lib().sf_load_solve(self._h, float(omega), float(rel_tol), int(max_sweeps)))

    def solve_begin(self, rel_tol=1e-6, omega=0.0, max_sweeps=200000):
        """start the same solve on the batch's own stream and return at once"""
        check(lib().sf_load_solve_begin(self._h, float(omega), float(rel_tol), int(max_sweeps)))

    def ready(self):
        """True once a started solve has finished"""
        return check(lib().sf_load_solve_ready(self._h)) == 1

    def solve_end(self):
        """wait for a started solve; returns the number of structures that converged"""
        return check(lib().sf_load_solve_end(self._h))

    def info(self, sid):
        sw = np.zeros(1, np.int32)
        rs = np.zeros(1, np.float64)
        g = np.zeros(1, np.int32)
        check(lib().sf_load_info(self._h, int(sid), ptr(sw), ptr(rs), ptr(g)))
        return {"sweeps": int(sw[0]), "resid": float(rs[0]), "grounded": int(g[0])}

    def field(self, sid):
        out = np.zeros(self.sizes[sid], np.float64)
        check(lib().sf_load_field(self._h, int(sid), ptr(out)))
        return out

    def verdicts(self, sid):
        n = self.sizes[sid]
        r = np.zeros((n, 4), np.float64)
        m = np.zeros((n, 4), np.uint8)
        f = np.zeros((n, 4), np.uint8)
        check(lib().sf_load_verdicts(self._h, int(sid), ptr(r), ptr(m), ptr(f)))
        return r, m, f

    def topology(self, sid):
        n = self.sizes[sid]
        nb6 = np.zeros((n, 6), np.int32)
        bond = np.zeros((n, 4), np.float32)
        check(lib().sf_load_topology(self._h, int(sid), ptr(nb6), ptr(bond)))
        return nb6, bond

    def timing(self):
        ms = np.zeros(1, np.float32)
        s32 = np.zeros(1, np.int64)
        s64 = np.zeros(1, np.int64)
        nb = np.zeros(1, np.int64)
        check(lib().sf_load_last_timing(self._h, ptr(ms), ptr(s32), ptr(s64), ptr(nb)))
        return {"ms": float(ms[0]), "brick_sweeps32": int(s32[0]), "brick_sweeps64": int(s64[0]), "bricks": int(nb[0])}

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
BODY_F, SHAPE_F, MAN_F = 20, 11, 40

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
            "ob_read_bodies": (None, [vp, vp]), "ob_stats": (None, [vp, vp]), "ob_candidates": (ci, [vp, vp, ci]),
            "ob_manifolds": (ci, [vp, vp, vp, ci]), "ob_colours": (ci, [vp, vp, ci]), "ob_islands": (None, [vp, vp]),
            "ob_bruteforce_pairs": (ci, [vp, cf, vp, ci]),
            "ol_sor_sweep": (cf, [vp, vp, cf]), "ol_solve64": (ci, [vp, vp, cd, ci]), "ol_residual64": (cd, [vp, vp]),
            "ol_verdicts64": (None, [vp, vp, vp, vp, vp]), "ol_verdicts32": (None, [vp, vp, vp, vp]),
            "ol_grounded": (ci, [vp, vp]), "ol_finalize": (None, [ci, vp, vp, vp, vp, vp]),
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

    def __init__(self, bodies, shape_body, shapes, prm=None, order=ORDER_COLOURED):
        L = lib()
        self._h = L.ob_create(order)
        if prm is not None:
            prm = np.ascontiguousarray(prm, np.float32)
            L.ob_set_params(self._h, _p(prm), len(prm))
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

    def bodies(self):
        out = np.zeros((self.nb, BODY_F), np.float32)
        lib().ob_read_bodies(self._h, _p(out))
        return out

    def stats(self):
        out = np.zeros(8, np.int32)
        lib().ob_stats(self._h, _p(out))
        return dict(zip(["awake", "pairs", "manifolds", "points", "colours", "overflow", "ground", "wakes"], out.tolist()))

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

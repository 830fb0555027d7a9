"""The allocation-free load solve (sf_load_solve_static) against the handle path (sf_load_add + sf_load_solve +
sf_load_field + sf_load_verdicts) through the C ABI: the same FP64 field, verdicts, sweep count and residual, bit
for bit, on two calls in a row (the second smaller, so the scratch keeps the first call's larger contents), and a
structure over the scratch's limit refused with SF_E_CAP before anything runs."""
import numpy as np
import pytest

from shatterforge_amd import scenes
from shatterforge_amd._native import lib, ptr

pytestmark = pytest.mark.gpu


def _protos():
    return lib()  # the binding declares both entry points (shatterforge_amd/_native.py)


def _static(L, mat, xyz, inf, ext=None):
    n = len(mat)
    phi = np.zeros(n, np.float64)
    ratio = np.zeros((n, 4), np.float64)
    mode = np.zeros((n, 4), np.uint8)
    fail = np.zeros((n, 4), np.uint8)
    sw = np.zeros(1, np.int32)
    res = np.zeros(1, np.float64)
    rc = L.sf_load_solve_static(n, ptr(mat), ptr(xyz), ptr(inf), ptr(ext) if ext is not None else None, 0.0, 1e-6,
                                200000, None, ptr(phi), ptr(ratio), ptr(mode), ptr(fail), ptr(sw), ptr(res))
    return rc, phi, (ratio, mode, fail), int(sw[0]), float(res[0])


def _handle(mat, xyz, inf, ext=None):
    import shatterforge_amd as sf
    B = sf.LoadBatch()
    sid = B.add(mat, xyz, inf, ext=ext)
    ok = B.solve(rel_tol=1e-6)
    out = (ok, B.field(sid), B.verdicts(sid), B.info(sid)["sweeps"], B.info(sid)["resid"])
    B.close()
    return out


def test_static_solve_equals_the_handle_path_twice_then_smaller():
    L = _protos()
    rng = np.random.default_rng(21)
    # 50k bricks: the correction in global memory; 30k: in LDS from packed rows; each call smaller than the last
    big = scenes.building_for(50000, seed=5)
    mid = scenes.building_for(30000, seed=7)
    small = scenes.building_for(3000, seed=6)
    ext = rng.uniform(-20.0, 300.0, len(small[0])).astype(np.float32)
    for tag, (mat, xyz, inf), e in (("50k-brick building", big, None), ("30k-brick building", mid, None),
                                     ("3k-brick building with loads", small, ext)):
        print(f"case: {tag}", flush=True)
        mat = np.ascontiguousarray(mat, np.uint8)
        xyz = np.ascontiguousarray(xyz, np.int32)
        inf = np.ascontiguousarray(inf, np.uint8)
        rc, phi, v, sw, res = _static(L, mat, xyz, inf, e)
        ok, phi_h, v_h, sw_h, res_h = _handle(mat, xyz, inf, e)
        assert rc == 1 and ok == 1, (rc, ok)
        assert sw == sw_h and res == res_h, (sw, sw_h, res, res_h)
        assert np.array_equal(phi.view(np.uint64), phi_h.view(np.uint64))
        for a, b in zip(v, v_h):
            assert np.array_equal(a, b)


def test_static_solve_over_its_limit_is_refused():
    from shatterforge_amd._native import SfError, check
    L = _protos()
    cap = int(L.sf_load_static_capacity())
    n = cap + 1
    mat = np.full(n, 1, np.uint8)
    xyz = np.stack([np.arange(n) % 256, np.arange(n) // 65536, (np.arange(n) // 256) % 256], 1).astype(np.int32)
    inf = np.zeros(n, np.uint8)
    phi = np.zeros(n, np.float64)
    rc = L.sf_load_solve_static(n, ptr(mat), ptr(xyz), ptr(inf), None, 0.0, 1e-6, 1000, None, ptr(phi),
                                None, None, None, None, None)
    assert rc == -3
    with pytest.raises(SfError, match="at most"):
        check(rc)

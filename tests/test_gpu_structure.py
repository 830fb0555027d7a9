"""The structural analysis on the GPU (libsf_hip sf_net_*: one workgroup per case runs the whole multigrid-
preconditioned conjugate-gradient solve) against the oracle (oracle/structure.c, its device summation order): the
cells, the displacements (f64), the iteration counts and every joint's and inner face's ratio equal bit for bit; the
verdicts equal the reference order's; a batch in one launch equals each case alone."""
import numpy as np
import pytest

from shatterforge_amd import scenes, structure
from oracle import pyoracle as po
from tests.structure_cases import both, gravity, cantilever, span, slab_on_column, mixed_frame, C, M, T

pytestmark = pytest.mark.gpu


def _same(a, o, tag):
    assert a.nodes == o["nodes"] and a.links == o["links"], tag
    assert a.levels == o["levels"], (tag, a.levels, o["levels"])
    assert np.array_equal(a.cells, o["node"]), tag
    assert a.iterations == o["iterations"], (tag, a.iterations, o["iterations"])
    assert np.array_equal(a.x, o["x"]), f"{tag}: displacements differ by {np.abs(a.x - o['x']).max()}"
    assert np.array_equal(a.joint_ids, o["joint_ids"]), tag
    assert np.array_equal(a.joint_ratios, o["joint_ratios"]), tag
    assert np.array_equal(a.inner_faces, o["inner"]), tag
    assert np.array_equal(a.inner_ratios, o["inner_ratios"]), tag
    assert a.max_ratio == o["max_ratio"], tag
    assert a.worst_in[:5] == o["worst_in"], (tag, a.worst_in, o["worst_in"])


def _column(m, p):
    return [([0, 0, 0], [4, 24, 4], m, 0, 1), ([0, 24, 0], [4, 2, 4], m, 0, 1)], [(1, p)]


CASES = {
    "cantilever 3 m": (cantilever(24, 2), ()), "cantilever 7 m": (cantilever(56, 2), ()),
    "cantilever 5 m deep": (cantilever(40, 4), ()), "timber cantilever": (cantilever(40, 2, T), ()),
    "one-brick cantilever 7.5 m": (cantilever(60, 2), ()),
    "span 6 m": (span(48, 2), ()), "span 12 m": (span(96, 2), ()), "timber span": (span(96, 4, m=T), ()),
    "span 22.5 m": ([([0, 0, 0], [8, 8, 8], C, 0, 1), ([188, 0, 0], [8, 8, 8], C, 0, 1), ([0, 8, 0], [196, 2, 8], C, 0, 1)], ()),
    "concrete column": _column(C, 5.0e6), "masonry column": _column(M, 2.0e6),
    "slab on a corner column": (slab_on_column(0), ()), "slab on a centred column": (slab_on_column(20), ()),
    "mixed frame": (mixed_frame(), ()),
}


@pytest.mark.parametrize("name", list(CASES))
def test_validation_structures_bit_exact(oracle, name):
    bricks, extra = CASES[name]
    print(f"case: {name}", flush=True)
    b, o, ys = both(bricks)
    set_, load, sink = gravity(o, ys, extra)
    a = structure.analyse(b, set_, load, sink)
    _same(a, po.analyse_oracle(o, set_, load, sink, order=po.ORDER_DEVICE), name)


ROUTES = {"one workgroup": (-1, 16), "a kernel per step": (0, 16)}
# The kernel-per-step route made a GPU memory fault (HSA_STATUS_ERROR_MEMORY_APERTURE_VIOLATION) twice, in the
# warm-start test on a 4-storey tower; its cause is not found, so the route is switched off in the library and its
# tests do not run until it is.
PER_STEP_OFF = pytest.mark.skip(reason="kernel-per-step route switched off: it faulted the GPU (cause not found yet)")
ROUTE_PARAMS = [pytest.param("one workgroup"), pytest.param("a kernel per step", marks=PER_STEP_OFF)]


def _solve(b, set_, load, sink, route, dir3=(0.0, -1.0, 0.0), point=(), warm=None):
    net = structure.Net()
    net.set_route(*ROUTES[route])
    net.add(b, set_, load, sink, dir3, point, warm)
    net.solve()
    return net.result(0)


@pytest.mark.parametrize("route", ROUTE_PARAMS)
@pytest.mark.parametrize("storeys,remove", [(12, ()), (6, ((0, 0), (1, 1), (2, 1), (1, 0)))])
def test_frame_tower_bit_exact_and_verdicts_of_the_reference_order(oracle, storeys, remove, route):
    print(f"case: frame tower, {storeys} storeys, ground columns removed {remove}, {route}", flush=True)
    b, o, ys = both(scenes.frame_tower(storeys, 3, remove=set(remove)))
    set_, load, sink = gravity(o, ys)
    a = _solve(b, set_, load, sink, route)
    _same(a, po.analyse_oracle(o, set_, load, sink, order=po.ORDER_DEVICE), "tower")
    r = po.analyse_oracle(o, set_, load, sink, order=po.ORDER_REFERENCE)
    assert np.array_equal(a.joint_ratios > 1.0, r["joint_ratios"] > 1.0)
    assert np.array_equal(a.inner_ratios > 1.0, r["inner_ratios"] > 1.0)
    assert np.abs(a.joint_ratios - r["joint_ratios"]).max() <= 1e-5 * max(1.0, float(r["max_ratio"]))


def test_batch_in_one_launch_equals_each_alone(oracle):
    print("case: 8 seeded frame buildings in one launch, then each alone", flush=True)
    towers = scenes.frame_towers(8, seed=4, storeys=(3, 8))
    stores = [both(t) for t, _ in towers]
    net = structure.Net()
    for b, o, ys in stores:
        net.add(b, *gravity(o, ys))
    net.solve()
    for k, (b, o, ys) in enumerate(stores):
        set_, load, sink = gravity(o, ys)
        got = net.result(k)
        _same(got, po.analyse_oracle(o, set_, load, sink, order=po.ORDER_DEVICE), f"building {k} in the batch")
        alone = structure.analyse(b, set_, load, sink)
        assert np.array_equal(alone.x, got.x) and np.array_equal(alone.joint_ratios, got.joint_ratios)


@pytest.mark.parametrize("route", ROUTE_PARAMS)
def test_warm_start_point_loads_and_a_sideways_load(oracle, route):
    print(f"case: warm start twice, a point load, a sideways load, {route}", flush=True)
    b, o, ys = both(scenes.frame_tower(4, 3))
    set_, load, sink = gravity(o, ys)
    wg, wo = structure.Warm(), po.WarmOracle()
    for rnd in range(2):
        a = _solve(b, set_, load, sink, route, warm=wg)
        _same(a, po.analyse_oracle(o, set_, load, sink, order=po.ORDER_DEVICE, warm=wo), f"warm {rnd}")
    assert a.iterations == 0 and len(wg) == len(wo)
    # a warm start from a different load: iterations from a non-zero start
    a = _solve(b, set_, load * np.float32(1.5), sink, route, warm=wg)
    _same(a, po.analyse_oracle(o, set_, load * np.float32(1.5), sink, order=po.ORDER_DEVICE, warm=wo), "warm, heavier")
    assert a.iterations > 0
    roof = int(np.argmax(ys))
    pt = [(roof, 30.0, 50.0, 4.0e5)]
    a = _solve(b, set_, load, sink, route, point=pt)
    _same(a, po.analyse_oracle(o, set_, load, sink, point=pt, order=po.ORDER_DEVICE), "point")
    d = (0.6, -0.8, 0.0)
    a = _solve(b, set_, load * np.float32(3.0), sink, route, dir3=d)
    _same(a, po.analyse_oracle(o, set_, load * np.float32(3.0), sink, dir3=d, order=po.ORDER_DEVICE), "sideways")

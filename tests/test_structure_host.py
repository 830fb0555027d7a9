"""The structural analysis's host side in the product (libsf_hip's brick store; no GPU needed) against the oracle:
the joints of a store are the oracle's, in the oracle's numbering, and the masses are the oracle's bit for bit."""
import numpy as np
import pytest

from shatterforge_amd import scenes
from tests.structure_cases import both, mixed_frame, cantilever


@pytest.mark.parametrize("name", ["tower", "mixed", "cantilever", "random"])
def test_product_joints_equal_the_oracle_in_order(name):
    if name == "tower":
        bricks = scenes.frame_tower(5, 3, remove={(0, 0), (2, 1)})
    elif name == "mixed":
        bricks = mixed_frame()
    elif name == "cantilever":
        bricks = cantilever(40, 2)
    else:
        rng = np.random.default_rng(3)
        occ = np.zeros((24, 24, 24), bool)
        bricks = []
        for _ in range(400):
            mn = rng.integers(-8, 16, 3)
            size = rng.integers(1, 6, 3)
            lo = mn + 8
            hi = np.minimum(lo + size, 24)
            if occ[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]].any() or (hi - lo < size).any():
                continue
            occ[lo[0]:hi[0], lo[1]:hi[1], lo[2]:hi[2]] = True
            bricks.append((mn.tolist(), size.tolist(), int(rng.integers(0, 5)), 0, 1 | (4 if rng.random() < 0.1 else 0)))
    b, o, _ = both(bricks)
    gj, oj = b.joints(), o.joints()
    assert len(gj) > 0 and np.array_equal(gj, oj)
    for i in range(len(bricks)):
        assert b.mass(i) == o.mass(i)


def test_removed_brick_joints_and_rebuild():
    b, o, _ = both(scenes.frame_tower(3, 2))
    for i in (3, 10, 11):
        b.remove(i)
        o.remove(i)
    assert np.array_equal(b.joints(), o.joints())
    b.build_joints()
    o.build_joints()
    assert np.array_equal(b.joints(), o.joints())


def test_bad_bricks_are_refused():
    from shatterforge_amd import structure
    from shatterforge_amd._native import SfError
    b = structure.Bricks()
    with pytest.raises(SfError):
        b.add([0, 0, 0], [0, 4, 4], 0)
    with pytest.raises(SfError):
        b.add([0, 0, 0], [4, 256, 4], 0)
    with pytest.raises(SfError):
        b.add([0, 0, 0], [4, 4, 4], 7)

"""Collapse statistics of the device's solve order against the reference's sequential order, both run by the oracle
(the GPU equals the device order bit for bit, tests/test_gpu_lattice.py). Collapses are chaotic, so the check is the
reference's own: per-material fragment counts and masses after the fixture's charge set (tests/frac.mjs:14-16 with
the charge sets of tests/office_collapse.json and tests/masonry_collapse.json, on this repository's buildings), two
random seeds per order, averaged.

Stated tolerance: the structure's lost fraction within 0.02; per dominant material with at least 20 pieces, the
piece count and the bricks left in pieces within 40 % of the reference order's. Measured (900 frames, 4 seeds per
order): office lost 0.988-0.994 in both, concrete pieces -9 %, concrete bricks in pieces +35 %, masonry pieces
-11 %; masonry shaft lost 0.076 / 0.077, masonry pieces -34 %, bricks in pieces -4 %. Before the static chains
(each body's contacts against static bricks and the ground solved in generation order) the office lost 0.86 and
kept a column standing, so the lost fraction is what a plausible reordering breaks first.
"""
import numpy as np
import pytest

import lattice_cases as lc

FRAMES = 900


def _run(oracle, order, st, charges, seed):
    mat, xyz, inf = st
    O = oracle.LatticeOracle(order=order, brick_cap=len(mat) + 64, body_cap=16384, seed=seed)
    O.add_structure(mat, xyz, inf)
    O.finalize()
    O.warm_all()
    n = 0
    while O.stats()[14] > 0 or n < 30:
        O.step()
        n += 1
    for c in charges:
        O.explode(*c)
    for _ in range(FRAMES):
        O.step()
    s = O.structs()[0]
    return (s[4] - s[5]) / s[4], O.matstats().reshape(10, 5)[1:6].astype(np.float64)


@pytest.mark.parametrize("case", ["office", "shaft"])
def test_fragment_histograms_follow_the_reference_order(oracle, case):
    st, charges = lc.office_all_columns() if case == "office" else lc.masonry_shaft()
    res = {}
    for order in (oracle.ORDER_SEQ, oracle.ORDER_DEV):
        runs = [_run(oracle, order, st, charges, seed) for seed in (1, 2)]
        res[order] = (np.mean([r[0] for r in runs]), np.mean([r[1] for r in runs], axis=0))
    (lost_s, ms_s), (lost_d, ms_d) = res[oracle.ORDER_SEQ], res[oracle.ORDER_DEV]
    print(f"{case}: lost {lost_s:.3f} (reference order) vs {lost_d:.3f} (device order)")
    assert abs(lost_s - lost_d) <= 0.02
    names = ["concrete", "masonry", "steel", "glass", "timber"]
    for m in range(5):
        if ms_s[m, 0] < 20:
            continue
        dp = ms_d[m, 0] / ms_s[m, 0] - 1.0
        db = ms_d[m, 1] / ms_s[m, 1] - 1.0
        print(f"  {names[m]}: pieces {ms_s[m, 0]:.0f} vs {ms_d[m, 0]:.0f} ({dp:+.0%}), bricks {ms_s[m, 1]:.0f} vs "
              f"{ms_d[m, 1]:.0f} ({db:+.0%})")
        assert abs(dp) <= 0.4 and abs(db) <= 0.4, (names[m], ms_s[m], ms_d[m])

"""The gap rule of the device's SAS-circle exposure test.

The host decides whether closed arc intervals [s, e] leave part of [0, 2 pi) open by a
sweep in (s, e) order (a start beyond the running end is a gap, and so is a final end
below 2 pi). The device kernel (`k_circle_exposed`, `arcs_leave_gap`) instead tests the
points v in {0} and the interval ends below 2 pi: a gap exists iff some such v lies in
no interval as s <= v < e. These tests check that the two rules decide alike on random
interval sets, including touching, nested, repeated and wrapped intervals, and on the
edge cases (no intervals, one full circle, an interval ending exactly at 2 pi).
"""
import math
import random

TWO_PI = 2.0 * math.pi


def sweep_gap(iv):
    """The host's sorted sweep."""
    if not iv:
        return True
    cur = 0.0
    for s, e in sorted(iv):
        if s > cur:
            return True
        cur = max(cur, e)
    return cur < TWO_PI


def point_gap(iv):
    """The device's rule: some v in {0} and the ends below 2 pi held by no interval."""
    if not iv:
        return True
    for v in [0.0] + [e for _, e in iv]:
        if not v < TWO_PI:
            continue
        if not any(s <= v < e for s, e in iv):
            return True
    return False


def arc(start, length):
    """One covered arc as the host and the device store it: split at 2 pi."""
    s0 = start % TWO_PI
    e0 = s0 + length
    return [(s0, TWO_PI), (0.0, e0 - TWO_PI)] if e0 > TWO_PI else [(s0, e0)]


def test_edge_cases():
    assert point_gap([]) and sweep_gap([])
    full = arc(1.0, TWO_PI)
    assert not point_gap(full) and not sweep_gap(full)
    assert not point_gap([(0.0, TWO_PI)]) and not sweep_gap([(0.0, TWO_PI)])
    # touching closed intervals cover; a start just after the running end opens a gap
    touch = [(0.0, 1.0), (1.0, 3.0), (3.0, TWO_PI)]
    assert not point_gap(touch) and not sweep_gap(touch)
    split = [(0.0, 1.0), (math.nextafter(1.0, 2.0), TWO_PI)]
    assert point_gap(split) and sweep_gap(split)
    # nothing from 0: the point 0 is open
    late = [(0.5, TWO_PI)]
    assert point_gap(late) and sweep_gap(late)


def test_random_interval_sets_decide_alike():
    rng = random.Random(20261203)
    grid = [k * TWO_PI / 64 for k in range(65)]
    n_gap = n_cover = 0
    for trial in range(20000):
        iv = []
        for _ in range(rng.randint(0, 14)):
            if rng.random() < 0.5:  # ends on a coarse grid: many exact touches and repeats
                s = rng.choice(grid[:-1])
                length = rng.choice(grid[1:])
            else:
                s = rng.uniform(0.0, TWO_PI)
                length = rng.uniform(0.0, TWO_PI)
            iv += arc(s, length)
        g = sweep_gap(iv)
        assert point_gap(iv) == g, iv
        n_gap += g
        n_cover += not g
    # both outcomes are well represented
    assert n_gap > 1000 and n_cover > 1000

"""Pins the oracle's energy reductions (src/tetwild/LocalOperations.cpp:605-693)
and configuration evaluation (src/tetwild/EdgeRemover.cpp) with constructed
inputs whose answers are fixed by hand: values placed inside and exactly on
the bucket limits of getFilterEnergy, histograms that take each branch of its
decision, sums against exact rational arithmetic."""
from fractions import Fraction

import numpy as np
import pytest

from oracle import oracle as O
from tetwild_amd import workloads as W

MAX = 1e50


def test_buckets_inside_and_on_limits():
    # thres = 10: bucket j holds (9 + 10^j, 9 + 10^(j+1)], bucket 10 holds > 9 + 1e10
    e, expect = [], np.zeros(11, np.int64)
    for j in range(10):
        e += [9 + 2 * 10.0 ** j, 9 + 10.0 ** (j + 1)]          # inside, upper limit
        expect[j] += 2
        e.append(np.nextafter(9 + 10.0 ** j, np.inf))          # just above the lower limit
        expect[j] += 1
    e += [np.nextafter(9 + 1e10, np.inf), 1e20, MAX, np.inf]
    expect[10] += 4
    e += [10.0, 3.0, 9.5, np.nan]                               # in no bucket
    s, m, m2, cnt, b = O.energy_stats(np.array(e), thres=10.0)
    assert np.array_equal(b, expect)
    assert cnt == len(e) and m == np.inf and m2 == np.inf
    assert np.isnan(s)


def test_removed_locked_max_second_sum():
    e = np.array([3.0, 4.5, MAX, 7.25, 100.0, 5.0])
    removed = np.array([0, 0, 0, 0, 1, 0], np.uint8)
    tets = np.array([[0, 1, 2, 3], [4, 5, 6, 7], [8, 9, 10, 11], [0, 5, 9, 12],
                     [1, 2, 3, 4], [13, 14, 15, 16]], np.int32)
    locked = np.zeros(17, np.uint8)
    locked[12] = 1                                              # locks tet 3 only
    s, m, m2, cnt, b = O.energy_stats(e, removed, tets, locked, thres=10.0)
    assert cnt == 4 and m == MAX and m2 == 5.0
    assert s == MAX                                             # 3 + 4.5 + 1e50 + 5 rounds to 1e50
    assert b[10] == 1 and b.sum() == 1                          # 100 is removed, 7.25 <= 10
    s, m, m2, cnt, b = O.energy_stats(np.array([3.0, 4.5, 7.25]))
    assert (s, m, m2, cnt) == (14.75, 7.25, 7.25, 3)


def test_sum_against_exact():
    st = W.Stream(3, 0)
    e = 3.0 + st.uniform(20000) * 1e3
    s = O.energy_stats(e)[0]
    exact = sum(Fraction(float(v)) for v in e)
    # recursive summation of positive terms: |s - exact| <= (n - 1) u sum
    assert abs(Fraction(s) - exact) <= Fraction(len(e) - 1, 2**53) * exact


@pytest.mark.parametrize("buckets,value,clean", [
    ([10, 0, 0, 0, 0, 0, 0, 0, 0, 0, 0], 8.0, False),      # t1[0] >= t2[0]
    ([10, 0, 0, 0, 0, 0, 0, 1, 0, 0, 0], 8.0, True),       # ... and t2[5] > 0
    ([0, 0, 0, 0, 0, 0, 0, 0, 0, 1, 5], 1e11, None),       # t1[8] <= t2[8]
    ([0, 0, 10, 0, 0, 0, 0, 0, 0, 0, 0], 509.0, None),     # crossing at i = 1: 9 + 5 * 10^2
    ([1, 0, 0, 0, 5, 0, 0, 0, 0, 0, 0], 50009.0, None),    # crossing at i = 3: 9 + 5 * 10^4
    ([0, 5, 5, 0, 0, 0, 0, 0, 0, 0, 0], 8.0, None),        # a tie, no crossing: falls through
])
def test_filter_energy_branches(buckets, value, clean):
    v, c = O.filter_energy(np.array(buckets), thres=10.0)
    assert v == value and c == clean


def test_tet_groups_composition():
    x, y, z, tets, _ = W.kuhn_lattice(3, seed=2)
    inv = tets[:, [0, 1, 3, 2]]
    groups = [tets[0:2], tets[2:6], np.concatenate([tets[6:8], inv[8:9]]), tets[9:9], inv[10:12]]
    off = np.concatenate([[0], np.cumsum([len(g) for g in groups])]).astype(np.int32)
    allt = np.concatenate(groups).astype(np.int32)
    fl, mq, q = O.tet_groups(allt, off, x, y, z)
    assert list(fl) == [0, 0, 1, 0, 1]
    assert np.isnan(mq[2]) and np.isnan(mq[4]) and mq[3] == 0.0
    for g in (0, 1):
        qq = O.tet_qualities(allt[off[g]:off[g + 1]], x, y, z)
        assert np.array_equal(q[off[g]:off[g + 1]], qq) and mq[g] == qq.max()

"""The perturbation kernels' BLA level search against the oracle's Bla::find.

The kernels (fractodactyl_amd/csrc/bla_find.hpp) search the levels top-down from the
highest candidate in groups of four radii, where Bla::find (crates/fd-kernel/src/bla.rs:
234-249) walks up from level 1 and stops at the first failure.  The two agree because
validity only shrinks going up: a level-i block's radius is at most that of its first
half, the level i-1 block at the same start (Block::merge, bla.rs:91-111).  These tests
check that invariant on real tables and replay the kernels' search, step for step, on the
oracle's table against the oracle's own find at every block start, for several |delta|^2
and step budgets.  The device code itself is checked end to end, bit for bit, by the GPU
parity tests.
"""
import numpy as np
import pytest

import oracle

GROUP = 4  # kFindGroup


def table(centre, width, max_iter, eps, dc_scale):
    re_, im_ = centre
    zr, zi, _ = oracle.reference(re_, im_, width, 0.0, 64, 36, max_iter, 1e10)
    dc = float(width) * dc_scale
    levels = oracle.bla_build(zr, zi, eps, dc)
    # r^2 per level in the kernels' indexing: level i holds blocks of 2 << i steps
    r2 = [lv[:, 4] * lv[:, 4] for lv in levels]
    return zr, zi, dc, r2


def search(r2, m, d2, budget):
    """The kernels' lookup (find_cap + level 0 + find_level_down), in Python."""
    k = m - 1
    if k < 0 or k & 1:
        return None
    nlevels = len(r2)
    top = nlevels if k == 0 else (k & -k).bit_length() - 1
    top = min(top, nlevels)
    by_budget = budget.bit_length() - 2 if budget >= 2 else -1  # 2 << i <= budget
    cap = min(top - 1, by_budget)

    def radius(i):
        idx = k >> (i + 1)
        return r2[i][idx] if idx < len(r2[i]) else 0.0  # zero padding past the end

    if cap < 0 or not d2 < radius(0):
        return None
    hi = cap
    while hi >= 1:
        got = -1
        for q in range(GROUP - 1, -1, -1):
            lv = hi - q
            if lv >= 1 and d2 < radius(lv):
                got = lv
        if got >= 1:
            return 2 << got
        hi -= GROUP
    return 2


CASES = [
    # (centre, width, max_iter, eps, dc / width): the v0 deep-zoom centre
    (("-0.743643887037158704752191506114774", "0.131825904205311970493132056385139"), "1e-28",
     20000, 2.0 ** -50, 0.9),
    # the period-16116 ladder nucleus of the KERN-05 rung-7676 frame, shallow
    (("-0.7432918908524302029316243259725107571", "0.13124055230879760477084590658147814307"),
     "1e-30", 30000, 2.0 ** -40, 0.7),
]


@pytest.mark.parametrize("centre,width,max_iter,eps,dcs", CASES)
def test_radii_shrink_going_up(centre, width, max_iter, eps, dcs):
    _, _, _, r2 = table(centre, width, max_iter, eps, dcs)
    assert len(r2) >= 4
    for i in range(1, len(r2)):
        first_half = r2[i - 1][0:2 * len(r2[i]):2]
        assert np.all(r2[i] <= first_half)


@pytest.mark.parametrize("centre,width,max_iter,eps,dcs", CASES)
def test_top_down_search_equals_find(centre, width, max_iter, eps, dcs):
    zr, zi, dc, r2 = table(centre, width, max_iter, eps, dcs)
    ms = np.arange(1, len(zr) - 1, dtype=np.uint64)
    positive = np.concatenate([lv[lv > 0] for lv in r2])
    # |delta|^2 from below every radius to above most, including radii themselves (ties)
    d2s = [0.0] + list(np.quantile(positive, [0.01, 0.2, 0.5, 0.8, 0.99])) + \
        [float(positive.max()) * 2, float(r2[0][7])]
    checked = 0
    for d2 in d2s:
        for budget in (2, 3, 17, 1000, 1 << 40):
            want = oracle.bla_find(zr, zi, eps, dc, ms, d2, budget)
            got = [search(r2, int(m), d2, budget) for m in ms]
            assert got == want, (d2, budget)
            checked += sum(g is not None for g in got)
    assert checked > 1000


def kernel_block_index(tb, stride, k, i):
    """bla_find.hpp's bla_block_index: (tb + stride) - ((stride - k/2) >> i), 32-bit."""
    return ((tb + stride) - ((stride - (k >> 1)) >> i)) & 0xFFFFFFFF


def test_block_index_is_the_level_major_position():
    # The kernels address block k >> (i + 1) of level i, which sits at
    # tb + stride - (stride >> i) + (k >> (i + 1)) in the level-major table, with one shift:
    # exact wherever 2^(i+1) divides k, the only starts a level-i lookup or block reads.
    for lg in range(1, 13):
        stride = 1 << lg
        for tb in (0, 3 * stride):
            for k in range(0, stride, 2):
                tz = lg if k == 0 else (k & -k).bit_length() - 1
                for i in range(0, min(tz, lg)):
                    want = tb + stride - (stride >> i) + (k >> (i + 1))
                    assert kernel_block_index(tb, stride, k, i) == want, (stride, tb, k, i)
                    assert want < tb + stride  # inside this table's slots

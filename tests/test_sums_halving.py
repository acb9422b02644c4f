"""The lane algorithm of group_sum_scatter (mdl_internal.hpp), run on the CPU.

The device sums of the pair passes and the control variate's per-atom terms
use recursive halving over the lane bits: at each step a lane keeps one half of
its components and sends the other half to its partner. The step kernels rely
on two properties, both checked here on a lane-by-lane model of the exchange
for every component count the kernels use (and all others up to the group
width): each component's sum lands in exactly one lane, and it equals, bit for
bit, the sum a butterfly (pairs across the highest lane bit first) gives.
"""

from __future__ import annotations

import numpy as np
import pytest


def _halved(n, steps):
    for _ in range(steps):
        n = (n + 1) // 2
    return n


def _scatter(x, mistake=False):
    """x [W, NV] -> (value, index or -1) per lane, as group_sum_scatter; with
    ``mistake`` the partner sends the half it keeps (a planted error)."""
    W, NV = x.shape
    v = [list(row) for row in x]
    lo, hi = [0] * W, [NV] * W
    S = 0
    while (W >> (S + 1)) >= 1:
        A, OFF = _halved(NV, S), (W // 2) >> S
        H = (A + 1) // 2
        new = []
        for lane in range(W):
            up, p = (lane & OFF) != 0, lane ^ OFF
            row = []
            for i in range(H):
                a = v[lane][i]
                b = v[lane][H + i] if H + i < A else 0.0
                pa = v[p][i]
                pb = v[p][H + i] if H + i < A else 0.0
                keep = b if up else a
                recv = pa if (p & OFF) else pb      # the partner sends its other half
                if mistake:
                    recv = pb if (p & OFF) else pa
                row.append(keep + recv)
            new.append(row)
        for lane in range(W):
            mid = lo[lane] + H
            if lane & OFF:
                lo[lane] = mid
            else:
                hi[lane] = min(hi[lane], mid)
        v = new
        S += 1
    return [(v[lane][0], lo[lane] if lo[lane] < hi[lane] else -1) for lane in range(W)]


def _butterfly(x):
    W = x.shape[0]
    b = x.copy()
    off = W // 2
    while off >= 1:
        b = b + b[np.arange(W) ^ off]
        off //= 2
    return b[0]


@pytest.mark.parametrize("W", [4, 16, 64])
def test_every_component_once_and_butterfly_bits(W):
    rs = np.random.RandomState(W)
    for NV in range(1, W + 1):
        x = rs.standard_normal((W, NV)) * np.exp(rs.uniform(-20, 20, (W, NV)))
        out = _scatter(x)
        got = {}
        for val, idx in out:
            if idx >= 0:
                assert idx not in got, (W, NV, idx)
                got[idx] = val
        assert sorted(got) == list(range(NV)), (W, NV)
        ref = _butterfly(x)
        for c, val in got.items():
            assert val == ref[c], (W, NV, c)


def test_a_wrong_exchange_is_caught():
    rs = np.random.RandomState(1)
    x = rs.standard_normal((16, 4))
    ref = _butterfly(x)
    got = {i: v for v, i in _scatter(x, mistake=True) if i >= 0}
    assert any(not np.isclose(got.get(c, np.nan), ref[c]) for c in range(4))

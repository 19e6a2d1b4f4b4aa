"""Launch-shape tuning (host logic, no GPU): the committed gfx950 table covers every
(order, precision, lists, log2 ni) bucket with shapes the kernels are compiled for,
and the autotuner only proposes shapes inside its stated bounds."""
from __future__ import annotations

import math
import os

import pytest

from sapporo2_amd import autotune

TABLE = os.path.join(os.path.dirname(os.path.dirname(os.path.abspath(__file__))), "sapporo2_amd", "tune",
                     "gfx950.txt")
P_COMPILED = {1, 2, 4, 8, 16, 32, 64}


def _rows():
    rows = []
    with open(TABLE) as f:
        for line in f:
            line = line.strip()
            if not line or line.startswith("#"):
                continue
            v = [int(x) for x in line.split()]
            assert len(v) == 7, line
            rows.append(v)
    return rows


def test_table_covers_every_bucket():
    """Every tuned system size log2(nj) has one entry per (order, precision, lists,
    log2 ni) bucket, for ni up to 16384 and up to nj."""
    rows = _rows()
    ljs = sorted({r[4] for r in rows})
    assert 17 in ljs                               # the block-step metric's N = 131072
    for lj in ljs:
        keys = [(o, p, g, li) for o, p, g, li, l, _, _ in rows if l == lj]
        want = {(o, p, g, li) for o in (1, 2) for p in (1, 2) for g in (0, 1) for li in range(min(15, lj + 1))}
        assert set(keys) == want, lj
        assert len(keys) == len(want), lj          # one entry per bucket


def test_table_shapes_are_valid():
    for o, p, g, li, lj, P, jlog in _rows():
        assert P in P_COMPILED, (o, p, g, li, P)
        assert P <= max(1, 2 ** (li + 1)), (li, P)  # never wider than the bucket's active count
        Q = 64 // P
        jchunk = 2 ** jlog
        assert jchunk >= 64 and jchunk % Q == 0, (li, P, jchunk)
        assert jchunk <= 2 ** lj


@pytest.mark.parametrize("ni", [1, 3, 64, 100, 1024, 16384])
@pytest.mark.parametrize("nj", [8192, 131072])
def test_candidates_within_bounds(ni, nj):
    cands = autotune.candidates(ni, nj)
    assert cands, (ni, nj)
    # the wave floor holds for every candidate unless none can meet it (then it is dropped)
    fl = autotune.MIN_WAVES
    floor = fl if any(math.ceil(ni / P) * S >= fl for P, S in cands) else 1
    for P, S in cands:
        assert P in P_COMPILED
        Q = 64 // P
        assert math.ceil(nj / S) >= max(64, 16 * Q)
        waves = math.ceil(ni / P) * S
        assert floor <= waves <= 262144


def test_select_prefers_fewer_partial_rows_within_tie():
    """Among shapes within 1 % of the fastest the tuner keeps the one writing the fewest
    partial-sum rows; a shape more than 1 % faster wins regardless."""
    flat = [(1.5839, 64, 512, 256), (1.5849, 64, 256, 512), (1.5878, 64, 128, 1024), (1.6423, 64, 64, 2048)]
    assert autotune.select(flat)[2] == 128
    steep = [(3.244, 64, 1024, 256), (3.295, 64, 512, 512), (3.3935, 64, 256, 1024)]
    assert autotune.select(steep)[2] == 1024
    assert autotune.partial_rows(256) == 64 and autotune.partial_rows(6) == 6

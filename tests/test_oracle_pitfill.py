"""Pins the oracle's pit fill (PitFillingAlgorithmCPUOpenMP.h:68-242), CPU only.

- two independent restatements agree (bucket queues with the dead-cell rule
  versus Dijkstra over values with an explicit round-trip test), on grids that
  hit non-round-tripping mm levels (251, 253, 502, ...) and nodata;
- when every level round-trips the result equals the textbook priority-flood
  fill (Barnes, Lehman, Mulla 2014) of the mm-quantized DEM;
- a hand grid where a dead seed level changes the answer;
- invariants: filled >= quantized, seeds unchanged, unreached cells keep Z;
- the quantisation rule (:84-106) and the input domain.
"""
import heapq

import numpy as np
import pytest

import oracle
from helpers import NODATA, mm_grid, rng

ROUND_TRIP = np.array([q for q in range(0, 5000)
                       if int(np.float32(1000.0) * np.float32(np.float32(q) / np.float32(1000.0))) == q])


def quantize(z, nodata=NODATA):
    q = z.copy()
    m = q != nodata
    idx = (np.float32(1000.0) * q[m]).astype(np.uint64)
    idx[(q[m] != 0) & (idx == 0)] = 1
    q[m] = idx.astype(np.float32) / np.float32(1000.0)
    return q


def textbook_fill(z, nodata=NODATA):
    """Priority-flood: every reached cell relays (no dead levels)."""
    h, w = z.shape
    out = z.copy()
    done = np.zeros((h, w), bool)
    pq = []
    for y in range(h):
        for x in range(w):
            if z[y, x] == nodata:
                continue
            if x in (0, w - 1) or y in (0, h - 1) or z[y, x] == 0:
                done[y, x] = True
                heapq.heappush(pq, (float(z[y, x]), y, x))
    while pq:
        v, y, x = heapq.heappop(pq)
        for dy in (-1, 0, 1):
            for dx in (-1, 0, 1):
                ny, nx = y + dy, x + dx
                if (dy or dx) and 0 <= ny < h and 0 <= nx < w and not done[ny, nx] \
                        and z[ny, nx] != nodata:
                    done[ny, nx] = True
                    out[ny, nx] = max(z[ny, nx], np.float32(v))
                    heapq.heappush(pq, (float(out[ny, nx]), ny, nx))
    return out


def test_non_round_trip_levels_exist():
    q = np.arange(1, 500000, dtype=np.int64)
    back = (np.float32(1000.0) * (q.astype(np.float32) / np.float32(1000.0))).astype(np.int64)
    dead = q[back != q]
    assert len(dead) == 2949 and dead[0] == 251


@pytest.mark.parametrize("seed", range(40))
def test_two_restatements_agree(seed):
    r = rng(1000 + seed)
    h, w = int(r.integers(3, 45)), int(r.integers(3, 45))
    kind = seed % 4
    if kind == 0:
        z = r.uniform(0, 5, (h, w)).astype(np.float32)
    elif kind == 1:
        z = mm_grid(r, h, w, 0, 600, 0.05)
    elif kind == 2:
        z = mm_grid(r, h, w, 240, 300, 0.1)
    else:
        z = mm_grid(r, h, w, 0, 3000, 0.05, nodata_frac=0.05)
    a, qa = oracle.pit_fill(z, NODATA)
    b, qb = oracle.pit_fill(z, NODATA, method="dijkstra")
    assert np.array_equal(qa, qb) and np.array_equal(a, b)


@pytest.mark.parametrize("seed", range(15))
def test_equals_textbook_when_all_levels_round_trip(seed):
    r = rng(2000 + seed)
    h, w = int(r.integers(3, 40)), int(r.integers(3, 40))
    z = (r.choice(ROUND_TRIP, (h, w)) / 1000.0).astype(np.float32)
    z[r.random((h, w)) < 0.03] = 0.0
    z = quantize(z)
    assert all(int(np.float32(1000) * v) == int(round(v * 1000)) for v in z.ravel())
    filled, q = oracle.pit_fill(z, NODATA)
    assert np.array_equal(q, z)
    assert np.array_equal(filled, textbook_fill(z))


def test_dead_seed_level_hand_case():
    # 0.2515 quantizes to 0.251 (q = 251), which does not survive the bucket
    # round trip: those edge seeds never expand.  The interior 0.1 cell is
    # reached only through the 0.9 cell fed by the 0 seed.
    z = np.array([[0.2515, 0.2515, 0.2515, 0.0],
                  [0.2515, 0.1, 0.9, 0.2515],
                  [0.2515, 0.2515, 0.2515, 0.2515]], np.float32)
    filled, q = oracle.pit_fill(z, NODATA)
    assert q[0, 0] == np.float32(np.float32(251) / np.float32(1000))
    assert filled[1, 1] == np.float32(0.9) and filled[1, 2] == np.float32(0.9)
    # a textbook fill would take the lower dead seeds instead
    assert textbook_fill(q)[1, 1] == q[0, 0]
    # A cell enclosed by dead values is never reached and keeps Z.  The 0
    # seed keeps the minimum at 0 (so 0.251 stays dead); its flood reaches
    # (1,2) only at the dead value 0.251 and (1,3) at 0.3, neither of which
    # can lift (1,1).
    z2 = np.array([[0.2515, 0.2515, 0.2515, 0.2515, 0.2515],
                   [0.2515, 0.1, 0.2515, 0.3, 0.2515],
                   [0.2515, 0.2515, 0.2515, 0.0, 0.2515]], np.float32)
    filled2, q2 = oracle.pit_fill(z2, NODATA)
    assert filled2[1, 1] == q2[1, 1] == np.float32(0.1)
    assert filled2[1, 3] == np.float32(0.3)
    assert textbook_fill(q2)[1, 1] == q2[0, 0]
    assert np.array_equal(filled2, oracle.pit_fill(z2, NODATA, method="dijkstra")[0])


@pytest.mark.parametrize("seed", range(10))
def test_invariants(seed):
    r = rng(3000 + seed)
    h, w = int(r.integers(3, 60)), int(r.integers(3, 60))
    z = mm_grid(r, h, w, 0, 2000, 0.03, nodata_frac=0.02)
    filled, q = oracle.pit_fill(z, NODATA)
    m = q != NODATA
    assert (filled[m] >= q[m]).all()
    assert np.array_equal(filled[~m], q[~m])
    edge = np.zeros((h, w), bool)
    edge[0, :] = edge[-1, :] = edge[:, 0] = edge[:, -1] = True
    seeds = m & (edge | (q == 0))
    assert np.array_equal(filled[seeds], q[seeds])


def test_quantisation_rule():
    z = np.array([[1e-5, 0.0, 0.0009, 0.2515, 123.4567, NODATA, 7.0]], np.float32)
    _, q = oracle.pit_fill(z, NODATA)
    assert q[0, 0] == np.float32(0.001)  # non-zero rounding to 0 becomes 1 mm
    assert q[0, 1] == 0.0
    assert q[0, 2] == np.float32(0.001)
    assert q[0, 3] == np.float32(np.float32(251) / np.float32(1000))
    assert q[0, 4] == np.float32(np.float32(123456) / np.float32(1000))
    assert q[0, 5] == NODATA and q[0, 6] == 7.0


def test_domain_errors():
    with pytest.raises(RuntimeError):
        oracle.pit_fill(np.array([[1.0, -0.5], [1.0, 1.0]], np.float32), NODATA)
    # bucket range beyond the exact float span
    with pytest.raises(RuntimeError):
        oracle.pit_fill(np.array([[0.0, 20000.0], [1.0, 1.0]], np.float32), NODATA)

"""CPU check of the closed form the GPU IVF-binary search implements.

IVFFlatBinaryIndex::searchKNN (BinaryIVF.hpp:81-108) keeps a max-heap of (dist, id) with strict
admission over candidates in scan order (probe rank, then ascending id) and pops the largest
(dist, id). Its result is not "the k smallest (dist, id)" when distances tie. The GPU kernel uses
this closed form instead of a heap:

    T  = k-th smallest distance among the scanned candidates (all of them if there are <= k)
    F  = the first k candidates in scan order with dist <= T
    result = {dist < T}  U  (the k - |{dist < T}| smallest ids among F's candidates with dist == T)

emitted ascending by (dist, id). This test compares it with a literal restatement of the heap
on random sequences with heavy ties.
"""

import heapq

import numpy as np
import pytest


def heap_reference(seq, k):
    h = []  # max-heap of (dist, id) via negation
    for d, i in seq:
        if len(h) < k or d < -h[0][0]:
            heapq.heappush(h, (-d, -i))
            if len(h) > k:
                heapq.heappop(h)
    return sorted((-a, -b) for a, b in h)


def closed_form(seq, k):
    if len(seq) <= k:
        return sorted(seq)
    ds = sorted(d for d, _ in seq)
    T = ds[k - 1]
    lt = [(d, i) for d, i in seq if d < T]
    F = [(d, i) for d, i in seq if d <= T][:k]
    ties = sorted(i for d, i in F if d == T)[: k - len(lt)]
    return sorted(lt + [(T, i) for i in ties])


@pytest.mark.parametrize("seed", range(300))
def test_closed_form_matches_heap(seed):
    rng = np.random.default_rng(seed)
    n = int(rng.integers(1, 60))
    k = int(rng.integers(1, 20))
    dmax = int(rng.integers(1, 6))
    ids = rng.permutation(1000)[:n]
    # scan order: lists in probe order, ascending ids inside a list
    nl = int(rng.integers(1, 5))
    cut = np.sort(rng.integers(0, n + 1, nl - 1))
    parts = np.split(ids, cut)
    seq = []
    for p in parts:
        for i in sorted(p):
            seq.append((int(rng.integers(0, dmax)), int(i)))
    assert closed_form(seq, k) == heap_reference(seq, k)


def test_closed_form_differs_from_lexicographic_topk():
    # the reason the closed form exists: (5,10),(5,3),(1,7) with k=2 keeps id 3, but scan order
    # (5,10) first can keep 10 when the smaller tie arrives after the heap is full
    seq = [(5, 10), (1, 7), (5, 3)]
    assert heap_reference(seq, 2) == [(1, 7), (5, 10)]
    assert closed_form(seq, 2) == [(1, 7), (5, 10)]
    assert sorted(seq)[:2] == [(1, 7), (5, 3)]

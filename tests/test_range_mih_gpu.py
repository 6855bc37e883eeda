"""GPU Hamming range search (BKTreeBinaryIndex*.find_threshold) and multi-index hashing
(MIHBinaryIndex.searchKNN) against the oracle and the reference's recorded outputs.

Range search must return exactly the oracle's (row, distance) pairs in ascending row order; the
reference returns the same pairs in its tree's order (compared as sets). MIH must equal the oracle
(candidates: a sub-string within radius // m bits; k nearest in (distance, row) order) bit for bit;
the reference's recorded lists agree in distances, and in rows wherever the k-th distance does not
tie. The reference test cases (pynear/tests/test_bktree.py, test_binary_approx.py:120-220) run
through the same calls.
"""

import os

import numpy as np
import pytest

import oracle
import pynear_amd as P
from pynear_amd import _native as N

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
BK = [(P.BKTreeBinaryIndex64, 8), (P.BKTreeBinaryIndex128, 16), (P.BKTreeBinaryIndex256, 32),
      (P.BKTreeBinaryIndex512, 64), (P.BKTreeBinaryIndexN, 1)]


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


def _csr(g, key):
    c = g[f"{key}_count"]
    off = np.concatenate([[0], np.cumsum(c)])
    return [(g[f"{key}_idx"][off[i]:off[i + 1]], g[f"{key}_dist"][off[i]:off[i + 1]]) for i in range(len(c))]


def _near(X, rng, nq, max_flips):
    src = rng.integers(0, len(X), nq)
    Q = X[src].copy()
    nbits = X.shape[1] * 8
    for j in range(nq):
        bits = rng.permutation(nbits)[:rng.integers(0, max_flips + 1)]
        Q[j, bits // 8] ^= (np.uint8(1) << (bits % 8).astype(np.uint8))
    return Q


# ---------------------------------------------------------------- range search
@pytest.mark.parametrize("thr", [8, 40, 52])
def test_range_equals_recorded_bktree(g, thr):
    t = P.BKTreeBinaryIndex128()
    t.set(g["bk_X"])
    ind, dist, keys = t.find_threshold(g["bk_Q"], thr)
    oi, od = oracle.hamming_range(g["bk_X"], g["bk_Q"], thr)
    for q, (ri, rd) in enumerate(_csr(g, f"bk_thr{thr}")):
        assert ind[q] == oi[q].tolist() and dist[q] == od[q].tolist()
        assert sorted(zip(ri.tolist(), rd.tolist())) == list(zip(ind[q], dist[q]))
        assert keys[q] == g["bk_X"][ind[q]].tolist()


@pytest.mark.parametrize("nbytes", [1, 3, 8, 16, 20, 64, 100, 130, 300])
@pytest.mark.parametrize("thr_frac", [0.0, 0.1, 0.3, 0.45, 1.0])
def test_range_matches_oracle(nbytes, thr_frac):
    """Thresholds from none to every row: the larger ones overflow the per-slice buffer of 8 and
    take the second scan; widths above 128 bytes take the wide-row scan."""
    rng = np.random.default_rng(nbytes * 7 + int(thr_frac * 100))
    X = rng.integers(0, 256, (3000, nbytes), dtype=np.uint8)
    X[50] = X[10]
    Q = np.concatenate([_near(X, rng, 12, max(1, nbytes)), rng.integers(0, 256, (5, nbytes), dtype=np.uint8)])
    thr = int(thr_frac * nbytes * 8)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    off, idx, dist = h.range(Q, thr)
    oi, od = oracle.hamming_range(X, Q, thr)
    assert off[-1] == sum(len(a) for a in oi)
    for q in range(len(Q)):
        np.testing.assert_array_equal(idx[off[q]:off[q + 1]], oi[q])
        np.testing.assert_array_equal(dist[off[q]:off[q + 1]], od[q])


def test_range_capacity_protocol_and_negative_threshold():
    rng = np.random.default_rng(3)
    X = rng.integers(0, 256, (5000, 16), dtype=np.uint8)
    Q = X[:40]
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    full = h.range(Q, 50)
    small = h.range(Q, 50, capacity=3)  # first call reports the total, the second fills
    for a, b in zip(full, small):
        np.testing.assert_array_equal(a, b)
    off, idx, dist = h.range(Q, -1)
    assert off.tolist() == [0] * 41 and len(idx) == 0
    with pytest.raises(ValueError):
        h.range(X[:3, :8].copy(), 3)
    with pytest.raises(ValueError):
        N.ExactHandle(N.METRIC_L2).range(np.zeros((1, 4), np.uint8), 1)


@pytest.mark.parametrize("cls,dim", BK)
def test_bktree_reference_cases(cls, dim):
    """pynear/tests/test_bktree.py: empty index, empty set, find self, find all, duplicates."""
    rng = np.random.default_rng(42)
    data = rng.integers(0, 256, size=(2, dim), dtype=np.uint8)
    while len({tuple(r) for r in data.tolist()}) < 2:
        data = rng.integers(0, 256, size=(2, dim), dtype=np.uint8)
    tree = cls()
    ind, dist, keys = tree.find_threshold(data, 1)
    assert ind == [[], []] and dist == [[], []]
    assert tree.empty() and tree.values() == []
    tree.set(np.array([], dtype=np.uint8))
    ind, dist, keys = tree.find_threshold(data, 1)
    assert ind == [[], []] and dist == [[], []] and tree.empty() and tree.values() == []

    tree = cls()
    tree.set(data)
    ind, dist, keys = tree.find_threshold(data, 0)
    assert ind == [[0], [1]] and dist == [[0], [0]] and keys == data[:, None, :].tolist()
    assert tree.size() == 2 and sorted(tree.values()) == sorted(data.tolist())
    ind, dist, keys = tree.find_threshold(data, dim * 8)
    pair = np.bitwise_count(data[:, None, :] ^ data[None]).sum(-1)
    assert ind == [[0, 1]] * 2 and dist == pair.tolist()
    assert keys == np.broadcast_to(data, (2, 2, dim)).tolist()

    zeros = np.zeros((2, dim), np.uint8)
    tree = cls()
    tree.set(zeros)
    ind, dist, keys = tree.find_threshold(zeros, 255)
    assert ind == [[0, 1]] * 2 and dist == [[0, 0]] * 2 and tree.size() == 2


def test_bktree_set_adds_rows():
    """BKTree::update (BKTree.hpp:65-69): a second set keeps the first rows and numbers the new
    ones after them."""
    rng = np.random.default_rng(5)
    a = rng.integers(0, 256, (100, 16), dtype=np.uint8)
    b = rng.integers(0, 256, (50, 16), dtype=np.uint8)
    t = P.BKTreeBinaryIndex128()
    t.set(a)
    t.set(b)
    assert t.size() == 150
    ind, dist, _ = t.find_threshold(b[:3], 0)
    assert [r[-1] for r in ind] == [100, 101, 102]
    with pytest.raises(ValueError):
        t.set(rng.integers(0, 256, (4, 8), dtype=np.uint8))
    gen = P.BKTreeBinaryIndexN()
    gen.set(rng.integers(0, 256, (10, 5), dtype=np.uint8))
    with pytest.raises(ValueError):
        gen.set(rng.integers(0, 256, (4, 6), dtype=np.uint8))


@pytest.mark.timeout(300)
def test_range_1m_x_512bit_equals_knn_within_threshold():
    """1M x 512-bit rows, 1000 queries at up to 60 flipped bits: the range result at threshold 60 is
    exactly the exact k-NN (k = 32) entries with distance <= 60 wherever the 32nd distance exceeds
    60; 16 queries against the oracle."""
    import bench
    X, Q, src, nflip = bench.hamming_workload(nq=1000, flip_p=0.06, with_source=True)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    off, idx, dist = h.range(Q, 60)
    ki, kd = h.search(Q, 32)
    checked = 0
    for q in range(len(Q)):
        if kd[q, -1] > 60:
            sel = kd[q] <= 60
            assert sorted(ki[q][sel].tolist()) == idx[off[q]:off[q + 1]].tolist(), q
            checked += 1
    assert checked > 900
    assert all(src[q] in idx[off[q]:off[q + 1]] for q in range(len(Q)) if nflip[q] <= 60)
    sub = np.arange(0, 1000, 64)
    oi, od = oracle.hamming_range(X, Q[sub], 60, chunk=4)
    for j, q in enumerate(sub):
        np.testing.assert_array_equal(idx[off[q]:off[q + 1]], oi[j])
        np.testing.assert_array_equal(dist[off[q]:off[q + 1]], od[j])


# ---------------------------------------------------------------- MIH
@pytest.mark.parametrize("m,radius", [(4, 8), (2, 9), (8, 16)])
def test_mih_equals_recorded_and_oracle(g, m, radius):
    X, Q = g["bk_X"], g["bk_Q"]
    t = P.MIHBinaryIndex(m)
    t.set(X)
    ind, dist = t.searchKNN(Q, 5, radius)
    oi, od = oracle.mih_search(X, Q, m, radius, 5)
    for q, (ri, rd) in enumerate(_csr(g, f"mih_m{m}_r{radius}")):
        assert ind[q] == oi[q].tolist() and dist[q] == od[q].tolist()
        assert dist[q] == rd.tolist()
        if rd.size:
            below = rd < rd.max()
            assert ri[below].tolist() == list(np.array(ind[q])[below])


@pytest.mark.parametrize("nbytes,m", [(64, 8), (64, 16), (32, 4), (16, 4), (16, 2), (16, 8), (12, 4), (30, 6),
                                      (7, 7), (40, 5), (64, 32), (24, 3)])
@pytest.mark.parametrize("k", [1, 10, 32, 100])
def test_mih_matches_oracle(nbytes, m, k):
    """Sub-strings of 1 to 8 bytes (one or two words per group), k up to 32 on the filtered scan and
    k = 100 on the distance-matrix path; radii below, at and above m."""
    rng = np.random.default_rng(nbytes * 31 + m)
    X = rng.integers(0, 256, (2500, nbytes), dtype=np.uint8)
    X[7] = X[3]
    Q = np.concatenate([_near(X, rng, 10, 3 * m), rng.integers(0, 256, (3, nbytes), dtype=np.uint8)])
    t = P.MIHBinaryIndex(m)
    t.set(X)
    for radius in (0, m - 1, 2 * m, 5 * m):
        ind, dist = t.searchKNN(Q, k, radius)
        oi, od = oracle.mih_search(X, Q, m, radius, k)
        for q in range(len(Q)):
            assert ind[q] == oi[q].tolist(), (radius, q)
            assert dist[q] == od[q].tolist(), (radius, q)


def _make_db(n, nbytes, seed=0):
    return np.random.default_rng(seed).integers(0, 256, (n, nbytes), dtype=np.uint8)


def test_mih_reference_cases():
    """pynear/tests/test_binary_approx.py:122-220 through the same calls."""
    db = _make_db(2000, 64)
    rng = np.random.default_rng(1)
    ti = list(range(0, 200, 20))
    q = db[ti].copy()
    for j in range(len(ti)):
        bits = rng.permutation(512)[:5]
        q[j, bits // 8] ^= (np.uint8(1) << (bits % 8).astype(np.uint8))
    idx = P.MIHBinaryIndex(m=8)
    idx.set(db)
    res_idx, res_dist = idx.searchKNN(q, k=10, radius=8)
    for i, true_i in enumerate(ti):
        assert true_i in res_idx[i]
    assert all(d == sorted(d) for d in res_dist)
    assert idx.m() == 8 and idx.n() == 2000 and idx.nbytes() == 64
    with pytest.raises(ValueError):
        P.MIHBinaryIndex(m=5).set(_make_db(100, 64))
    with pytest.raises(ValueError):
        P.MIHBinaryIndex(m=4).set(_make_db(100, 64))
    r0, d0 = idx.searchKNN(db[:1], k=1, radius=0)
    assert 0 in r0[0] and d0[0][0] == 0
    e = P.MIHBinaryIndex(m=8)
    e.set(_make_db(0, 64))
    assert e.n() == 0
    idx2 = P.MIHBinaryIndex(m=8)
    small = _make_db(10, 64)
    idx2.set(small)
    r, _ = idx2.searchKNN(small[:2], k=5, radius=0)
    assert 0 in r[0] and 1 in r[1]


def test_mih_errors():
    idx = P.MIHBinaryIndex(m=8)
    with pytest.raises(RuntimeError):
        idx.searchKNN(_make_db(2, 64), 3)
    idx.set(_make_db(50, 64))
    with pytest.raises(ValueError):
        idx.searchKNN(_make_db(2, 32), 3)
    with pytest.raises(ValueError):
        idx.searchKNN(_make_db(2, 64), 3, radius=-1)
    with pytest.raises(ValueError):
        P.MIHBinaryIndex(m=64).set(_make_db(10, 64))  # 64 sub-strings: beyond the GPU layout
    with pytest.raises(ValueError):
        P.MIHBinaryIndex(m=0)


@pytest.mark.timeout(300)
def test_mih_1m_x_512bit():
    """1M x 512-bit rows, m = 8, radius 16, k = 10, 1000 queries: every candidate within radius is
    a source row at <= 16 flips, found first; 16 queries against the oracle."""
    import bench
    X, Q, src, nflip = bench.hamming_workload(nq=1000, flip_p=0.02, with_source=True)
    t = P.MIHBinaryIndex(8)
    t.set(X)
    ind, dist = t.searchKNN(Q, 10, 16)
    for q in range(len(Q)):
        assert dist[q] == sorted(dist[q])
        if nflip[q] <= 16:
            assert ind[q][0] == src[q] and dist[q][0] == nflip[q]
    sub = np.arange(0, 1000, 64)
    oi, od = oracle.mih_search(X, Q[sub], 8, 16, 10, chunk=2)
    for j, q in enumerate(sub):
        assert ind[q] == oi[j].tolist() and dist[q] == od[j].tolist()

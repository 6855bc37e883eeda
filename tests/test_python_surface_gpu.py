"""The pynear-shaped Python classes on the GPU against outputs recorded from the reference module.

tests/golden/ref_vectors.npz holds what the reference's own ``_pynear`` classes returned on seeded
inputs (tests/golden/make_golden.py). These tests run the same calls through ``pynear_amd`` and
compare the Python results as a user of pynear sees them: farthest-first lists from ``searchKNN``
(VPTree.hpp:621-631), flat lists from ``search1NN``, ``TypeError``/``ValueError`` from the binary
wrapper (pynear/__init__.py:56-104), the free distance functions and pickling.

Distances must be equal. Indices must be equal wherever the distance does not tie; the reference
VP-tree orders equal distances by its randomly seeded tree walk (VPTree.hpp:120-124), so at a tie
only the set of rows matters, and at the k-th distance any rows of that distance are correct.
"""

import os
import pickle

import numpy as np
import pytest

import oracle
import pynear_amd as P
from pynear_amd import _native as N

pytestmark = pytest.mark.gpu

GOLD = os.path.join(os.path.dirname(__file__), "golden")
FLOAT = {"l2": (P.VPTreeL2Index, N.METRIC_L2), "l1": (P.VPTreeL1Index, N.METRIC_L1),
         "linf": (P.VPTreeChebyshevIndex, N.METRIC_CHEBYSHEV)}


@pytest.fixture(scope="module")
def g():
    return np.load(os.path.join(GOLD, "ref_vectors.npz"))


def _true_hamming(X, Q):
    return np.bitwise_count(Q[:, None, :] ^ X[None, :, :]).sum(-1).astype(np.int64)


def assert_same_up_to_ties(ours_idx, ours_dist, ref_idx, ref_dist, true_dist):
    """Row by row: equal distance lists; every listed index has its listed distance; rows of a
    distance below the last one listed agree as sets (the last distance may be cut by k)."""
    ours_idx, ref_idx = np.asarray(ours_idx), np.asarray(ref_idx)
    ours_dist, ref_dist = np.asarray(ours_dist), np.asarray(ref_dist)
    np.testing.assert_array_equal(ours_dist, ref_dist)
    for q in range(ours_idx.shape[0]):
        np.testing.assert_array_equal(true_dist[q, ours_idx[q]], ours_dist[q])
        np.testing.assert_array_equal(true_dist[q, ref_idx[q]], ref_dist[q])
        edge = ours_dist[q].max()
        below = ours_dist[q] != edge
        assert sorted(ours_idx[q][below]) == sorted(ref_idx[q][ref_dist[q] != edge]), q


def _float_true(metric, X, Q):
    """The oracle's distance of every (query, row) pair: reference arithmetic, bit for bit."""
    i, d = oracle.knn_float(metric, X, Q, X.shape[0])
    out = np.empty_like(d)
    np.put_along_axis(out, i, d, axis=1)
    return out


@pytest.mark.parametrize("name", ["l2", "l1", "linf"])
def test_float_vptree_classes_equal_recorded_lists(g, name):
    cls, metric = FLOAT[name]
    X, Q = g[f"{name}_X"], g[f"{name}_Q"]
    t = cls()
    t.set(X)
    idx, dist = t.searchKNN(Q, 5)
    assert isinstance(idx, list) and isinstance(idx[0], list) and isinstance(idx[0][0], int)
    assert isinstance(dist[0][0], float)
    true = _float_true(metric, X, Q)
    assert_same_up_to_ties(idx, np.array(dist, np.float32), g[f"{name}_idx"], g[f"{name}_dist"], true)
    # farthest-first, as the reference returns them
    assert all(r[0] >= r[-1] for r in dist)
    i1, d1 = t.search1NN(Q)
    assert isinstance(i1, list) and isinstance(i1[0], int) and isinstance(d1[0], float)
    np.testing.assert_array_equal(np.array(d1, np.float32), g[f"{name}_dist1"])
    np.testing.assert_array_equal(np.array(i1), g[f"{name}_idx1"])


@pytest.mark.parametrize("name", ["l2", "l1", "linf"])
def test_five_point_fixture(g, name):
    """pynear/tests/test_vptree.py:282-301: 5 points on a line, k = 4 > the half the split keeps."""
    cls, _ = FLOAT[name]
    t = cls()
    t.set(g["five_X"])
    idx, dist = t.searchKNN(g["five_Q"], 4)
    assert idx == [[3, 2, 0, 1]]
    assert idx == g[f"five_{name}_idx"].tolist()
    np.testing.assert_array_equal(np.array(dist, np.float32), g[f"five_{name}_dist"])


@pytest.mark.parametrize("nb", [64, 16, 8, 12, 6, 17])
def test_binary_width_dispatch_equals_recorded_lists(g, nb):
    """The fixed classes at 64/16/8 bytes and the generic class at 12, 6 and 17 bytes, through the
    dispatching VPTreeBinaryIndex and through the class the reference dispatches to."""
    X, Q = g[f"binw{nb}_X"], g[f"binw{nb}_Q"]
    true = _true_hamming(X, Q)
    direct = {64: P.VPTreeBinaryIndex512, 16: P.VPTreeBinaryIndex128, 8: P.VPTreeBinaryIndex64}.get(nb, P.VPTreeBinaryIndexN)
    for t in (P.VPTreeBinaryIndex(), direct()):
        t.set(X)
        idx, dist = t.searchKNN(Q, 4)
        assert isinstance(dist[0][0], int)
        assert_same_up_to_ties(idx, dist, g[f"binw{nb}_idx"], g[f"binw{nb}_dist"], true)
        i1, d1 = t.search1NN(Q)
        np.testing.assert_array_equal(d1, g[f"binw{nb}_dist1"])
        np.testing.assert_array_equal(true[np.arange(len(Q)), i1], d1)
    if nb in P.VPTreeBinaryIndex._BY_WIDTH:
        t = P.VPTreeBinaryIndex()
        t.set(X)
        assert type(t._index) is direct


def test_binary_256_class_equals_recorded_lists(g):
    X, Q = g["bin_X"], g["bin_Q"]
    t = P.VPTreeBinaryIndex256()
    t.set(X)
    idx, dist = t.searchKNN(Q, 3)
    assert_same_up_to_ties(idx, dist, g["bin_idx"], g["bin_dist"], _true_hamming(X, Q))


def test_binary_odd_width_above_127_bits_is_true_popcount(g):
    """At 33 bytes the reference's generic Hamming sums into an int8_t (DistanceFunctions.hpp:125-146),
    so its distances above 127 wrap negative and its neighbours are not the nearest rows. This
    library returns the true popcount ranking; the recorded reference distances are the int8 wrap
    of the true distance of the rows it returned."""
    X, Q = g["binw33_X"], g["binw33_Q"]
    true = _true_hamming(X, Q)
    ref_idx, ref_dist = g["binw33_idx"], g["binw33_dist"]
    np.testing.assert_array_equal(np.take_along_axis(true, ref_idx, 1).astype(np.int8), ref_dist)
    t = P.VPTreeBinaryIndex()
    t.set(X)
    idx, dist = t.searchKNN(Q, 4)
    oi, od = oracle.knn_hamming(X, Q, 4)
    np.testing.assert_array_equal(np.array(idx)[:, ::-1], oi)
    np.testing.assert_array_equal(np.array(dist)[:, ::-1], od)
    assert (ref_dist < 0).all() and (np.array(dist) > 0).all()


def test_binary_wrapper_errors():
    t = P.VPTreeBinaryIndex()
    with pytest.raises(TypeError):
        t.set(np.zeros((4, 8), np.float32))
    with pytest.raises(ValueError):
        t.set(np.zeros(8, np.uint8))
    t.set(np.zeros((4, 8), np.uint8))
    with pytest.raises(ValueError):  # query width differs from the indexed width
        t.searchKNN(np.zeros((2, 16), np.uint8), 1)
    with pytest.raises(TypeError):
        t.search1NN(np.zeros((2, 8), np.int32))
    with pytest.raises(ValueError):
        P.VPTreeBinaryIndex512().set(np.zeros((4, 32), np.uint8))
    assert P.VPTreeBinaryIndex().search1NN(np.zeros((1, 8), np.uint8)) == ([], [])
    with pytest.raises(ValueError):
        P.dist_hamming([1, 2, 3], [1, 2])
    with pytest.raises(ValueError):
        P.dist_hamming([0] * 12, [0] * 12)


def test_binary_duplicates():
    """pynear/tests/test_vptree.py:183-195: two equal rows, k = 2, both returned at distance 0."""
    data = np.zeros((2, 32), np.uint8)
    t = P.VPTreeBinaryIndex()
    t.set(data)
    idx, dist = t.searchKNN(data, 2)
    assert [sorted(i) for i in idx] == [[0, 1]] * 2
    assert dist == [[0, 0]] * 2


def test_free_distance_functions_equal_recorded_values(g):
    fa, fb = g["dist_fa"], g["dist_fb"]
    for name, fn in (("l2", P.dist_l2), ("l1", P.dist_l1), ("linf", P.dist_chebyshev)):
        got = np.array([fn(fa[j], fb[j]) for j in range(len(fa))], np.float32)
        np.testing.assert_array_equal(got, g[f"dist_{name}"])
    for nb, fn in ((64, P.dist_hamming_512), (32, P.dist_hamming_256), (16, P.dist_hamming_128), (8, P.dist_hamming_64)):
        a, b = g[f"dist_h{nb}_a"], g[f"dist_h{nb}_b"]
        exp = g[f"dist_h{nb}"]
        assert [fn(list(a[j]), list(b[j])) for j in range(len(a))] == exp.tolist()
        assert [P.dist_hamming(list(a[j]), list(b[j])) for j in range(len(a))] == exp.tolist()


@pytest.mark.parametrize("cls,data", [
    (P.VPTreeL2Index, "f"), (P.VPTreeL1Index, "f"), (P.VPTreeChebyshevIndex, "f"),
    (P.VPTreeBinaryIndex512, 64), (P.VPTreeBinaryIndex256, 32), (P.VPTreeBinaryIndex128, 16),
    (P.VPTreeBinaryIndex64, 8), (P.VPTreeBinaryIndexN, 20)])
def test_pickle_round_trip(cls, data):
    rng = np.random.default_rng(7)
    if data == "f":
        X = rng.random((3000, 12)).astype(np.float32)
        Q = rng.random((17, 12)).astype(np.float32)
    else:
        X = rng.integers(0, 256, (3000, data), dtype=np.uint8)
        Q = rng.integers(0, 256, (17, data), dtype=np.uint8)
    t = cls()
    t.set(X)
    before = t.searchKNN(Q, 6)
    t2 = pickle.loads(pickle.dumps(t))
    assert t2.searchKNN(Q, 6) == before
    assert t2.search1NN(Q) == t.search1NN(Q)
    empty = pickle.loads(pickle.dumps(cls()))
    with pytest.raises(RuntimeError):
        empty.searchKNN(Q, 1)


@pytest.mark.parametrize("name", ["l2", "l1", "linf"])
def test_k_larger_than_dataset(name):
    """pynear/tests/test_vptree.py:305-328 (query larger than the dataset): k > n returns all n rows."""
    cls, metric = FLOAT[name]
    rng = np.random.default_rng(8)
    X = rng.random((5, 8)).astype(np.float32)
    Q = rng.random((8, 8)).astype(np.float32)
    t = cls()
    t.set(X)
    idx, dist = t.searchKNN(Q, 9)
    assert all(len(r) == 5 for r in idx)
    oi, od = oracle.knn_float(metric, X, Q, 5)
    np.testing.assert_array_equal(np.array(idx)[:, ::-1], oi)
    np.testing.assert_array_equal(np.array(dist, np.float32)[:, ::-1], od)


def test_search1nn_at_100k_rows_all_queries():
    """100k x 32 U[0,1) rows, 1000 queries (seed 0), L2 search1NN: every query against the oracle."""
    import bench
    X, Q = bench.l2_workload()
    t = P.VPTreeL2Index()
    t.set(X)
    i1, d1 = t.search1NN(Q)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, 1)
    np.testing.assert_array_equal(np.array(d1, np.float32), od[:, 0])
    np.testing.assert_array_equal(np.array(i1), oi[:, 0])

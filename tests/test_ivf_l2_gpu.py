"""GPU parity of kmeans_l2 and IVFFlatL2Index.

kmeans_l2 must be identical to the reference (labels and centroids bit for bit). The IVF search
returns fp32 squared L2 distances over the probed clusters. Its neighbours must be the exact
(fp64) top-k of the rows in the probed clusters up to near ties (tests/neighbours.py): a row may
be swapped only within TOL[precision] of the k-th distance, the width over which the candidate
precision cannot order rows before the fp32 rerank.
"""

import os
import pickle

import numpy as np
import pytest

import oracle
import pynear_amd
from neighbours import near_tie_check, nearest_lists, sq_l2_rows
from pynear_amd import IVFFlatL2Index

pytestmark = pytest.mark.gpu
GOLD = os.path.join(os.path.dirname(__file__), "golden")
# relative width around the k-th distance inside which rows may be swapped, per candidate precision
TOL = {"bf16x3": 1e-5, "bf16": 1e-2, "fp8": 5e-2}


@pytest.mark.parametrize("n,d,k,seed", [(500, 32, 20, 42), (1000, 8, 10, 1), (300, 37, 7, 3), (2000, 64, 40, 42),
                                        (800, 16, 12, 42), (3000, 100, 30, 9), (64, 5, 64, 42)])
def test_kmeans_identical_to_reference_restatement(n, d, k, seed):
    rng = np.random.default_rng(n + d)
    X = rng.random((n, d)).astype(np.float32)
    gl, gc = pynear_amd.kmeans_l2(X, k, 100, seed)
    ol, oc = oracle.kmeans_l2(X, k, 100, seed)
    np.testing.assert_array_equal(gl, ol)
    np.testing.assert_array_equal(gc, oc)


def test_kmeans_matches_recorded_reference_outputs():
    g = np.load(os.path.join(GOLD, "ref_vectors.npz"))
    gl, gc = pynear_amd.kmeans_l2(g["km_X"], 12, 100, 42)
    np.testing.assert_array_equal(gl, g["km_labels"])
    np.testing.assert_array_equal(gc, g["km_centroids"])


def test_kmeans_errors():
    with pytest.raises(RuntimeError):
        pynear_amd.kmeans_l2(np.zeros((5, 3), np.float32), 6)
    with pytest.raises(RuntimeError):
        pynear_amd.kmeans_l2(np.zeros(5, np.float32), 2)


def _gmm(n, d, comps, seed, sigma=0.3):
    rng = np.random.default_rng(seed)
    centres = rng.standard_normal((comps, d)).astype(np.float32)
    lab = rng.integers(0, comps, n)
    return (centres[lab] + sigma * rng.standard_normal((n, d))).astype(np.float32)


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp8"])
@pytest.mark.parametrize("d", [8, 32, 100])
def test_full_probe_is_exact(precision, d):
    X = _gmm(4000, d, 30, seed=d)
    Q = _gmm(60, d, 30, seed=d + 1)
    idx = IVFFlatL2Index(n_clusters=25, n_probe=25)
    idx.precision = precision
    idx.set(X)
    gi, gd = idx.searchKNN(Q, 10)
    rows = np.arange(len(X))
    for i in range(len(Q)):
        near_tie_check(gi[i], gd[i], rows, sq_l2_rows(X, Q[i], rows), 10, TOL[precision])


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
@pytest.mark.parametrize("n_probe", [1, 3, 8])
def test_partial_probe_is_exact_over_probed_lists(n_probe, precision):
    """The probed lists are the n_probe nearest centroids (fp64, queries without a near tie at the
    n_probe-th centroid); the neighbours are the exact top-k of their rows up to near ties, and
    the numpy restatement of pynear/forest.py:63-111 returns the same distances for shared rows."""
    X = _gmm(5000, 48, 40, seed=7)
    Q = _gmm(50, 48, 40, seed=8)
    idx = IVFFlatL2Index(n_clusters=40, n_probe=n_probe)
    idx.precision = precision
    idx.set(X)
    gi, gd = idx.searchKNN(Q, 10)
    lists = np.arange(len(idx._list_centroids))
    chosen, clear = nearest_lists(Q, idx._centroids[idx._list_centroids], lists, n_probe)
    assert clear.mean() > 0.9
    for i in np.flatnonzero(clear):
        rows = np.sort(np.concatenate([idx._orig_indices[c] for c in chosen[i]]))
        near_tie_check(gi[i], gd[i], rows, sq_l2_rows(X, Q[i], rows), 10, TOL[precision])
    cent = idx._centroids[idx._list_centroids]
    oi, od = oracle.ivf_l2_search(cent, idx._cluster_data, idx._cluster_norms, idx._orig_indices, Q, n_probe, 10)
    for a, b, da, db in zip(gi, oi, gd, od):
        for c in set(a) & set(b):  # the restatement uses the expanded form ||x||^2 - 2 q.x + ||q||^2
            assert da[a.index(c)] == pytest.approx(db[b.index(c)], rel=1e-4, abs=1e-3)


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp8"])
def test_bounded_second_pass_keeps_results(precision):
    """The list scan runs the first probe, then the other probes with a per-query bound from the
    first (ivfl2.pass_margin_pct); switching the bound off (one pass over all probes) must give
    the same neighbours and distances. Small clusters (fewer rows than k) leave a query unbounded."""
    from pynear_amd import _native as N
    X = _gmm(12000, 40, 60, seed=11, sigma=0.4)
    Q = _gmm(300, 40, 60, seed=12, sigma=0.4)
    h = N.IvfL2Handle()
    h.build(X, 80, 20, 5)
    old = N.tune_get("ivfl2.pass_margin_pct")
    try:
        for k, n_probe in ((1, 4), (10, 8), (32, 6), (10, 80)):
            N.tune_set("ivfl2.pass_margin_pct", 50)
            a = h.search(Q, n_probe, k, precision)
            N.tune_set("ivfl2.pass_margin_pct", 0)
            b = h.search(Q, n_probe, k, precision)
            for x, y in zip(a, b):
                assert np.array_equal(x, y), (k, n_probe)
    finally:
        N.tune_set("ivfl2.pass_margin_pct", old)


@pytest.mark.parametrize("k,n_probe,nlist", [(10, 8, 40), (20, 6, 40), (1, 3, 40), (10, 3, 3), (40, 5, 12)])
def test_pruned_second_pass_equals_unpruned(k, n_probe, nlist):
    """The fp8 scan's second pass skips the row groups whose lower bound (the group's fp8 centre
    and radius) rules out every query of a query group (ivfl2.prune). Its neighbours and
    distances must equal those of the same pass without the skipping, bit for bit; nlist = 3 gives
    lists longer than one LDS segment of row norms (8,192 rows)."""
    from pynear_amd import _native as N
    X = _gmm(30000, 96, 24, seed=41, sigma=0.3)
    Q = _gmm(2000, 96, 24, seed=42, sigma=0.3)
    h = N.IvfL2Handle()
    h.build(X, nlist, 10, 3)
    old = N.tune_get("ivfl2.prune")
    try:
        N.tune_set("ivfl2.prune", 1)
        a = h.search(Q, n_probe, k, "fp8")
        N.tune_set("ivfl2.prune", 0)
        b = h.search(Q, n_probe, k, "fp8")
    finally:
        N.tune_set("ivfl2.prune", old)
    for x, y in zip(a, b):
        assert np.array_equal(x, y), (k, n_probe, nlist)


@pytest.mark.parametrize("precision", ["bf16", "bf16x3", "fp8"])
def test_repeated_searches_identical(precision):
    """Lists probed by a few hundred queries each split into several work items that workgroups
    take in whatever order they reach the queue: repeated searches still return identical
    neighbours and distances."""
    from pynear_amd import _native as N
    X = _gmm(30000, 64, 24, seed=31, sigma=0.3)
    Q = _gmm(3000, 64, 24, seed=32, sigma=0.3)
    h = N.IvfL2Handle()
    h.build(X, 24, 10, 7)
    for k, n_probe in ((10, 4), (32, 3)):
        first = h.search(Q, n_probe, k, precision)
        for _ in range(2):
            again = h.search(Q, n_probe, k, precision)
            for x, y in zip(first, again):
                assert np.array_equal(x, y), (k, n_probe)


@pytest.mark.parametrize("precision", ["bf16", "fp8"])
@pytest.mark.parametrize("k", [5, 20, 40])
def test_given_probes_with_holes_equal_fixed_probes(precision, k):
    """Probes given per query with varying counts and unusable slots (ids of empty clusters, -1)
    give the same neighbours as the fixed-nprobe search over the same lists; k above 16 takes the
    wider per-lane lists."""
    from pynear_amd import _native as N, hipdev
    X = _gmm(15000, 48, 50, seed=21)
    Q = _gmm(200, 48, 50, seed=22)
    h = N.IvfL2Handle()
    labels, cent = h.build(X, 60, 15, 3)
    keep = np.flatnonzero(np.bincount(labels, minlength=60))
    d2 = ((Q[:, None, :].astype(np.float64) - cent[keep][None].astype(np.float64)) ** 2).sum(-1)
    order = keep[np.argsort(d2, axis=1, kind="stable")]
    rng = np.random.default_rng(23)
    counts = rng.integers(1, 9, len(Q)).astype(np.int32)
    np_max = 12
    probes = np.full((len(Q), np_max), -1, np.int32)
    empty = np.setdiff1d(np.arange(60), keep)
    for i in range(len(Q)):
        pr = order[i, :counts[i]].tolist()
        if empty.size and i % 3 == 0:  # an empty cluster's id inside the span is skipped
            pr.insert(1, int(empty[0]))
        probes[i, :len(pr)] = pr
        counts[i] = len(pr)
    dq, dpr, dc = hipdev.DeviceArray.from_numpy(Q), hipdev.DeviceArray.from_numpy(probes), hipdev.DeviceArray.from_numpy(counts)
    oi = hipdev.DeviceArray((len(Q), k), np.int32)
    od = hipdev.DeviceArray((len(Q), k), np.float32)
    oc = hipdev.DeviceArray((len(Q),), np.int32)
    h.search_probes_device(dq.ptr, len(Q), dpr.ptr, dc.ptr, np_max, k, oi.ptr, od.ptr, oc.ptr, precision)
    gi, gd, gc = oi.numpy(), od.numpy(), oc.numpy()
    # reference: exact top-k over the same lists per query (fp64), then agreement up to near ties
    for i in range(len(Q)):
        lists = [c for c in probes[i, :counts[i]] if c >= 0 and c in set(keep.tolist())]
        rows = np.flatnonzero(np.isin(labels, lists))
        assert gc[i] == min(k, len(rows))
        near_tie_check(gi[i, :gc[i]], gd[i, :gc[i]], rows, sq_l2_rows(X, Q[i], rows), k, TOL[precision])


def test_reference_ivfflat_behaviour():
    """pynear/tests/test_ivfflat.py:44-161 through the same calls."""
    rng = np.random.default_rng(0)
    data = rng.random((500, 32)).astype(np.float32)
    queries = rng.random((20, 32)).astype(np.float32)
    index = IVFFlatL2Index(n_clusters=20, n_probe=20)
    index.set(data)
    ai, ad = index.searchKNN(queries, 5)
    ei = [np.argsort(((data - q) ** 2).sum(1))[:5].tolist() for q in queries]
    assert all(set(a) == set(e) for a, e in zip(ai, ei))
    assert all(d == sorted(d) for d in ad)
    # recall at n_probe 15 of 20 on low-dim data
    rng = np.random.default_rng(1)
    data = rng.random((2000, 8)).astype(np.float32)
    queries = rng.random((50, 8)).astype(np.float32)
    index = IVFFlatL2Index(n_clusters=20, n_probe=15)
    index.set(data)
    ai, _ = index.searchKNN(queries, 10)
    ei = [np.argsort(((data - q) ** 2).sum(1))[:10] for q in queries]
    assert np.mean([len(set(a) & set(e.tolist())) / 10 for a, e in zip(ai, ei)]) >= 0.9
    # search1NN == searchKNN(k=1); 1-D query; k larger than a cluster; n_clusters capped at N
    rng = np.random.default_rng(2)
    data = rng.random((500, 16)).astype(np.float32)
    q = rng.random((10, 16)).astype(np.float32)
    index = IVFFlatL2Index(n_clusters=10, n_probe=10)
    index.set(data)
    assert index.search1NN(q) == index.searchKNN(q, 1)
    i1, _ = index.searchKNN(q[0], 3)
    assert len(i1) == 1 and len(i1[0]) == 3
    small = IVFFlatL2Index(n_clusters=20, n_probe=20)
    small.set(rng.random((50, 4)).astype(np.float32))
    si, _ = small.searchKNN(rng.random((1, 4)).astype(np.float32), k=10)
    assert len(si[0]) == 10
    tiny = IVFFlatL2Index(n_clusters=100, n_probe=100)
    tiny.set(rng.random((10, 4)).astype(np.float32))
    assert tiny.n_clusters <= 10
    wide = IVFFlatL2Index(n_clusters=40, n_probe=10)
    wide.set(rng.random((2000, 1024)).astype(np.float32))
    wi, _ = wide.searchKNN(rng.random((5, 1024)).astype(np.float32), k=5)
    assert len(wi) == 5 and all(len(r) == 5 for r in wi)
    with pytest.raises(RuntimeError):
        IVFFlatL2Index().searchKNN(q, 3)


def test_pickle_roundtrip_rebuilds_device_index():
    rng = np.random.default_rng(8)
    data = rng.random((500, 32)).astype(np.float32)
    queries = rng.random((10, 32)).astype(np.float32)
    index = IVFFlatL2Index(n_clusters=20, n_probe=7)
    index.set(data)
    before = index.searchKNN(queries, 5)
    index2 = pickle.loads(pickle.dumps(index))
    assert index2._dev is None
    after = index2.searchKNN(queries, 5)
    assert before[0] == after[0]
    np.testing.assert_allclose(before[1], after[1])


def test_device_entry_point_matches_host():
    from pynear_amd import _native as N, hipdev
    X = _gmm(6000, 64, 50, seed=3)
    Q = _gmm(200, 64, 50, seed=4)
    h = N.IvfL2Handle()
    h.build(X, 50, 100, 42)
    hi, hd, hc = h.search(Q, 8, 10)
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((200, 10), np.int32)
    od = hipdev.DeviceArray((200, 10), np.float32)
    oc = hipdev.DeviceArray((200,), np.int32)
    st = hipdev.Stream()
    h.search_device(dq.ptr, 200, 8, 10, oi.ptr, od.ptr, oc.ptr, "bf16", st.handle)
    st.synchronize()
    np.testing.assert_array_equal(oc.numpy(), hc)
    np.testing.assert_array_equal(oi.numpy(), hi)
    np.testing.assert_array_equal(od.numpy(), hd)


@pytest.mark.parametrize("k", [65, 100, 500])
@pytest.mark.parametrize("n_probe", [1, 5, 40])
def test_k_above_64_scores_every_probed_row(k, n_probe):
    """k > 64 (no limit in forest.py:84-105): every row of the probed lists is scored in fp32, so
    the result is the exact top-k of those rows up to fp32 near ties, the same for every
    precision; a query with fewer probed rows than k gets all of them."""
    X = _gmm(8000, 48, 40, seed=41)
    Q = _gmm(40, 48, 40, seed=42)
    idx = IVFFlatL2Index(n_clusters=40, n_probe=n_probe)
    idx.set(X)
    gi, gd = idx.searchKNN(Q, k)
    idx.precision = "fp8"
    fi, fd = idx.searchKNN(Q, k)
    assert gi == fi and gd == fd
    lists = np.arange(len(idx._list_centroids))
    chosen, clear = nearest_lists(Q, idx._centroids[idx._list_centroids], lists, n_probe)
    assert clear.mean() > 0.9
    for i in np.flatnonzero(clear):
        rows = np.sort(np.concatenate([idx._orig_indices[c] for c in chosen[i]]))
        assert len(gi[i]) == min(k, len(rows))
        near_tie_check(gi[i], gd[i], rows, sq_l2_rows(X, Q[i], rows), k, 1e-5)


def test_k_above_64_device_entry_rejects():
    from pynear_amd import _native as N, hipdev
    X = _gmm(2000, 16, 10, seed=43)
    h = N.IvfL2Handle()
    h.build(X, 10, 10, 42)
    dq = hipdev.DeviceArray.from_numpy(X[:4])
    o = hipdev.DeviceArray((4, 100), np.int32)
    with pytest.raises(ValueError):
        h.search_device(dq.ptr, 4, 3, 100, o.ptr, o.ptr, o.ptr)
    idx, dist, cnt = h.search(X[:4], 10, 2000)  # every row, k = n
    assert (cnt == 2000).all() and (idx[:, 0] == np.arange(4)).all() and (dist[:, 0] == 0).all()

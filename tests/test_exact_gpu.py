"""GPU parity of the exact search (pn_exact) against the CPU oracle.

Hamming results must be bit-exact, indices included, with ties broken by lower index. Float
results must match the oracle's distances bit for bit (the kernels use the reference's operation
order) and its (distance, index) order.
"""

import numpy as np
import pytest

import oracle
from pynear_amd import _native as N

pytestmark = pytest.mark.gpu


def _hamming_case(n, nbytes, nq, seed, dup_frac=0.0):
    rng = np.random.default_rng(seed)
    X = rng.integers(0, 256, (n, nbytes), dtype=np.uint8)
    if dup_frac:
        m = int(n * dup_frac)
        X[rng.integers(0, n, m)] = X[rng.integers(0, n, m)]
    Q = X[rng.integers(0, n, nq)].copy()
    flips = rng.integers(0, 256, Q.shape, dtype=np.uint8) & rng.integers(0, 256, Q.shape, dtype=np.uint8)
    Q[: nq // 2] ^= flips[: nq // 2]
    return X, Q


@pytest.mark.parametrize("nbytes", [4, 8, 16, 20, 32, 64, 100, 130])
@pytest.mark.parametrize("k", [1, 5, 10, 32])
def test_hamming_widths_and_k(nbytes, k):
    X, Q = _hamming_case(3001, nbytes, 70, seed=nbytes * 100 + k)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    gi, gd = h.search(Q, k)
    oi, od = oracle.knn_hamming(X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("k", [1, 7, 33, 300])
def test_hamming_ties_lower_index_wins(k):
    # few distinct values -> massive ties at every rank
    rng = np.random.default_rng(7)
    X = (rng.integers(0, 4, (5000, 8)) * 3).astype(np.uint8)
    Q = X[:40].copy()
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    gi, gd = h.search(Q, k)
    oi, od = oracle.knn_hamming(X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


def test_hamming_config2_shape_sample():
    """The 512-bit workload (uniform rows, Binomial(512, 0.1) bit flips) at 200k rows."""
    import bench
    X, Q = bench.hamming_workload(200_000, 2_000, seed=1)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    gi, gd = h.search(Q, 10)
    sub = np.arange(0, len(Q), 10)
    oi, od = oracle.knn_hamming(X, Q[sub], 10)
    np.testing.assert_array_equal(gd[sub], od)
    np.testing.assert_array_equal(gi[sub], oi)
    # size-independent properties on every query: sorted, source row first at its flip distance
    assert (np.diff(gd, axis=1) >= 0).all()


@pytest.mark.parametrize("metric", [N.METRIC_L2, N.METRIC_L1, N.METRIC_CHEBYSHEV])
@pytest.mark.parametrize("d", [1, 2, 3, 5, 8, 13, 16, 31, 32, 33, 64, 65, 100])
def test_float_bit_exact(metric, d):
    rng = np.random.default_rng(1000 + d * 7 + metric)
    X = rng.random((2500, d)).astype(np.float32)
    Q = rng.random((45, d)).astype(np.float32)
    h = N.ExactHandle(metric)
    h.set(X)
    for k in (1, 5, 16):
        gi, gd = h.search(Q, k)
        oi, od = oracle.knn_float(metric, X, Q, k)
        np.testing.assert_array_equal(gd, od)
        np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("metric", [N.METRIC_L2, N.METRIC_L1, N.METRIC_CHEBYSHEV])
def test_float_vs_fp64(metric):
    """Distances within rtol 1e-5 of FP64 and equal index sets (no near-ties in this data)."""
    rng = np.random.default_rng(5)
    X = rng.standard_normal((4000, 24)).astype(np.float32)
    Q = rng.standard_normal((30, 24)).astype(np.float32)
    h = N.ExactHandle(metric)
    h.set(X)
    gi, gd = h.search(Q, 8)
    fi, fd = oracle.knn_fp64(metric, X, Q, 8)
    np.testing.assert_allclose(gd, fd, rtol=1e-5)
    np.testing.assert_array_equal(gi, fi)


@pytest.mark.parametrize("metric", [N.METRIC_L2, N.METRIC_L1, N.METRIC_CHEBYSHEV, N.METRIC_HAMMING])
def test_k_at_and_above_n(metric):
    rng = np.random.default_rng(11)
    if metric == N.METRIC_HAMMING:
        X = rng.integers(0, 256, (37, 8), dtype=np.uint8)
        Q = rng.integers(0, 256, (5, 8), dtype=np.uint8)
        ref = lambda k: oracle.knn_hamming(X, Q, k)  # noqa: E731
    else:
        X = rng.random((37, 6)).astype(np.float32)
        Q = rng.random((5, 6)).astype(np.float32)
        ref = lambda k: oracle.knn_float(metric, X, Q, k)  # noqa: E731
    h = N.ExactHandle(metric)
    h.set(X)
    for k in (36, 37, 38, 100):
        gi, gd = h.search(Q, k)
        oi, od = ref(k)
        assert gi.shape == (5, 37 if k >= 37 else k)
        np.testing.assert_array_equal(gd, od)
        np.testing.assert_array_equal(gi, oi)


def test_empty_and_mismatch_errors():
    h = N.ExactHandle(N.METRIC_L2)
    with pytest.raises(RuntimeError):
        h.search(np.zeros((1, 4), np.float32), 1)
    h.set(np.zeros((0, 4), np.float32))
    with pytest.raises(RuntimeError):
        h.search(np.zeros((1, 4), np.float32), 1)
    h.set(np.ones((10, 4), np.float32))
    with pytest.raises(ValueError):
        h.search(np.zeros((1, 5), np.float32), 1)


@pytest.mark.parametrize("use_null_stream", [False, True])
def test_device_entry_point_matches_host(use_null_stream):
    from pynear_amd import hipdev
    X, Q = _hamming_case(50_000, 64, 600, seed=3)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    hi, hd = h.search(Q, 10)
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((600, 10), np.int32)
    od = hipdev.DeviceArray((600, 10), np.int32)
    st = None if use_null_stream else hipdev.Stream()
    h.search_device(dq.ptr, 600, 10, oi.ptr, od.ptr, st.handle if st else None)
    if st:
        st.synchronize()
    hipdev.synchronize()
    np.testing.assert_array_equal(oi.numpy(), hi)
    np.testing.assert_array_equal(od.numpy(), hd)


@pytest.mark.parametrize("metric", [N.METRIC_L2, N.METRIC_L1])
def test_device_entry_point_pads_queries(metric):
    """Queries with a row length that is not a padded width (d=13) go through the device API."""
    from pynear_amd import hipdev
    rng = np.random.default_rng(21)
    X = rng.random((9000, 13)).astype(np.float32)
    Q = rng.random((77, 13)).astype(np.float32)
    h = N.ExactHandle(metric)
    h.set(X)
    dq = hipdev.DeviceArray.from_numpy(Q)
    oi = hipdev.DeviceArray((77, 5), np.int32)
    od = hipdev.DeviceArray((77, 5), np.float32)
    st = hipdev.Stream()
    h.search_device(dq.ptr, 77, 5, oi.ptr, od.ptr, st.handle)
    st.synchronize()
    ri, rd = oracle.knn_float(metric, X, Q, 5)
    np.testing.assert_array_equal(oi.numpy(), ri)
    np.testing.assert_array_equal(od.numpy(), rd)


@pytest.mark.parametrize("wave,lean", [(0, 0), (1, 0), (1, 1), (1, 2), (1, 3)])
@pytest.mark.parametrize("wpc", [1, 8])
def test_merge_and_slicing_variants(wave, lean, wpc):
    """Every merge (lane per query; wave per query with private lists; the lean wave merge and
    its rounds path) and different slice counts give identical results."""
    X, Q = _hamming_case(40_000, 32, 300, seed=wpc * 10 + wave)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    with _Tune(merge__wave_per_query=wave, merge__lean=lean, hamming__wg_per_cu=wpc):
        gi, gd = h.search(Q, 16)
    oi, od = oracle.knn_hamming(X, Q, 16)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("lean", [0, 1, 2, 3])
@pytest.mark.parametrize("case", ["hamming", "l2_mfma", "l2_valu"])
def test_all_rows_equal_ties_by_index(case, lean):
    """Every row equal: every distance ties, so the k results are rows 0 .. k - 1. With tens of
    thousands of tied keys the wave merge's sampled bound admits thousands of them: the merge with
    private lists takes its round-by-round path, the lean merge lowers its bound (or, forced, takes
    its rounds)."""
    rng = np.random.default_rng(17)
    k = 32 if case != "hamming" else 16
    if case == "hamming":
        X = np.tile(rng.integers(0, 256, (1, 32), dtype=np.uint8), (50_000, 1))
        Q = rng.integers(0, 256, (70, 32), dtype=np.uint8)
        h = N.ExactHandle(N.METRIC_HAMMING)
    else:
        X = np.tile(rng.standard_normal((1, 32)).astype(np.float32), (30_000, 1))
        Q = rng.standard_normal((70, 32)).astype(np.float32)
        h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma=0 if case == "l2_valu" else 1, merge__lean=lean):
        gi, gd = h.search(Q, k)
    assert (gi == np.arange(k)[None, :]).all()
    assert (gd == gd[:, :1]).all()
    if case == "hamming":
        oi, od = oracle.knn_hamming(X, Q, k)
    else:
        oi, od = oracle.knn_float(N.METRIC_L2, X, Q, k)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)


@pytest.mark.parametrize("fuse", [0, 1])
@pytest.mark.parametrize("lean", [0, 1, 2, 3])
@pytest.mark.parametrize("nq,k", [(1, 32), (3, 10), (200, 32), (200, 5)])
def test_merge_rows_sorted_by_cluster(fuse, lean, nq, k):
    """Rows stored cluster by cluster: a query's neighbours crowd a few consecutive slices (a few
    lanes of the merge hold most keys under the sampled bound), with one to a few hundred queries
    (up to eight slices per lane); merge and rerank in one launch or two."""
    rng = np.random.default_rng(nq * 100 + k)
    C = rng.standard_normal((40, 32)).astype(np.float32) * 4
    lab = np.sort(rng.integers(0, 40, 120_000))
    X = (C[lab] + 0.3 * rng.standard_normal((lab.size, 32))).astype(np.float32)
    Q = (C[rng.integers(0, 40, nq)] + 0.3 * rng.standard_normal((nq, 32))).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma=1, merge__lean=lean, l2__fuse_rerank=fuse):
        gi, gd = h.search(Q, k)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


def test_pairwise_distances_bit_exact():
    rng = np.random.default_rng(9)
    for d in (1, 3, 8, 13, 32, 77):
        a = rng.standard_normal((200, d)).astype(np.float32)
        b = rng.standard_normal((200, d)).astype(np.float32)
        for metric, m in ((N.METRIC_L2, oracle.L2), (N.METRIC_L1, oracle.L1), (N.METRIC_CHEBYSHEV, oracle.LINF)):
            g = N.pairwise_distance(metric, a, b)
            o = np.array([oracle.dist(m, a[i], b[i]) for i in range(200)], np.float32)
            np.testing.assert_array_equal(g, o)
    a = rng.integers(0, 256, (300, 23), dtype=np.uint8)
    b = rng.integers(0, 256, (300, 23), dtype=np.uint8)
    np.testing.assert_array_equal(N.pairwise_distance(N.METRIC_HAMMING, a, b), np.bitwise_count(a ^ b).sum(1))


class _Tune:
    def __init__(self, **kv):
        self.kv = {k.replace("__", "."): v for k, v in kv.items()}

    def __enter__(self):
        self.old = {k: N.tune_get(k) for k in self.kv}
        for k, v in self.kv.items():
            N.tune_set(k, v)

    def __exit__(self, *a):
        for k, v in self.old.items():
            N.tune_set(k, v)


@pytest.mark.parametrize("d", [1, 3, 16, 32, 48, 100, 128, 192, 256])
@pytest.mark.parametrize("k", [1, 5, 10, 32])
def test_l2_mfma_path_bit_exact(d, k):
    """The MFMA candidate scan + exact rerank returns the oracle's result bit for bit."""
    rng = np.random.default_rng(d * 1000 + k)
    X = rng.standard_normal((6000, d)).astype(np.float32)
    Q = rng.standard_normal((130, d)).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma=1):
        gi, gd = h.search(Q, k)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("d", [33, 39, 71, 135, 196, 263])
@pytest.mark.parametrize("k", [1, 10, 32])
@pytest.mark.parametrize("reg", [0, 1])
def test_l2_mfma_rerank_row_tails(d, k, reg):
    """Rows of 8 u + r floats (r = 1..7) through both rerank gathers (query chains in registers or
    re-read): the fused 4-wide and masked tails follow the reference's order bit for bit; clustered
    rows put many candidates near the k-th distance, so few are skipped."""
    rng = np.random.default_rng(d * 7 + k)
    C = rng.standard_normal((20, d)).astype(np.float32)
    X = (C[rng.integers(0, 20, 5000)] + 0.05 * rng.standard_normal((5000, d))).astype(np.float32)
    Q = (C[rng.integers(0, 20, 97)] + 0.05 * rng.standard_normal((97, d))).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma=1, l2__rerank_reg=reg):
        gi, gd = h.search(Q, k)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("d", [8, 32, 64, 128])
def test_l2_fallback_path_exact(d):
    """With single-product bf16 the error bound is wide, queries fall back to the exact scan,
    and the result is still the oracle's (and identical to the VALU scan's)."""
    rng = np.random.default_rng(d)
    X = rng.random((20000, d)).astype(np.float32)
    X[5000:5100] = X[:100]  # exact duplicates: ties broken by lower index
    Q = np.concatenate([X[:40] + 1e-4, rng.random((40, d)).astype(np.float32)])
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma_terms=1):
        gi, gd = h.search(Q, 10)
    with _Tune(l2__mfma=0):
        vi, vd = h.search(Q, 10)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, 10)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(vd, od)
    np.testing.assert_array_equal(vi, oi)


def test_l2_fallback_many_flagged_queries():
    """More flagged queries than fallback workgroups (two per CU): each workgroup takes several
    queries in turn; every result is the oracle's."""
    rng = np.random.default_rng(77)
    X = rng.random((6000, 32)).astype(np.float32)
    Q = rng.random((1500, 32)).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma_terms=1):
        gi, gd = h.search(Q, 10)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, 10)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("mfma", [0, 1])
def test_l2_sqrt_tie_orders_by_index(mfma):
    """Two rows whose squared distances differ (1 and 1 + 2^-23) but whose float square roots are
    equal (1.0) tie on the returned distance, so the lower index comes first, although its squared
    distance is the larger one; the other rows are far. Both selection paths (VALU scan with the
    sqrt-ordered merge, and the MFMA scan with its exact rerank) keep that order."""
    rng = np.random.default_rng(5)
    X = (rng.standard_normal((3000, 2)) * 0.1 + 10.0).astype(np.float32)
    X[2] = [1.0, 0.00028]  # squared distance to the origin 1.0000001, root 1.0
    X[5] = [1.0, 0.0]      # squared distance 1.0, root 1.0
    Q = np.zeros((3, 2), np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__mfma=mfma):
        gi, gd = h.search(Q, 2)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, 2)
    np.testing.assert_array_equal(gi, oi)
    np.testing.assert_array_equal(gd, od)
    assert gi[0].tolist() == [2, 5] and gd[0].tolist() == [1.0, 1.0]


@pytest.mark.parametrize("d,n", [(100, 3000), (256, 4096), (768, 1500), (80, 37)])
@pytest.mark.parametrize("k", [1, 10, 32])
def test_l2_small_database_of_wide_rows(d, n, k):
    """At most 4096 rows wider than 64 columns take the key-matrix path (a product per query chunk
    with a four-step load pipeline, then selection, rerank and fallback): bit-exact to the oracle."""
    rng = np.random.default_rng(d + n + k)
    X = rng.standard_normal((n, d)).astype(np.float32)
    Q = rng.standard_normal((150, d)).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    gi, gd = h.search(Q, k)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q, k)
    np.testing.assert_array_equal(gd, od)
    np.testing.assert_array_equal(gi, oi)


@pytest.mark.parametrize("d", [128, 192, 256])
def test_l2_wide_scan_equals_generic_scan(d):
    """Rows of 128, 192 and 256 columns: the scan with register-resident queries and the generic
    scan give the same result (several slices per query block, a ragged last slice, k = 7)."""
    rng = np.random.default_rng(d + 7)
    X = rng.standard_normal((9001, d)).astype(np.float32)
    Q = rng.standard_normal((260, d)).astype(np.float32)
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    with _Tune(l2__wide=1):
        wi, wd = h.search(Q, 7)
    with _Tune(l2__wide=0):
        gi, gd = h.search(Q, 7)
    np.testing.assert_array_equal(wi, gi)
    np.testing.assert_array_equal(wd, gd)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q[::13], 7)
    np.testing.assert_array_equal(wi[::13], oi)
    np.testing.assert_array_equal(wd[::13], od)


def test_l2_config1_shape_sample():
    """100k x 32 U[0,1), 1000 queries, k=5 (seed 0): every query vs the oracle."""
    import bench
    X, Q = bench.l2_workload()
    h = N.ExactHandle(N.METRIC_L2)
    h.set(X)
    gi, gd = h.search(Q, 5)
    sub = np.arange(0, 1000, 4)
    oi, od = oracle.knn_float(N.METRIC_L2, X, Q[sub], 5)
    np.testing.assert_array_equal(gd[sub], od)
    np.testing.assert_array_equal(gi[sub], oi)
    fi, fd = oracle.knn_fp64(N.METRIC_L2, X, Q[sub], 5)
    np.testing.assert_allclose(gd[sub], fd, rtol=1e-6)

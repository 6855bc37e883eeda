"""pn_exact_search_device_noalloc: the exact search without device allocation (fixed per-GPU
scratch) returns what pn_exact_search_device returns, call after call, and rejects the shapes its
scratch cannot serve."""

import numpy as np
import pytest

import oracle
from pynear_amd import _native as N
from pynear_amd import hipdev

pytestmark = pytest.mark.gpu


def _both(h, Q, k, dist_dtype):
    """(idx, dist) from search_device and from search_device_noalloc on the same device queries."""
    nq = len(Q)
    n, _ = h.info()
    kk = min(k, n)
    dq = hipdev.DeviceArray.from_numpy(Q)
    out = []
    for fn in (h.search_device, h.search_device_noalloc):
        oi = hipdev.DeviceArray((nq, kk), np.int32)
        od = hipdev.DeviceArray((nq, kk), dist_dtype)
        fn(dq.ptr, nq, k, oi.ptr, od.ptr)
        hipdev.synchronize()
        out.append((oi.numpy(), od.numpy()))
    return out


def _flips(rng, X, nq, p):
    src = rng.integers(0, len(X), nq)
    mask = (rng.random((nq, X.shape[1] * 8)) < p).astype(np.uint8)
    return X[src] ^ np.packbits(mask, axis=1, bitorder="little")


@pytest.mark.parametrize("nbytes", [64, 20])
def test_hamming_equal_to_search_device_two_calls(nbytes):
    rng = np.random.default_rng(11 + nbytes)
    X = rng.integers(0, 256, (60_000, nbytes), dtype=np.uint8)
    h = N.ExactHandle(N.METRIC_HAMMING)
    h.set(X)
    for nq, k in ((3_000, 10), (257, 7)):  # the second call smaller
        Q = _flips(rng, X, nq, 0.1)
        (ai, ad), (bi, bd) = _both(h, Q, k, np.int32)
        assert np.array_equal(ai, bi) and np.array_equal(ad, bd)
        oi, od = oracle.knn_hamming(X, Q[:64], k)
        assert np.array_equal(bi[:64], oi) and np.array_equal(bd[:64], od)


@pytest.mark.parametrize("metric,d", [(N.METRIC_L2, 32), (N.METRIC_L1, 17), (N.METRIC_CHEBYSHEV, 64)])
def test_float_metrics_equal_to_search_device_two_calls(metric, d):
    rng = np.random.default_rng(5 + d)
    X = rng.random((40_000, d), dtype=np.float32)
    h = N.ExactHandle(metric)
    h.set(X)
    for nq, k in ((1_000, 5), (33, 1)):
        Q = rng.random((nq, d), dtype=np.float32)
        (ai, ad), (bi, bd) = _both(h, Q, k, np.float32)
        assert np.array_equal(ai, bi) and np.array_equal(ad, bd)
        oi, od = oracle.knn_float(metric, X, Q[:50], k)
        assert np.array_equal(bi[:50], oi) and np.array_equal(bd[:50], od)


def test_rejects_what_the_scratch_cannot_serve():
    rng = np.random.default_rng(3)
    h = N.ExactHandle(N.METRIC_HAMMING)
    X = rng.integers(0, 256, (5_000, 8), dtype=np.uint8)
    h.set(X)
    nq = 200_000  # 32 keys x 200,192 padded queries x 8 B per slice > 32 MiB
    dq = hipdev.DeviceArray((nq, 8), np.uint8)
    oi = hipdev.DeviceArray((nq, 32), np.int32)
    od = hipdev.DeviceArray((nq, 32), np.int32)
    with pytest.raises(ValueError, match="scratch"):
        h.search_device_noalloc(dq.ptr, nq, 32, oi.ptr, od.ptr)
    with pytest.raises(ValueError):
        h.search_device_noalloc(dq.ptr, 10, 33, oi.ptr, od.ptr)  # k > 32
    wide = N.ExactHandle(N.METRIC_L2)
    wide.set(rng.random((1_000, 100), dtype=np.float32))  # more than 64 floats per row
    dq2 = hipdev.DeviceArray((4, 100), np.float32)
    with pytest.raises(ValueError):
        wide.search_device_noalloc(dq2.ptr, 4, 5, oi.ptr, od.ptr)

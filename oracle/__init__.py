"""CPU oracle for the pynear_amd search path -- TEST INFRASTRUCTURE ONLY.

Only ``tests/``, ``__graft_entry__.smoke()`` and the ``cpu_baseline`` leg of ``bench.py`` may import
this package, and only as the checker. The product package ``pynear_amd`` never imports it and
raises when its HIP library is missing instead of falling back to anything here.

Contents
--------
* ``liboracle.so`` (``oracle.cpp``): the reference's distance arithmetic in its own summation order,
  brute-force k-NN with (dist, idx) order, ``IVFFlatBinaryIndex`` build and search, ``kmeans_l2``.
* numpy restatements: ``ivf_l2_search`` (``pynear/forest.py:63-111``), FP64 brute force,
  ``hamming_range`` (BK-tree ``find_threshold``) and ``mih_search`` (``MIHBinaryIndex.searchKNN``).
* ``load_ref()``: the reference's own ``_pynear`` module compiled by ``oracle/Makefile`` into
  ``oracle/_ref`` (absent when it was not built).
"""

import ctypes
import glob
import heapq
import importlib.util
import os

import numpy as np

HERE = os.path.dirname(os.path.abspath(__file__))
_LIB = None

L2, L1, LINF = 0, 1, 2

_f32p = np.ctypeslib.ndpointer(np.float32, flags="C_CONTIGUOUS")
_u8p = np.ctypeslib.ndpointer(np.uint8, flags="C_CONTIGUOUS")
_i64p = np.ctypeslib.ndpointer(np.int64, flags="C_CONTIGUOUS")
_i32p = np.ctypeslib.ndpointer(np.int32, flags="C_CONTIGUOUS")
_i64 = ctypes.c_int64


def lib():
    """Load ``oracle/liboracle.so`` (build it with ``make -C oracle``)."""
    global _LIB
    if _LIB is not None:
        return _LIB
    path = os.path.join(HERE, "liboracle.so")
    if not os.path.exists(path):
        raise RuntimeError(f"oracle library missing: {path} (run `make -C oracle`)")
    L = ctypes.CDLL(path)
    L.or_dist_l2.restype = ctypes.c_float
    L.or_dist_l1.restype = ctypes.c_float
    L.or_dist_linf.restype = ctypes.c_float
    for f in (L.or_dist_l2, L.or_dist_l1, L.or_dist_linf):
        f.argtypes = [_f32p, _f32p, _i64]
    L.or_dist_hamming.restype = ctypes.c_int64
    L.or_dist_hamming.argtypes = [_u8p, _u8p, _i64]
    L.or_knn_float.argtypes = [ctypes.c_int, _f32p, _i64, _i64, _f32p, _i64, _i64, _i64p, _f32p]
    L.or_knn_hamming.argtypes = [_u8p, _i64, _i64, _u8p, _i64, _i64, _i64p, _i64p]
    L.or_ivfbin_build.restype = ctypes.c_int32
    L.or_ivfbin_build.argtypes = [_u8p, _i64, _i64, ctypes.c_int32, ctypes.c_int32, ctypes.c_uint32, _u8p, _i32p]
    L.or_ivfbin_search.argtypes = [_u8p, ctypes.c_int32, _i64p, _i32p, _u8p, _i64, _u8p, _i64,
                                   ctypes.c_int32, _i64, _i64p, _i64p, _i32p]
    L.or_kmeans_l2.argtypes = [_f32p, _i64, _i64, _i64, _i64, ctypes.c_uint32, _i32p, _f32p]
    _LIB = L
    return L


def _f32(a):
    return np.ascontiguousarray(a, dtype=np.float32)


def _u8(a):
    return np.ascontiguousarray(a, dtype=np.uint8)


# ---------------------------------------------------------------- distances
def dist(metric, a, b):
    a, b = _f32(a).ravel(), _f32(b).ravel()
    f = {L2: lib().or_dist_l2, L1: lib().or_dist_l1, LINF: lib().or_dist_linf}[metric]
    return float(f(a, b, a.size))


def dist_hamming(a, b):
    a, b = _u8(a).ravel(), _u8(b).ravel()
    return int(lib().or_dist_hamming(a, b, a.size))


# ---------------------------------------------------------------- exact k-NN
def knn_float(metric, X, Q, k):
    """Exact k-NN, ascending (dist, idx); distances in the reference's float arithmetic."""
    X, Q = _f32(X), _f32(Q)
    n, d = X.shape
    kk = min(k, n)
    oi = np.empty((len(Q), kk), np.int64)
    od = np.empty((len(Q), kk), np.float32)
    lib().or_knn_float(metric, X, n, d, Q, len(Q), k, oi, od)
    return oi, od


def knn_hamming(X, Q, k):
    X, Q = _u8(X), _u8(Q)
    n, nb = X.shape
    kk = min(k, n)
    oi = np.empty((len(Q), kk), np.int64)
    od = np.empty((len(Q), kk), np.int64)
    lib().or_knn_hamming(X, n, nb, Q, len(Q), k, oi, od)
    return oi, od


def knn_fp64(metric, X, Q, k):
    """FP64 brute force (the tolerance oracle): exact distances in double, (dist, idx) order."""
    X = np.asarray(X, np.float64)
    Q = np.asarray(Q, np.float64)
    kk = min(k, len(X))
    oi = np.empty((len(Q), kk), np.int64)
    od = np.empty((len(Q), kk), np.float64)
    for s in range(0, len(Q), 64):
        diff = Q[s:s + 64, None, :] - X[None, :, :]
        if metric == L2:
            D = np.sqrt((diff * diff).sum(-1))
        elif metric == L1:
            D = np.abs(diff).sum(-1)
        else:
            D = np.abs(diff).max(-1)
        order = np.lexsort((np.broadcast_to(np.arange(len(X)), D.shape), D), axis=-1)[:, :kk]
        oi[s:s + 64] = order
        od[s:s + 64] = np.take_along_axis(D, order, -1)
    return oi, od


# ---------------------------------------------------------------- IVF binary
def ivfbin_build(X, nlist, max_iter=20, seed=42):
    X = _u8(X)
    n, nb = X.shape
    kc = min(nlist, n)
    cent = np.empty((kc, nb), np.uint8)
    labels = np.empty(n, np.int32)
    lib().or_ivfbin_build(X, n, nb, nlist, max_iter, seed, cent, labels)
    return cent, labels


def lists_from_labels(labels, nc):
    """Inverted lists in CSR form, ids ascending within a list (BinaryIVF.hpp:223-226)."""
    labels = np.asarray(labels)
    order = np.argsort(labels, kind="stable").astype(np.int32)
    off = np.zeros(nc + 1, np.int64)
    np.cumsum(np.bincount(labels, minlength=nc), out=off[1:])
    return off, order


def ivfbin_search(cent, off, ids, X, Q, nprobe, k):
    cent, X, Q = _u8(cent), _u8(X), _u8(Q)
    nq, nb = Q.shape
    oi = np.empty((nq, k), np.int64)
    od = np.empty((nq, k), np.int64)
    cnt = np.empty(nq, np.int32)
    lib().or_ivfbin_search(cent, len(cent), np.ascontiguousarray(off, np.int64), np.ascontiguousarray(ids, np.int32),
                           X, nb, Q, nq, nprobe, k, oi, od, cnt)
    return [oi[q, :cnt[q]].tolist() for q in range(nq)], [od[q, :cnt[q]].tolist() for q in range(nq)]


# ---------------------------------------------------------------- k-means L2
def kmeans_l2(X, k, max_iter=100, seed=42):
    X = _f32(X)
    n, d = X.shape
    labels = np.empty(n, np.int32)
    cent = np.empty((k, d), np.float32)
    lib().or_kmeans_l2(X, n, d, k, max_iter, seed, labels, cent)
    return labels, cent


# ---------------------------------------------------------------- IVF L2 (numpy)
def ivf_l2_search(centroids, cluster_data, cluster_norms, orig_indices, queries, n_probe, k):
    """numpy restatement of ``_IVFFlatIndex.searchKNN`` (pynear/forest.py:63-111).

    Coarse GEMM distances (:76, :200-204), nprobe nearest centroids, per-list GEMV distances in the
    ``|q|^2 + |x|^2 - 2 q.x`` form clamped at 0 (:189-194), per-list local top-k and a global
    heap (:84-105); ascending squared L2. The reference selects probes with ``argpartition`` and
    breaks ties by heap order, both unspecified; here probes and ties follow (dist, id) order.
    """
    queries = np.asarray(queries, np.float32)
    if queries.ndim == 1:
        queries = queries[None]
    C = np.asarray(centroids, np.float32)
    cn = np.einsum("ij,ij->i", C, C)
    qn = np.einsum("ij,ij->i", queries, queries)
    cdist = np.maximum(0.0, qn[:, None] + cn[None, :] - 2.0 * (queries @ C.T))
    nc = len(C)
    n_probe = min(n_probe, nc)
    all_i, all_d = [], []
    for qi, q in enumerate(queries):
        order = np.lexsort((np.arange(nc), cdist[qi]))[:n_probe]
        heap = []
        q_norm = float(np.dot(q, q))
        for ci in order:
            data = cluster_data[ci]
            if len(data) == 0:
                continue
            dd = np.maximum(0.0, q_norm + cluster_norms[ci] - 2.0 * (data @ q))
            for li in range(len(dd)):
                key = (-float(dd[li]), -int(orig_indices[ci][li]))
                if len(heap) < k:
                    heapq.heappush(heap, key)
                elif key > heap[0]:
                    heapq.heapreplace(heap, key)
        res = sorted((-a, -b) for a, b in heap)
        all_i.append([b for _, b in res])
        all_d.append([a for a, _ in res])
    return all_i, all_d


# ---------------------------------------------------------------- Hamming range / MIH (numpy)
def _popcounts(X, Q):
    """(nq, n, nbytes) per-byte popcounts of q xor x."""
    return np.bitwise_count(np.asarray(Q, np.uint8)[:, None, :] ^ np.asarray(X, np.uint8)[None, :, :])


def hamming_range(X, Q, threshold, chunk=64):
    """Every row within ``threshold`` of each query: the set BKTree::find returns
    (pynear/include/BKTree.hpp:71-115; the BK-tree is complete, a row enters iff dist <= threshold),
    here in ascending row order. Returns (list of int64 row arrays, list of int64 distance arrays)."""
    out_i, out_d = [], []
    for s in range(0, len(Q), chunk):
        d = _popcounts(X, Q[s:s + chunk]).sum(-1, dtype=np.int64)
        for r in d:
            rows = np.flatnonzero(r <= threshold)
            out_i.append(rows.astype(np.int64))
            out_d.append(r[rows])
    return out_i, out_d


def mih_search(X, Q, m, radius, k, chunk=64):
    """MIHBinaryIndex::searchKNN (pynear/include/MIH.hpp:98-152) by brute force: a row is a candidate
    when one of its m sub-strings (nbytes / m bytes each, :165-169) is within radius / m bits of the
    query's (:102-122: the enumerated keys are exactly those within that distance); the k nearest
    candidates by full Hamming distance, ascending (distance, row). The reference keeps ties at the
    k-th distance in hash-set order; here the lower row wins."""
    X = np.asarray(X, np.uint8)
    nb = X.shape[1]
    sb = nb // m
    r_sub = radius // m
    out_i, out_d = [], []
    for s in range(0, len(Q), chunk):
        pc = _popcounts(X, Q[s:s + chunk])
        d = pc.sum(-1, dtype=np.int64)
        cand = (pc[:, :, :m * sb].reshape(pc.shape[0], pc.shape[1], m, sb).sum(-1) <= r_sub).any(-1)
        for dq, cq in zip(d, cand):
            rows = np.flatnonzero(cq)
            o = rows[np.lexsort((rows, dq[rows]))][:k]
            out_i.append(o.astype(np.int64))
            out_d.append(dq[o])
    return out_i, out_d


# ---------------------------------------------------------------- reference module
def load_ref():
    """Import the reference's own ``_pynear`` built into ``oracle/_ref`` (None when absent)."""
    cands = glob.glob(os.path.join(HERE, "_ref", "_pynear*.so"))
    if not cands:
        return None
    spec = importlib.util.spec_from_file_location("_pynear", cands[0])
    mod = importlib.util.module_from_spec(spec)
    spec.loader.exec_module(mod)
    return mod

"""Index classes and free functions with the interface of the reference's ``_pynear`` module.

Each class mirrors one pybind11 class of pynear/src/PythonBindings.cpp:331-481 (same names,
methods, argument meaning, result layout and exception types) and runs on the GPU through the
C-ABI in ``include/pynear_amd.h``:

* ``VPTreeL2Index`` / ``VPTreeL1Index`` / ``VPTreeChebyshevIndex`` (:344-366): float32 rows,
  ``searchKNN`` -> (list[list[int]], list[list[float]]) farthest-first (VPTree.hpp:621-631),
  L2 distances are Euclidean (sqrt), ``search1NN`` -> (list[int], list[float]).
* ``VPTreeBinaryIndex{512,256,128,64}`` and ``VPTreeBinaryIndex`` (:368-406): uint8 rows,
  integer Hamming distances.

The VP-tree walk is replaced by an exact brute-force scan, so the tree-shaped parts of the
interface change meaning: ``to_string`` describes the scan index and pickling stores the rows
(the reference pickles tree state, PythonBindings.cpp:97-142, :185-197). Neighbours come back in
a total order: ascending (distance, index), reversed to farthest-first as in the reference.
"""

import numpy as np

from . import _native as N


def _as_2d_float(arr, what):
    a = np.ascontiguousarray(np.asarray(arr), dtype=np.float32)  # py::array::forcecast
    if a.ndim != 2:
        raise RuntimeError(f"{what}() expects a 2D float32 array of shape (n, d)")
    return a


def _as_2d_u8(arr):
    a = np.asarray(arr)
    if a.ndim != 2:
        raise RuntimeError("binary index expects a 2D array of uint8 rows")
    if a.size and (a.dtype.kind not in "iub" or a.min() < 0 or a.max() > 255):
        raise TypeError("binary index expects uint8 values")
    return np.ascontiguousarray(a, dtype=np.uint8)


class _ExactIndexBase:
    _metric = None

    def __init__(self):
        self._h = N.ExactHandle(self._metric)
        self._rows = None

    # -- state
    def _prepare(self, arr, what):
        raise NotImplementedError

    def set(self, vectors):
        rows = self._prepare(vectors, "set")
        self._h.set(rows)
        self._rows = rows if rows.shape[0] else None

    def _search(self, vectors, k, what):
        q = self._prepare(vectors, what)
        n, _ = self._h.info()
        if n == 0:
            raise RuntimeError("index must be first initialized with .set() function and non empty dataset")
        return self._h.search(q, k)

    def searchKNN(self, vectors, k):
        """Batch find top-k vectors in index and return indices and distances (farthest-first)."""
        idx, dist = self._search(vectors, k, "searchKNN")
        return idx[:, ::-1].tolist(), dist[:, ::-1].tolist()

    def search1NN(self, vectors):
        """Batch find closest vectors in index and return indices and distances."""
        idx, dist = self._search(vectors, 1, "search1NN")
        return idx[:, 0].tolist(), dist[:, 0].tolist()

    def to_string(self):
        n, d = self._h.info()
        return (f"####################\n# [{type(self).__name__} state]\nNum Data Points: {n}\n"
                f"Dimension: {d}\nSearch: exact GPU scan\n####################\n")

    # -- pickle: rows are stored and re-uploaded to the device on load
    def __getstate__(self):
        return {"rows": self._rows}

    def __setstate__(self, state):
        self.__init__()
        if state.get("rows") is not None:
            self.set(state["rows"])


class _FloatIndex(_ExactIndexBase):
    def _prepare(self, arr, what):
        return _as_2d_float(arr, what)


class _BinaryIndex(_ExactIndexBase):
    _metric = N.METRIC_HAMMING
    _width = None  # bytes per row for the fixed-width classes

    def _prepare(self, arr, what):
        a = _as_2d_u8(arr)
        if self._width is not None and a.shape[0] and a.shape[1] != self._width:
            raise ValueError(f"{type(self).__name__} expects rows of {self._width} bytes, got {a.shape[1]}")
        return a


class VPTreeL2Index(_FloatIndex):
    _metric = N.METRIC_L2


class VPTreeL1Index(_FloatIndex):
    _metric = N.METRIC_L1


class VPTreeChebyshevIndex(_FloatIndex):
    _metric = N.METRIC_CHEBYSHEV


class VPTreeBinaryIndex512(_BinaryIndex):
    _width = 64


class VPTreeBinaryIndex256(_BinaryIndex):
    _width = 32


class VPTreeBinaryIndex128(_BinaryIndex):
    _width = 16


class VPTreeBinaryIndex64(_BinaryIndex):
    _width = 8


class VPTreeBinaryIndex(_BinaryIndex):
    _width = None


class IVFFlatBinaryIndex:
    """Inverted-file index for binary descriptors (approximate Hamming k-NN) on the GPU.

    Mirrors ``_pynear.IVFFlatBinaryIndex`` (PythonBindings.cpp:454-466, BinaryIVF.hpp:34-227):
    ``IVFFlatBinaryIndex(nlist=256, nprobe=8, max_iter=20, seed=42)``, ``set``, ``searchKNN`` ->
    (list[list[int]], list[list[int]]) ascending by distance, ``nlist()``, ``nprobe()``,
    ``set_nprobe()``. The build reproduces the reference's clustering draw for draw, and the
    search returns what the reference's heap returns, ties included. Unlike the reference, the
    index can be pickled (the rows and parameters are stored and the index is rebuilt on load).
    """

    def __init__(self, nlist: int = 256, nprobe: int = 8, max_iter: int = 20, seed: int = 42):
        self._args = (int(nlist), int(nprobe), int(max_iter), int(seed))
        self._h = N.IvfBinHandle(*self._args)
        self._rows = None

    def set(self, vectors):
        rows = _as_2d_u8(vectors) if np.asarray(vectors).size else np.zeros((0, 0), np.uint8)
        self._h.set(rows)
        self._rows = rows if rows.shape[0] else None

    def searchKNN(self, vectors, k):
        """Batch find top-k vectors in index and return indices and distances."""
        q = _as_2d_u8(vectors)
        idx, dist, cnt = self._h.search(q, k)
        return ([idx[i, :c].tolist() for i, c in enumerate(cnt)], [dist[i, :c].tolist() for i, c in enumerate(cnt)])

    def nlist(self):
        return self._h.params()[0]

    def nprobe(self):
        return self._h.params()[1]

    def set_nprobe(self, nprobe):
        self._h.set_nprobe(nprobe)

    def __getstate__(self):
        nlist, nprobe = self._h.params()
        return {"args": (nlist, nprobe) + self._args[2:], "rows": self._rows}

    def __setstate__(self, state):
        self.__init__(*state["args"])
        if state.get("rows") is not None:
            self.set(state["rows"])


class _BKTreeIndex:
    """Hamming range search over uint8 rows with the interface of ``BKTreeBinaryIndex*``
    (PythonBindings.cpp:209-230, :408-451; BKTree.hpp:37-166).

    ``set`` adds rows (BKTree::update: the index keeps earlier rows and numbers new ones after
    them); ``find_threshold(vectors, threshold)`` returns, per query, every stored row within
    ``threshold`` bits as (indices, distances, keys) lists. The rows are scored on the GPU by a
    full scan instead of a tree walk, which finds the same set; matches come in ascending row
    order (the reference's order is its tree's breadth-first order). ``values`` returns the
    stored rows in insertion order (the reference: tree order; both documented as arbitrary).
    """

    _width = None

    def __init__(self):
        self._h = N.ExactHandle(N.METRIC_HAMMING)
        self._rows = None

    def _check_width(self, a):
        want = self._width if self._width is not None else (None if self._rows is None else self._rows.shape[1])
        if want is not None and a.shape[1] != want:
            raise ValueError(f"{type(self).__name__} holds rows of {want} bytes, got {a.shape[1]}")

    def set(self, vectors):
        a = np.asarray(vectors)
        if a.size == 0:
            return
        a = _as_2d_u8(a)
        self._check_width(a)
        self._rows = a.copy() if self._rows is None else np.concatenate([self._rows, a])
        self._h.set(self._rows)

    def find_threshold(self, vectors, threshold):
        """Batch find all vectors below the distance threshold (inclusive)."""
        a = np.asarray(vectors)
        nq = a.shape[0] if a.ndim >= 1 else 0
        if self._rows is None or nq == 0:
            return [[] for _ in range(nq)], [[] for _ in range(nq)], [[] for _ in range(nq)]
        q = _as_2d_u8(a)
        self._check_width(q)
        off, idx, dist = self._h.range(q, int(threshold))
        rows = self._rows
        ind = [idx[off[i]:off[i + 1]].tolist() for i in range(nq)]
        return (ind, [dist[off[i]:off[i + 1]].tolist() for i in range(nq)], [rows[r].tolist() for r in ind])

    def empty(self):
        return self._rows is None

    def size(self):
        return 0 if self._rows is None else int(self._rows.shape[0])

    def values(self):
        """Return all stored vectors in arbitrary order (here: insertion order)."""
        return [] if self._rows is None else self._rows.tolist()


class BKTreeBinaryIndex512(_BKTreeIndex):
    _width = 64


class BKTreeBinaryIndex256(_BKTreeIndex):
    _width = 32


class BKTreeBinaryIndex128(_BKTreeIndex):
    _width = 16


class BKTreeBinaryIndex64(_BKTreeIndex):
    _width = 8


class BKTreeBinaryIndex(_BKTreeIndex):
    _width = None


class MIHBinaryIndex:
    """Multi-Index Hashing for binary descriptors (approximate Hamming k-NN) on the GPU, with the
    interface of ``_pynear.MIHBinaryIndex`` (PythonBindings.cpp:253-272, :468-481; MIH.hpp:50-195):
    ``MIHBinaryIndex(m=8)``, ``set``, ``searchKNN(vectors, k, radius=8)`` -> (list[list[int]],
    list[list[int]]) ascending by distance, possibly fewer than k per query, ``m()``, ``n()``,
    ``nbytes()``. The candidates are the reference's (rows with a sub-string within radius // m
    bits of the query's); every row is tested on the GPU instead of looking keys up in tables.
    Ties at the k-th distance keep the lower row (the reference: hash-set order).
    """

    def __init__(self, m: int = 8):
        self._h = N.MihHandle(int(m))

    def set(self, vectors):
        a = np.asarray(vectors)
        if a.size == 0:
            self._h.set(np.zeros((0, 0), np.uint8))
            return
        self._h.set(_as_2d_u8(a))

    def searchKNN(self, vectors, k, radius=8):
        """Batch find top-k vectors in index and return indices and distances."""
        q = _as_2d_u8(vectors)
        idx, dist, cnt = self._h.search(q, int(k), int(radius))
        return ([idx[i, :c].tolist() for i, c in enumerate(cnt)], [dist[i, :c].tolist() for i, c in enumerate(cnt)])

    def m(self):
        return self._h.info()[0]

    def n(self):
        return self._h.info()[1]

    def nbytes(self):
        return self._h.info()[2]


# ---------------------------------------------------------------- free functions
def _pair_float(metric, a, b):
    a = np.ascontiguousarray(np.asarray(a), dtype=np.float32).reshape(1, -1)
    b = np.ascontiguousarray(np.asarray(b), dtype=np.float32).reshape(1, -1)
    if a.shape != b.shape:
        raise ValueError("a and b must have the same number of elements")
    return float(N.pairwise_distance(metric, a, b)[0])


def kmeans_l2(data, k, max_iter=100, seed=42):
    """Lloyd k-means with k-means++ initialisation on the GPU (PythonBindings.cpp:305-329).

    Returns (labels int32[n], centroids float32[k, d]), identical to the reference's kmeans_l2.
    """
    a = np.ascontiguousarray(np.asarray(data), dtype=np.float32)
    if a.ndim != 2:
        raise RuntimeError("kmeans_l2: data must be a 2D float32 array (N, D)")
    if k == 0 or k > a.shape[0]:
        raise RuntimeError("kmeans_l2: k must be between 1 and N")
    return N.kmeans_l2(a, k, max_iter, seed)


def dist_l2(a, b):
    """Euclidean distance, reference arithmetic (DistanceFunctions.hpp:304-345)."""
    return _pair_float(N.METRIC_L2, a, b)


def dist_l1(a, b):
    """Manhattan distance (DistanceFunctions.hpp:347-382)."""
    return _pair_float(N.METRIC_L1, a, b)


def dist_chebyshev(a, b):
    """Chebyshev distance (DistanceFunctions.hpp:384-420)."""
    return _pair_float(N.METRIC_CHEBYSHEV, a, b)


def _hamming_fixed(a, b, nbytes):
    a = np.asarray(a, dtype=np.uint8).reshape(1, -1)
    b = np.asarray(b, dtype=np.uint8).reshape(1, -1)
    if a.shape[1] < nbytes or b.shape[1] < nbytes:
        raise ValueError(f"hamming distance on {nbytes * 8} bits needs {nbytes} bytes per operand")
    a = np.ascontiguousarray(a[:, :nbytes])
    b = np.ascontiguousarray(b[:, :nbytes])
    return int(N.pairwise_distance(N.METRIC_HAMMING, a, b)[0])


def dist_hamming_512(a, b):
    return _hamming_fixed(a, b, 64)


def dist_hamming_256(a, b):
    return _hamming_fixed(a, b, 32)


def dist_hamming_128(a, b):
    return _hamming_fixed(a, b, 16)


def dist_hamming_64(a, b):
    return _hamming_fixed(a, b, 8)
